// One angular-momentum class of the adjoint gradient kernel; built once per class, basis type
// and precision with -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. -DHFG_PURE=0|1
// [-DHFG_MIXED=1] so the variants compile in parallel.
#include "grad_impl.hpp"

#ifndef HFG_MIXED
#define HFG_MIXED 0
#endif

namespace hfg {

#if HFG_MIXED
using Real = float;
#define HFG_LAUNCH launch_grad_class32
#else
using Real = double;
#define HFG_LAUNCH launch_grad_class64
#endif

template <>
hipError_t HFG_LAUNCH<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0)>(const GradArgs& a, int nitems,
                                                                     LaunchShape sh, hipStream_t s) {
    GradArgs b = a;
    b.nitems = nitems;
    constexpr int LT = HFG_LA + HFG_LB + HFG_LC + HFG_LD;
    constexpr size_t est = sizeof(Real) * vrr_entries(HFG_LA + HFG_LB + 1, HFG_LC + HFG_LD + 1, LT + 1);
    auto k = grad_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0), Real>;
    const LaunchShape ls = bounded_shape(k, nitems, sh, est);
    hipLaunchKernelGGL(k, dim3(ls.grid), dim3(ls.block), 0, s, b);
    return hipGetLastError();
}

}  // namespace hfg
