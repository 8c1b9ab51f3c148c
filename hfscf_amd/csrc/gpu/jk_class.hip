// One angular-momentum class of the J/K kernel; built once per class and precision with
// -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. [-DHFG_MIXED=1] so the classes compile in
// parallel. The mixed build evaluates and digests the quartets in FP32 (Boys values, pair
// data and screening in FP64) and accumulates the LDS tiles, the slabs and the gather in FP64.
#include "jk_impl.hpp"

#ifndef HFG_MIXED
#define HFG_MIXED 0
#endif

namespace hfg {

#if HFG_MIXED
using Real = float;
#define HFG_LAUNCH launch_jk_class32
#else
using Real = double;
#define HFG_LAUNCH launch_jk_class64
#endif

template <>
hipError_t HFG_LAUNCH<HFG_LA, HFG_LB, HFG_LC, HFG_LD>(bool pure, const JkArgs& a, int nitems, hipStream_t s) {
    JkArgs b = a;
    b.nitems = nitems;
    const int block = 256;
    if (pure) {
        auto k = jk_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, true, Real, double>;
        hipLaunchKernelGGL(k, dim3(grid_for(k, nitems, block)), dim3(block), 0, s, b);
    } else {
        auto k = jk_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, false, Real, double>;
        hipLaunchKernelGGL(k, dim3(grid_for(k, nitems, block)), dim3(block), 0, s, b);
    }
    return hipGetLastError();
}

}  // namespace hfg
