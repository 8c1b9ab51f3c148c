// One angular-momentum class of the adjoint gradient kernel; built once per class with
// -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. so the classes compile in parallel.
#include "grad_impl.hpp"

namespace hfg {

template <>
hipError_t launch_grad_class<HFG_LA, HFG_LB, HFG_LC, HFG_LD>(bool pure, const GradArgs& a, int nitems, hipStream_t s) {
    const dim3 grid(nitems), block(256);
    if (pure)
        hipLaunchKernelGGL((grad_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, true>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((grad_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace hfg
