// One angular-momentum class of the J/K kernel; built once per class with
// -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. so the classes compile in parallel.
#include "jk_impl.hpp"

namespace hfg {

template <>
hipError_t launch_jk_class<HFG_LA, HFG_LB, HFG_LC, HFG_LD>(bool pure, int mixed, const JkArgs& a, int nitems,
                                                         hipStream_t s) {
    (void)mixed;
    const dim3 grid(nitems), block(256);
    if (pure)
        hipLaunchKernelGGL((jk_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, true, double, double>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((jk_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, false, double, double>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace hfg
