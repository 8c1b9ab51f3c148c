// One class of the J/K kernel on the shell-group lists; built once per class, basis type and
// precision with -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. -DHFG_PURE=0|1 [-DHFG_MIXED=1].
#include "jk_gc_impl.hpp"

#ifndef HFG_MIXED
#define HFG_MIXED 0
#endif

namespace hfg {

#if HFG_MIXED
using Real = float;
#define HFG_LAUNCH launch_jk_gc_class32
#else
using Real = double;
#define HFG_LAUNCH launch_jk_gc_class64
#endif

template <>
hipError_t HFG_LAUNCH<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0)>(const JkArgs& a, const GcArgs& g, int nitems,
                                                                     LaunchShape sh, hipStream_t s) {
    JkArgs b = a;
    b.nitems = nitems;
    const int block = sh.block > 0 ? sh.block : 256;
    const int grid = sh.grid > 0 ? std::min(sh.grid, nitems) : nitems;
    hipLaunchKernelGGL((jk_gc_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0), Real>), dim3(grid), dim3(block), 0,
                       s, b, g);
    return hipGetLastError();
}

}  // namespace hfg
