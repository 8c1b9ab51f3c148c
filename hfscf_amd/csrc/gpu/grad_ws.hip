// One class of the workspace gradient kernel ((dd|dd), (dd|dp), (dp|dd), FP64); built once per
// class and basis type with -DHFG_LA=.. -DHFG_LB=.. -DHFG_LC=.. -DHFG_LD=.. -DHFG_PURE=0|1.
#include "grad_ws.hpp"

namespace hfg {

template <>
hipError_t launch_grad_ws<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0)>(const GradArgs& a, double* ws, int nitems,
                                                                         hipStream_t s) {
    GradArgs b = a;
    b.nitems = nitems;
    auto k = grad_ws_kernel<HFG_LA, HFG_LB, HFG_LC, HFG_LD, (HFG_PURE != 0)>;
    hipLaunchKernelGGL(k, dim3(kWsBlocks), dim3(kWsBlock), 0, s, b, ws);
    return hipGetLastError();
}

}  // namespace hfg
