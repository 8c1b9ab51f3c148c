// The mixed-precision gradient launchers of the (dd|dd) class run the FP64 kernel: its FP32
// instantiation did not finish compiling within an hour (hipcc for gfx950), so this class keeps
// FP64 recursions in the mixed mode. Every other class has its FP32 kernel.
#include "grad_impl.hpp"

namespace hfg {

template <>
hipError_t launch_grad_class64<2, 2, 2, 2, (HFG_PURE != 0)>(const GradArgs&, int, LaunchShape, hipStream_t);

template <>
hipError_t launch_grad_class32<2, 2, 2, 2, (HFG_PURE != 0)>(const GradArgs& a, int nitems, LaunchShape sh,
                                                           hipStream_t s) {
    return launch_grad_class64<2, 2, 2, 2, (HFG_PURE != 0)>(a, nitems, sh, s);
}

}  // namespace hfg
