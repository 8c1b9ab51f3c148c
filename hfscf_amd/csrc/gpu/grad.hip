// Two-electron RHF gradient driver (adjoint derivative-ERI kernels).
#include "context.hpp"

namespace hfg {

int run_gradient(hfg_ctx* c, const double* D, double* grad) {
    (void)D;
    (void)grad;
    c->err = "hfg_energy_backward: gradient kernels not built";
    return HFG_ERR_UNSUPPORTED;
}

}  // namespace hfg
