// Two-electron RHF gradient driver (adjoint derivative-ERI kernels).
#include <vector>

#include "context.hpp"
#include "grad_impl.hpp"

namespace hfg {

int run_gradient(hfg_ctx* c, const double* D, double* grad) {
    if (c->natom > kMaxAtoms) {
        c->err = "hfg_energy_backward: more atoms than the gradient kernel's LDS table holds";
        return HFG_ERR_UNSUPPORTED;
    }
    const size_t nn = (size_t)c->nbf * c->nbf;
    const int nout = c->natom * 3;
    HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, c->d_grad_slab.ensure((size_t)std::max<long long>(c->total_items, 1) * nout));
    HFG_CHECK(c, c->d_grad.ensure((size_t)nout));
    GradArgs a{};
    a.pairs = c->d_pairs.p;
    a.psmean = c->d_psmean.p;
    a.bp_off = c->d_bp_off.p;
    a.bp_cnt = c->d_bp_cnt.p;
    a.pp0 = c->d_pp0.p;
    a.pp1 = c->d_pp1.p;
    a.pp2 = c->d_pp2.p;
    a.boys = c->d_boys.p;
    a.D = c->d_D.p;
    a.shell_atom = c->d_shell_atom.p;
    a.n = c->nbf;
    a.natom = c->natom;
    a.tol = c->tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = c->d_grad_slab.p;
    HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
    HFG_CHECK(c, for_each_class_concurrent(c, [&](const ClassWork& cw, hipStream_t st) {
        GradArgs b = a;
        b.items = c->d_items.p + cw.dev_off;
        b.item_base = (int)cw.dev_off;
        return launch_grad(cw.la, cw.lb, cw.lc, cw.ld, c->pure, c->opt.precision == HFG_MIXED ? 1 : 0, b,
                           (int)cw.items.size(), st);
    }));
    HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
    HFG_CHECK(c, launch_grad_reduce(c->d_grad_slab.p, (int)c->total_items, nout, c->d_grad.p, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(grad, c->d_grad.p, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HFG_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_seconds = ms * 1e-3;
    c->stats.model_flops = c->total_grad_flops;
    c->stats.quartets = c->total_quartets;
    return HFG_OK;
}

}  // namespace hfg
