// Two-electron RHF gradient driver (adjoint derivative-ERI kernels) and the single-class runs
// of the launch-shape autotuner.
#include <vector>

#include "context.hpp"
#include "coop_grad.hpp"

namespace hfg {

namespace {

GradArgs grad_args(hfg_ctx* c) {
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
    a.mols = c->d_mols.p;
    a.shell_atom = c->d_shell_atom.p;
    a.tol = c->tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = c->d_grad_slab.p;
    return a;
}

JkArgs jk_args(hfg_ctx* c) {
    JkArgs a{};
    a.pairs = c->d_pairs.p;
    a.psmean = c->d_psmean.p;
    a.bp_off = c->d_bp_off.p;
    a.bp_cnt = c->d_bp_cnt.p;
    a.blk_nf = c->d_blk_nf.p;
    a.pp0 = c->d_pp0.p;
    a.pp1 = c->d_pp1.p;
    a.pp2 = c->d_pp2.p;
    a.boys = c->d_boys.p;
    a.D = c->d_D.p;
    a.mols = c->d_mols.p;
    a.tol = c->tol;
    a.psq = c->d_psq.p;
    a.elem_tol = c->elem_tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = c->d_slab.p;
    a.dmax = c->d_dmax.p;
    a.dtol = 0.0;
    return a;
}

// The gradient launch of one class: the cooperative kernel (L >= 7) or the class kernel.
hipError_t launch_grad_any(hfg_ctx* c, const ClassWork& cw, const GradArgs& b, LaunchShape sh, int mixed,
                           hipStream_t st) {
    if (cw.gcoop) return launch_grad_coop(b, cw.gcoop_args, cw.gcoop_lds[mixed], mixed, (int)cw.items.size(), sh, st);
    return launch_grad(cw.la, cw.lb, cw.lc, cw.ld, c->pure, mixed, b, (int)cw.items.size(), sh, st);
}

}  // namespace

int run_gradient(hfg_ctx* c, const double* D, double* grad) {
    if (int r = check_private_segments(c, true)) return r;
    for (const auto& M : c->mols)
        if (M.natom > kMaxAtoms) {
            c->err = "hfg_energy_backward: more atoms in a molecule than the gradient kernel's LDS table holds";
            return HFG_ERR_UNSUPPORTED;
        }
    const size_t nn = (size_t)c->nn_total;
    const int nout = c->nout_total;
    const int ncls = (int)c->classes.size();
    HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, c->d_grad_slab.ensure((size_t)std::max<long long>(c->grad_slab_doubles, 1)));
    HFG_CHECK(c, c->d_grad.ensure((size_t)std::max(nout, 1)));
    GradArgs a = grad_args(c);
    const int mixed = c->opt.precision == HFG_MIXED ? 1 : 0;
    HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
    HFG_CHECK(c, launch_zero_ints(c->d_queue.p + ncls, ncls, c->stream));
    HFG_CHECK(c, for_each_class_concurrent(c, c->stream, [&](const ClassWork& cw, hipStream_t st, int qi) {
        GradArgs b = a;
        b.items = c->d_items.p + cw.dev_off;
        b.queue = c->d_queue.p + ncls + qi;
        return launch_grad_any(c, cw, b, cw.grad_shape, mixed, st);
    }, true));
    HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
    GradReduceArgs r{};
    r.slab = c->d_grad_slab.p;
    r.items = c->d_items.p;
    r.mol_items = c->d_mol_items.p;
    r.mol_items_off = c->d_mol_items_off.p;
    r.out_mol = c->d_out_mol.p;
    r.out_k = c->d_out_k.p;
    r.out = c->d_grad.p;
    HFG_CHECK(c, launch_grad_reduce(r, nout, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(grad, c->d_grad.p, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HFG_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_seconds = ms * 1e-3;
    c->stats.model_flops = c->total_grad_flops;
    c->stats.quartets = c->total_quartets;
    return HFG_OK;
}

int run_one_class(hfg_ctx* c, size_t k, bool grad, LaunchShape sh) {
    if (int r = check_private_segments(c, grad)) return r;
    const ClassWork& cw = c->classes[k];
    const int mixed = c->opt.precision == HFG_MIXED ? 1 : 0;
    int* q = c->d_queue.p + (grad ? c->classes.size() : 0) + k;
    hipStream_t st = class_launch_stream(c, k, grad);
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    HFG_CHECK(c, launch_zero_ints(q, 1, st));
    HFG_CHECK(c, hipEventRecord(c->ev0, st));
    if (grad) {
        HFG_CHECK(c, c->d_grad_slab.ensure((size_t)std::max<long long>(c->grad_slab_doubles, 1)));
        GradArgs b = grad_args(c);
        b.slab = c->d_grad_slab.p;
        b.items = c->d_items.p + cw.dev_off;
        b.queue = q;
        HFG_CHECK(c, launch_grad_any(c, cw, b, sh, mixed, st));
    } else {
        JkArgs b = jk_args(c);
        b.items = c->d_items.p + cw.jk_dev_off();
        b.queue = q;
        HFG_CHECK(c, launch_jk_class(c, cw, b, sh, st));
    }
    HFG_CHECK(c, hipEventRecord(c->ev1, st));
    HFG_CHECK(c, hipStreamSynchronize(st));
    return HFG_OK;
}

}  // namespace hfg
