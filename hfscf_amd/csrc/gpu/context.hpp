// Context of the hfgpu library (host state and device buffers), shared by the
// J/K driver (context.hip) and the gradient driver (grad.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../../include/hfgpu.h"
#include "kernels.hpp"

namespace hfg {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<size_t>(count, 1);
        return e;
    }
    hipError_t upload(const std::vector<T>& v, hipStream_t s) {
        hipError_t e = ensure(v.size());
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

constexpr int kSideStreams = 4;  // class launches run concurrently on these

inline int nfun(int L, bool pure) { return pure ? 2 * L + 1 : (L + 1) * (L + 2) / 2; }


struct ClassWork {
    int la, lb, lc, ld;
    std::vector<WorkItem> items;
    size_t dev_off = 0;
    long long quartets = 0;
    double flops = 0;
    double grad_flops = 0;
};

}  // namespace hfg

struct hfg_ctx {
    using ClassWork = hfg::ClassWork;
    template <typename T>
    using DevBuf = hfg::DevBuf<T>;
    using PairDev = hfg::PairDev;
    using WorkItem = hfg::WorkItem;
    using GatherEntry = hfg::GatherEntry;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t side[hfg::kSideStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[hfg::kSideStreams] = {};
    std::vector<int> class_stream;  // side stream of each class launch
    hfg_options opt{};
    std::string err;
    bool have_basis = false, have_geom = false;
    // basis
    int nshell = 0, nbf = 0, ncart = 0;
    bool pure = false;
    std::vector<int> L, atom, nprim, prim_off, cart_off, pure_off;
    std::vector<double> alpha, coef;
    // geometry
    int natom = 0;
    std::vector<double> xyz;
    // blocks
    int nblk = 0;
    std::vector<int> blk_L, blk_nf, blk_func_off, blk_func_ao, shell_blk, shell_slot;
    std::vector<std::vector<int>> blk_shells;
    // pairs
    std::vector<PairDev> pairs;
    std::vector<double> psmean;
    std::vector<double> psq;  // sqrt|(pq|pq)| per pair component
    std::vector<int> bp_off, bp_cnt;
    std::vector<double4> pp0, pp1, pp2;
    std::vector<double> screen;
    double tol = 0.0;
    // work
    std::vector<ClassWork> classes;
    std::vector<int2> tiles;
    std::vector<int> gJ_off, gK_off;
    std::vector<GatherEntry> gJ, gK;
    long long slab_doubles = 0, total_quartets = 0, total_items = 0;
    double total_flops = 0, total_grad_flops = 0;
    // device
    DevBuf<double> d_psq, d_boys, d_D, d_J, d_K, d_F, d_H, d_G, d_slab, d_psmean, d_grad_slab, d_grad;
    DevBuf<float> d_dmax;  // largest |density| per shell block (density-weighted screening)
    std::vector<float> dmax;
    DevBuf<double4> d_pp0, d_pp1, d_pp2;
    DevBuf<PairDev> d_pairs;
    DevBuf<int> d_bp_off, d_bp_cnt, d_blk_nf, d_blk_func_off, d_blk_func_ao, d_gJ_off, d_gK_off, d_shell_atom;
    DevBuf<WorkItem> d_items;
    DevBuf<int2> d_tiles;
    DevBuf<GatherEntry> d_gJ, d_gK;
    // incremental Fock state
    bool have_prev = false;
    std::vector<double> Dprev;
    hfg_stats stats{};
};

#define HFG_CHECK(ctx, call)                                                              \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return e_ == hipErrorOutOfMemory ? HFG_ERR_NOMEM : HFG_ERR_HIP;               \
        }                                                                                 \
    } while (0)


namespace hfg {
// Two-electron gradient (natom x 3) through the adjoint derivative kernels.
int run_gradient(hfg_ctx* c, const double* D, double* grad);
}  // namespace hfg

namespace hfg {
// Launch fn(class_work, stream) for every class on the side streams (fork/join with the main
// stream): classes are independent until the gather.
template <typename F>
inline hipError_t for_each_class_concurrent(hfg_ctx* c, F&& fn) {
    hipError_t e = hipEventRecord(c->ev_fork, c->stream);
    if (e != hipSuccess) return e;
    for (int k = 0; k < kSideStreams; ++k)
        if ((e = hipStreamWaitEvent(c->side[k], c->ev_fork, 0)) != hipSuccess) return e;
    for (size_t i = 0; i < c->classes.size(); ++i)
        if ((e = fn(c->classes[i], c->side[c->class_stream[i]])) != hipSuccess) return e;
    for (int k = 0; k < kSideStreams; ++k) {
        if ((e = hipEventRecord(c->ev_join[k], c->side[k])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(c->stream, c->ev_join[k], 0)) != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace hfg
