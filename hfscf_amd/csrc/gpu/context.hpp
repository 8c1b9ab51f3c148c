// Context of the hfgpu library (host state and device buffers), shared by the
// J/K driver (context.hip) and the gradient driver (grad.hip).
#pragma once
#include <algorithm>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../../include/hfgpu.h"
#include "coop_args.hpp"
#include "gather.hpp"
#include "gc.hpp"
#include "kernels.hpp"
#include "launch_bounds.hpp"
#include "oneint.hpp"
#include "private_segments.hpp"

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

constexpr int kSideStreams = 8;  // most side streams a context may use (hfg_options.side_streams)
// Default side streams: HIP gives a process GPU_MAX_HW_QUEUES (4) hardware queues and the
// streams share them; measured FP64 / mixed builds with 1, 2, 3, 4, 6 side streams: (H2O)16
// 43.8/25.8, 34.3/18.0, 33.6/18.9, 34.6/18.9, 33.4/17.6 ms; (H2O)8 15.8/11.0, 10.1/6.3, 9.2/5.6,
// 10.8/6.4, 8.9/5.5 ms.
constexpr int kDefaultSideStreams = 6;
// Classes from L = la+lb+lc+ld = kCoopRequiredL have only the cooperative kernels; from
// kCoopTuneMinL up both forms exist and hfg_options.coop_min_l / hfg_autotune choose.
constexpr int kCoopRequiredL = 7;
constexpr int kCoopTuneMinL = 4;
// Autotuning tries the cooperative form from this L only: below it the per-lane register kernels
// won every measurement on (H2O)8 and (H2O)16 / cc-pVDZ (by up to 5x at L = 4; the (H2O)16
// gradient's L = 5 classes too), and the trials of large classes cost seconds in a batch.
constexpr int kCoopTrialMinL = 6;

inline int nfun(int L, bool pure) { return pure ? 2 * L + 1 : (L + 1) * (L + 2) / 2; }


struct ClassWork {
    int la, lb, lc, ld;
    std::vector<WorkItem> items;
    size_t dev_off = 0;
    long long quartets = 0;
    double flops = 0;
    double grad_flops = 0;
    LaunchShape jk_shape{0, 0};    // persistent workgroups / block of the J/K launch (0 = automatic)
    LaunchShape grad_shape{0, 0};  // the same for the gradient launch
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // brackets of the last direct launch
    // J/K on the wavefront-cooperative kernel: its program (device pointers into the context's
    // d_coop buffer) and the LDS bytes of its FP64 and mixed launches
    bool coop = false;
    bool coop_ok = false;  // a program exists (the class may move between the two kernel forms)
    CoopArgs coop_args{};
    size_t coop_lds[2] = {0, 0};
    // the gradient on the cooperative kernel (program in d_gcoop)
    bool gcoop = false;
    bool gcoop_ok = false;
    CoopGradArgs gcoop_args{};
    size_t gcoop_lds[2] = {0, 0};
    // J/K of (ss|ss) on the shell-group kernel: its own work items over the group lists (the
    // per-shell items above then serve the gradient only) and the model FLOPs of that kernel
    bool gc = false;
    std::vector<WorkItem> gc_items;
    size_t gc_dev_off = 0;
    double gc_flops = 0;
    // the work items, their offset in d_items and the model FLOPs of the J/K launch
    const std::vector<WorkItem>& jk_items() const { return gc ? gc_items : items; }
    size_t jk_dev_off() const { return gc ? gc_dev_off : dev_off; }
    double jk_flops() const { return gc ? gc_flops : flops; }
};

// One molecule of a context (a context holds one, or many in batching mode).
struct MolHost {
    int shell0 = 0, nshell = 0;  // global shell range
    int atom0 = 0, natom = 0;    // global atom range (geometry)
    int nbf = 0, ncart = 0;
    int blk0 = 0, nblk = 0;      // global block range
    int bp0 = 0, nbp = 0;        // global block-pair range
    long long moff = 0;          // offset of its n x n matrices in concatenated arrays
    int dm_off = 0;              // offset of its nshell x nshell density-maximum table
    int gout0 = 0;               // offset of its natom x 3 gradient in concatenated output
};

// Default split of a class launch into work items (hfg_options.work_split replaces it): four times
// the items of one per 1/1536 of the build's model cost, which keeps several wavefronts per SIMD
// busy in the wide s and p classes (measured on (H2O)16 / cc-pVDZ: 48 ms per FP64 build at 1,
// 37 at 2, 36 at 4, 38 at 8).
constexpr double kDefaultWorkSplit = 4.0;

// Cached graph of one J/K build configuration.
struct JkGraph {
    bool valid = false;
    int mixed = 0;
    bool dscreen = false;
    double dtol = 0.0;  // the density threshold baked into the captured launch arguments
    hipGraphExec_t exec = nullptr;
};

}  // namespace hfg

struct hfg_ctx {
    using ClassWork = hfg::ClassWork;
    template <typename T>
    using DevBuf = hfg::DevBuf<T>;
    using PairDev = hfg::PairDev;
    using WorkItem = hfg::WorkItem;
    using GatherEntry = hfg::GatherEntry;
    using GatherChunk = hfg::GatherChunk;
    int device = 0;
    int max_lds = 65536;  // LDS bytes one workgroup may use on the device
    double* scratch = nullptr;  // the device's fixed scratch of hfg_build_jk_noalloc
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nside = hfg::kDefaultSideStreams;  // side streams in use, class launches run concurrently on these
    hipStream_t side[hfg::kSideStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[hfg::kSideStreams] = {};
    // the device's shared stream for kernels that may spill to scratch (see shared_scratch_stream)
    hipStream_t heavy = nullptr;
    hipEvent_t ev_heavy_join = nullptr;
    std::vector<int> class_stream;  // side stream of each class's J/K launch (-1: the shared stream)
    std::vector<int> grad_stream;   // the same for the gradient launches
    std::vector<int> class_order, grad_order;  // launch order of the classes (J/K, gradient)
    // hardware queue of each side stream and of the shared stream (equal values: one queue, whose
    // kernels run one after another); the stream balance spreads the classes over the queues
    int side_queue[hfg::kSideStreams] = {};
    int heavy_queue = -1;
    hfg_options opt{};
    std::string err;
    bool have_basis = false, have_geom = false;
    bool batch = false;  // molecules added with hfg_batch_add_molecule
    // molecules
    std::vector<hfg::MolHost> mols;
    long long nn_total = 0;  // sum of n^2 over molecules
    int nout_total = 0;      // sum of 3 natom over molecules
    // basis (concatenated over molecules; offsets within each molecule)
    int nshell = 0, nbf = 0, ncart = 0;  // nbf / ncart: of molecule 0 (single-molecule calls)
    bool pure = false;
    std::vector<int> L, atom, atom_local, nprim, prim_off, cart_off, pure_off, shell_mol;
    std::vector<double> alpha, coef;
    // geometry (concatenated)
    int natom = 0;
    std::vector<double> xyz;
    std::vector<double> Zn;  // nuclear charges of a single molecule (hfg_set_geometry), for hfg_oneint
    // blocks
    int nblk = 0;
    std::vector<int> blk_L, blk_nf, blk_func_off, blk_func_ao, blk_mol, shell_blk, shell_slot;
    std::vector<std::vector<int>> blk_shells;
    // block pairs (X >= Y, same molecule)
    std::vector<int> bp_X, bp_Y;
    // pairs
    std::vector<PairDev> pairs;
    std::vector<double> psmean;  // sort key of each pair (mean Schwarz value, or the primitive bound)
    std::vector<double> psq;  // sqrt|(pq|pq)| per pair component
    std::vector<int> bp_off, bp_cnt;
    std::vector<double> screen;  // Cartesian (pq|pq) of molecule 0 (single-molecule contexts)
    // shell groups of the s blocks (hfg_options.group_shells): per s-s block pair the sorted list
    // of group pairs, their keys and primitive data
    std::vector<hfg::GcPairDev> gpairs;
    std::vector<double> gkey;
    std::vector<int> goff, gcnt;
    std::vector<int> gc_nfull;  // primitive pairs stored per group pair of each list
    double tol = 0.0;       // quartets with psmean_ab psmean_cd below this are skipped
    double elem_tol = 0.0;  // hfscf's element screen and Cartesian diagonal (0 without screening)
    // work
    std::vector<ClassWork> classes;
    std::vector<int4> tiles;
    std::vector<int> gJ_off, gK_off;
    std::vector<GatherEntry> gJ, gK;
    std::vector<GatherChunk> gchunks;  // first-pass chunks: every tile's J chunks, then its K chunks
    std::vector<int> cJ_off, cK_off;    // per tile, its J / K chunks in gchunks
    long long gpart_doubles = 0;        // chunk sums of one build
    long long slab_doubles = 0, grad_slab_doubles = 0, total_quartets = 0, total_items = 0;
    double total_flops = 0, total_grad_flops = 0;
    // device
    DevBuf<double> d_psq, d_boys, d_D, d_J, d_K, d_F, d_H, d_G, d_slab, d_psmean, d_grad_slab, d_grad;
    DevBuf<int4> d_coop;   // programs of the classes on the cooperative J/K kernel
    DevBuf<int4> d_gcoop;  // programs of the classes on the cooperative gradient kernel
    DevBuf<float> d_dmax;  // largest |density| per shell block (density-weighted screening)
    std::vector<float> dmax;
    DevBuf<double4> d_pp0, d_pp1, d_pp2;
    DevBuf<PairDev> d_pairs;
    DevBuf<hfg::MolDev> d_mols;
    DevBuf<int> d_bp_off, d_bp_cnt, d_blk_nf, d_blk_func_off, d_blk_func_ao, d_gJ_off, d_gK_off, d_shell_atom;
    DevBuf<int> d_queue;  // one work-item counter per class launch (J/K, then gradient)
    DevBuf<int> d_mol_items, d_mol_items_off, d_out_mol, d_out_k;
    DevBuf<WorkItem> d_items;
    DevBuf<int4> d_tiles;
    DevBuf<GatherEntry> d_gJ, d_gK;
    DevBuf<GatherChunk> d_gchunks;
    DevBuf<int> d_cJ_off, d_cK_off;
    DevBuf<double> d_gpart;
    DevBuf<hfg::GcPairDev> d_gpairs;
    DevBuf<double> d_gkey;
    DevBuf<int> d_goff, d_gcnt;
    DevBuf<double4> d_gc_p0, d_gc_p1, d_gc_K;
    hfg::GcArgs gc_args{};  // device pointers of the group tables
    DevBuf<double> d_eri_out;  // the packed vector of hfg_eri_vector (freed after each call)
    // one-electron integrals (hfg_oneint, hfg_oneint_gradient)
    DevBuf<hfg::OneShell> d_one_sh;
    DevBuf<int2> d_one_pairs;
    DevBuf<double> d_one_alpha, d_one_coef, d_one_xyz, d_one_Z, d_one_buf;
    DevBuf<int> d_hess_cart;  // first Cartesian function per shell (hfg_hessian_2e)
    DevBuf<int4> d_hess_items;  // its work items (hfg::HItem, four ints)
    // launch graphs of the J/K build (keyed by precision and density screening)
    std::vector<hfg::JkGraph> graphs;
    bool class_timed = false;  // the last J/K build bracketed each class launch with events
    // incremental Fock state: the previous density, on the host (hfg_fock) or on the device
    // (hfg_fock_device, prev_on_device)
    bool have_prev = false, prev_on_device = false;
    std::vector<double> Dprev;
    DevBuf<double> d_Dprev;
    // per shell (global index): output-basis offset within its molecule, function count, molecule
    DevBuf<int> d_shell_ao, d_shell_nf, d_shell_mol;
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
// Two-electron gradients (concatenated natom x 3 per molecule) through the adjoint
// derivative kernels, for the concatenated densities D.
int run_gradient(hfg_ctx* c, const double* D, double* grad);
}  // namespace hfg

namespace hfg {
// Launch fn(class_work, stream) for every class on the side streams (fork from and join back
// to `main`): classes are independent until the gather. Under stream capture the fork/join
// events make the side-stream launches part of the captured graph.
// Which class launches a call enqueues: every class, only those on the side streams (the part of
// a J/K build that is captured in a graph), or only those on the shared stream.
enum class ClassSet { All, Side, Shared };

// Number of classes that run on the shared stream (J/K, or the gradient when grad). Their
// device-queue counters come first ([0, n)), the side-stream classes' after them.
inline int shared_class_count(const hfg_ctx* c, bool grad) {
    const std::vector<int>& streams = grad ? c->grad_stream : c->class_stream;
    return (int)std::count(streams.begin(), streams.end(), -1);
}

// Forks `main` to the streams of the selected classes, calls fn(class, stream, counter index)
// for each in the launch order, and joins them back into `main` (the shared stream only with
// join_shared). Counter indices: shared-stream classes 0.., then the side-stream classes, in
// launch order.
template <typename F>
inline hipError_t for_each_class_concurrent(hfg_ctx* c, hipStream_t main, F&& fn, bool grad = false,
                                            ClassSet set = ClassSet::All, bool join_shared = true) {
    const std::vector<int>& streams = grad ? c->grad_stream : c->class_stream;
    const bool side = set != ClassSet::Shared, shared = set != ClassSet::Side;
    hipError_t e = hipEventRecord(c->ev_fork, main);
    if (e != hipSuccess) return e;
    if (side)
        for (int k = 0; k < c->nside; ++k)
            if ((e = hipStreamWaitEvent(c->side[k], c->ev_fork, 0)) != hipSuccess) return e;
    if (shared && (e = hipStreamWaitEvent(c->heavy, c->ev_fork, 0)) != hipSuccess) return e;
    // the shared-stream classes first: they are few workgroups of long latency, and launched
    // after the side streams' persistent grids they would wait for compute units to drain
    const std::vector<int>& order = grad ? c->grad_order : c->class_order;
    int q = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (size_t n = 0; n < c->classes.size(); ++n) {
            const size_t i = order.size() == c->classes.size() ? (size_t)order[n] : n;
            if ((streams[i] < 0) == (pass == 0)) {
                const int qi = q++;
                if ((pass == 0 && !shared) || (pass == 1 && !side)) continue;
                if ((e = fn(c->classes[i], streams[i] < 0 ? c->heavy : c->side[streams[i]], qi)) != hipSuccess)
                    return e;
            }
        }
    if (side)
        for (int k = 0; k < c->nside; ++k) {
            if ((e = hipEventRecord(c->ev_join[k], c->side[k])) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(main, c->ev_join[k], 0)) != hipSuccess) return e;
        }
    if (!shared) return hipSuccess;
    if ((e = hipEventRecord(c->ev_heavy_join, c->heavy)) != hipSuccess) return e;
    // without join_shared the caller makes main wait for ev_heavy_join itself, later
    return join_shared ? hipStreamWaitEvent(main, c->ev_heavy_join, 0) : hipSuccess;
}

// Largest scratch per lane of a gradient register kernel that runs concurrently with others on
// the side streams: the J/K register kernels of (dp|dp) and (dd|pp), 4,088 bytes per lane, run
// there in every build.
constexpr int kGradConcurrentBytes = 4096;

// Whether a class launch runs on the device's shared stream (one kernel at a time): J/K register
// kernels whose scratch bound is above kJkSharedBytes; gradient register kernels whose private
// segment in this build (the larger of the FP64 and mixed kernels) is above kGradConcurrentBytes
// (the scratch bound decides where the table has no entry).
inline bool on_shared_stream(const hfg_ctx* c, const ClassWork& cw, bool grad) {
    if (grad ? cw.gcoop : cw.coop) return false;
    if (!grad) return class_spills(cw.la, cw.lb, cw.lc, cw.ld, false);
    const int b64 = private_segment_bytes(true, false, cw.la, cw.lb, cw.lc, cw.ld, c->pure);
    const int b32 = private_segment_bytes(true, true, cw.la, cw.lb, cw.lc, cw.ld, c->pure);
    if (b64 < 0 || b32 < 0) return class_spills(cw.la, cw.lb, cw.lc, cw.ld, true);
    return std::max(b64, b32) > kGradConcurrentBytes;
}

// Stream of a class launch (J/K or gradient): a kernel of the shared stream always runs there,
// also while autotuning tries it for a class that is on a side stream.
inline hipStream_t class_launch_stream(hfg_ctx* c, size_t k, bool grad) {
    const ClassWork& cw = c->classes[k];
    const int s = (grad ? c->grad_stream : c->class_stream)[k];
    if (s < 0 || on_shared_stream(c, cw, grad)) return c->heavy;
    return c->side[s];
}

// The J/K launch of one class (the cooperative kernel or the class kernel) on stream st.
hipError_t launch_jk_class(hfg_ctx* c, const ClassWork& cw, const JkArgs& a, LaunchShape sh, hipStream_t st);

// HFG_OK when every class kernel that a J/K build (or a gradient when grad) dispatches has a
// private segment within kMaxPrivateBytes; HFG_ERR_UNSUPPORTED with the class named otherwise.
int check_private_segments(hfg_ctx* c, bool grad);

// Run class k alone (J/K, or the gradient when grad) with launch shape sh on the main stream,
// bracketed by ev0 / ev1, and wait for it (the autotuner's measurement).
int run_one_class(hfg_ctx* c, size_t k, bool grad, LaunchShape sh);
}  // namespace hfg
