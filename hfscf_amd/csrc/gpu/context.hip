// Host side of the hfgpu C ABI: set-up of blocks, primitive pairs, Schwarz screening, work
// decomposition and gather lists for one molecule or a batch of molecules; the J/K, Fock,
// energy and gradient drivers; HIP-graph replay of the J/K build and the launch-shape autotuner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "boys.hpp"
#include "context.hpp"
#include "coop.hpp"
#include "coop_grad.hpp"
#include "flopmodel.hpp"
#include "launch_bounds.hpp"
#include "misc.hpp"
#include "noalloc.hpp"
#include "private_segments.hpp"

using namespace hfg;

namespace {

int fail(hfg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// One stream per device, shared by every context of the process, for the kernels that may
// spill to scratch (the d-shell class kernels and the Schwarz diagonals). The runtime gives each
// hardware queue its own scratch from a limited per-device pool and keeps it; spilling kernels
// spread over several queues (several streams, several contexts) exhausted that pool
// (HSA_STATUS_ERROR_OUT_OF_RESOURCES). On one stream only one queue holds scratch. It lives
// until the process ends.
hipStream_t shared_scratch_stream(int device) {
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)streams.size() <= device) streams.resize(device + 1, nullptr);
    if (!streams[device] && hipStreamCreate(&streams[device]) != hipSuccess) streams[device] = nullptr;
    return streams[device];
}

// Streams and launch order of the class launches (J/K, or the gradient when grad): the classes
// of the shared stream (on_shared_stream) run there (-1); the others are balanced over the side
// streams, onto the least loaded one, by the weights w (the FLOP model at set-up, the measured
// kernel times after autotuning). The launch order is the cooperative-kernel classes first (one
// quartet per workgroup at a time: few workgroups on long serial chains, whose latency then
// hides under the wide launches instead of trailing the build), then heaviest first.
void assign_streams(hfg_ctx* c, bool grad, const std::vector<double>& w) {
    std::vector<int>& map = grad ? c->grad_stream : c->class_stream;
    map.assign(c->classes.size(), 0);
    std::vector<size_t> ord(c->classes.size());
    std::iota(ord.begin(), ord.end(), 0);
    auto coop = [&](size_t k) { return grad ? c->classes[k].gcoop : c->classes[k].coop; };
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
        if (coop(x) != coop(y)) return coop(x);
        return w[x] > w[y];
    });
    std::vector<int>& order = grad ? c->grad_order : c->class_order;
    order.assign(ord.begin(), ord.end());
    // the balance is over hardware queues: side streams on one queue count as one, the first of
    // them takes the queue's classes, and a queue shared with the shared stream starts with the
    // classes that run there
    const int nside = std::max(1, std::min(c->nside, kSideStreams));
    double load[kSideStreams] = {0};
    bool rep[kSideStreams] = {};
    for (int s = 0; s < nside; ++s) rep[s] = c->side_queue[s] == s;
    for (size_t k = 0; k < c->classes.size(); ++k) {
        const ClassWork& cw = c->classes[k];
        if (on_shared_stream(c, cw, grad) && c->heavy_queue >= 0) load[c->heavy_queue] += w[k];
    }
    for (size_t k : ord) {
        const ClassWork& cw = c->classes[k];
        if (on_shared_stream(c, cw, grad)) {
            map[k] = -1;
            continue;
        }
        int best = -1;
        for (int s = 0; s < nside; ++s)
            if (rep[s] && (best < 0 || load[s] < load[best])) best = s;
        map[k] = best;
        load[best] += w[k];
    }
}

int ket_count_host(const double* s, int n, double si, double tol) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid] * si >= tol)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Blocks: shells of one molecule, one angular momentum and one primitive count, at most
// kBlockFuncs output functions each. Every shell pair of a block pair then has the same number
// of primitive pairs, so the lanes of a wavefront run primitive loops of equal length. Blocks
// of a molecule are contiguous.
void build_blocks(hfg_ctx* c) {
    std::vector<int> order(c->nshell);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (c->shell_mol[x] != c->shell_mol[y]) return c->shell_mol[x] < c->shell_mol[y];
        if (c->L[x] != c->L[y]) return c->L[x] < c->L[y];
        return c->nprim[x] > c->nprim[y];
    });
    c->blk_L.clear();
    c->blk_nf.clear();
    c->blk_mol.clear();
    c->blk_shells.clear();
    c->shell_blk.assign(c->nshell, -1);
    c->shell_slot.assign(c->nshell, -1);
    size_t k = 0;
    while (k < order.size()) {
        const int Lb = c->L[order[k]], m = c->shell_mol[order[k]], np = c->nprim[order[k]];
        const int nf = nfun(Lb, c->pure);
        const int per = std::max(1, kBlockFuncs / nf);
        std::vector<int> sh;
        while (k < order.size() && c->L[order[k]] == Lb && c->shell_mol[order[k]] == m && c->nprim[order[k]] == np &&
               (int)sh.size() < per)
            sh.push_back(order[k++]);
        const int b = (int)c->blk_L.size();
        for (size_t s = 0; s < sh.size(); ++s) {
            c->shell_blk[sh[s]] = b;
            c->shell_slot[sh[s]] = (int)s;
        }
        c->blk_L.push_back(Lb);
        c->blk_nf.push_back((int)sh.size() * nf);
        c->blk_mol.push_back(m);
        c->blk_shells.push_back(sh);
    }
    c->nblk = (int)c->blk_L.size();
    for (auto& M : c->mols) M.nblk = 0;
    for (int b = c->nblk - 1; b >= 0; --b) {
        c->mols[c->blk_mol[b]].blk0 = b;
        c->mols[c->blk_mol[b]].nblk++;
    }
    c->blk_func_off.assign(c->nblk + 1, 0);
    c->blk_func_ao.clear();
    for (int b = 0; b < c->nblk; ++b) {
        const int nf = nfun(c->blk_L[b], c->pure);
        for (int s : c->blk_shells[b]) {
            const int off = c->pure ? c->pure_off[s] : c->cart_off[s];
            for (int p = 0; p < nf; ++p) c->blk_func_ao.push_back(off + p);
        }
        c->blk_func_off[b + 1] = (int)c->blk_func_ao.size();
    }
}

// Block pair index of blocks R >= C of one molecule.
inline int bp_index(const hfg_ctx* c, int R, int C) {
    const MolHost& M = c->mols[c->blk_mol[R]];
    const int r = R - M.blk0, q = C - M.blk0;
    return M.bp0 + r * (r + 1) / 2 + q;
}

// A candidate shell pair: device record plus its global shell ids and the offset of its
// Cartesian Schwarz diagonal.
struct CandPair {
    PairDev p;
    int ga, gb;
    long long diag_off;
};

// Primitive-pair data of global shells (a, b) appended to pp0/pp1/pp2.
PairDev make_pair(hfg_ctx* c, int a, int b, std::vector<double4>& pp0, std::vector<double4>& pp1,
                  std::vector<double4>& pp2) {
    static const double kSqrt2Pi25 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    const int s0 = c->mols[c->shell_mol[a]].shell0;
    PairDev p{};
    p.sa = a - s0;
    p.sb = b - s0;
    p.slot_a = c->shell_slot[a];
    p.slot_b = c->shell_slot[b];
    p.ao_a = c->pure ? c->pure_off[a] : c->cart_off[a];
    p.ao_b = c->pure ? c->pure_off[b] : c->cart_off[b];
    const double* A = &c->xyz[3 * c->atom[a]];
    const double* B = &c->xyz[3 * c->atom[b]];
    double r2 = 0;
    for (int d = 0; d < 3; ++d) {
        p.AB[d] = A[d] - B[d];
        r2 += p.AB[d] * p.AB[d];
    }
    p.off = (int)pp0.size();
    p.pstride = 1;
    p.np = c->nprim[a] * c->nprim[b];
    for (int i = 0; i < c->nprim[a]; ++i)
        for (int j = 0; j < c->nprim[b]; ++j) {
            const double ai = c->alpha[c->prim_off[a] + i], bj = c->alpha[c->prim_off[b] + j];
            const double g = ai + bj;
            double P[3];
            for (int d = 0; d < 3; ++d) P[d] = (ai * A[d] + bj * B[d]) / g;
            const double K = std::exp(-ai * bj / g * r2) * c->coef[c->prim_off[a] + i] * c->coef[c->prim_off[b] + j] *
                             kSqrt2Pi25 / g;
            pp0.push_back(make_double4(g, P[0], P[1], P[2]));
            pp1.push_back(make_double4(P[0] - A[0], P[1] - A[1], P[2] - A[2], K));
            pp2.push_back(make_double4(0.5 / g, ai, bj, 0.0));
        }
    return p;
}

// Shells a and b on one atom with the same angular momentum and identical exponents.
bool same_exponents(const hfg_ctx* c, int a, int b) {
    if (c->atom[a] != c->atom[b] || c->L[a] != c->L[b] || c->nprim[a] != c->nprim[b]) return false;
    for (int k = 0; k < c->nprim[a]; ++k)
        if (c->alpha[c->prim_off[a] + k] != c->alpha[c->prim_off[b] + k]) return false;
    return true;
}

// Shell groups of the s blocks: consecutive shells of a block on one atom with identical
// exponents, at most kGcMaxShells per group (a block's shells keep the basis order, so the
// contracted s shells of one atom are neighbours). For every s-s block pair, the list of group
// pairs (GA, GB) (GB <= GA when both blocks are one) sorted by descending key: members are the
// shell pairs (a in GA, b in GB) in canonical order (b <= a within one group), each the index
// of its surviving per-shell pair (-1 when screened out), and the key is the largest member key,
// so the group lists are screened like the per-shell lists and each member quartet again on its
// own keys. The primitive pairs of a group pair are those of its exponents with one K per
// member; the ones where some member's K passes the significance test of the per-shell primitive
// pairs (|K| mq >= sig_cut) come first, in basis order, and np counts them. Returns false (no
// group lists) when no group has two shells or group_shells is off.
bool build_groups(hfg_ctx* c, double mq, double sig_cut, std::vector<double4>& gp0, std::vector<double4>& gp1,
                  std::vector<double4>& gK) {
    static const double kSqrt2Pi25 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    const int nbp = (int)c->bp_X.size();
    c->gpairs.clear();
    c->gkey.clear();
    c->goff.assign(nbp, 0);
    c->gcnt.assign(nbp, 0);
    c->gc_nfull.assign(nbp, 0);
    gp0.clear();
    gp1.clear();
    gK.clear();
    if (c->opt.group_shells < 0) return false;
    std::vector<std::vector<std::vector<int>>> groups(c->nblk);  // slots of each group of each block
    bool twins = false;
    for (int X = 0; X < c->nblk; ++X) {
        const std::vector<int>& sh = c->blk_shells[X];
        for (size_t s = 0; s < sh.size();) {
            std::vector<int> g{(int)s};
            while (c->blk_L[X] == 0 && (int)g.size() < kGcMaxShells && s + g.size() < sh.size() &&
                   same_exponents(c, sh[s], sh[s + g.size()]))
                g.push_back((int)(s + g.size()));
            twins = twins || g.size() > 1;
            s += g.size();
            groups[X].push_back(std::move(g));
        }
    }
    if (!twins) return false;
    struct Entry {
        double key = 0.0;
        GcPairDev gp{};
        std::vector<double4> p0, p1, K;
    };
    std::vector<int> pidx(kBlockFuncs * kBlockFuncs);
    std::vector<double4> rest0, rest1, restK;  // a group pair's insignificant primitive pairs
    for (int g = 0; g < nbp; ++g) {
        c->goff[g] = (int)c->gpairs.size();
        const int X = c->bp_X[g], Y = c->bp_Y[g];
        std::fill(pidx.begin(), pidx.end(), -1);
        for (int k = c->bp_off[g]; k < c->bp_off[g] + c->bp_cnt[g]; ++k)
            pidx[c->pairs[k].slot_a * kBlockFuncs + c->pairs[k].slot_b] = k;
        std::vector<Entry> ents;
        for (size_t ia = 0; ia < groups[X].size(); ++ia)
            for (size_t ib = 0; ib < groups[Y].size(); ++ib) {
                if (X == Y && ib > ia) continue;
                Entry e;
                int sha[kGcMaxMembers], shb[kGcMaxMembers];
                bool live = false;
                for (int sa : groups[X][ia])
                    for (int sb : groups[Y][ib]) {
                        if (X == Y && ia == ib && sb > sa) continue;
                        const int m = e.gp.nmem++;
                        e.gp.mem[m] = pidx[sa * kBlockFuncs + sb];
                        sha[m] = c->blk_shells[X][sa];
                        shb[m] = c->blk_shells[Y][sb];
                        if (e.gp.mem[m] >= 0) {
                            live = true;
                            e.key = std::max(e.key, c->psmean[e.gp.mem[m]]);
                        }
                    }
                if (!live) continue;
                for (int m = e.gp.nmem; m < kGcMaxMembers; ++m) e.gp.mem[m] = -1;
                const int a0 = sha[0], b0 = shb[0];
                const double* A = &c->xyz[3 * c->atom[a0]];
                const double* B = &c->xyz[3 * c->atom[b0]];
                double r2 = 0;
                for (int d = 0; d < 3; ++d) {
                    e.gp.AB[d] = A[d] - B[d];
                    r2 += e.gp.AB[d] * e.gp.AB[d];
                }
                for (int i = 0; i < c->nprim[a0]; ++i)
                    for (int j = 0; j < c->nprim[b0]; ++j) {
                        const double ai = c->alpha[c->prim_off[a0] + i], bj = c->alpha[c->prim_off[b0] + j];
                        const double p = ai + bj;
                        double K[kGcMaxMembers] = {0.0, 0.0, 0.0, 0.0};
                        bool sig = false;
                        for (int m = 0; m < e.gp.nmem; ++m) {
                            if (e.gp.mem[m] < 0) continue;
                            // the per-shell primitive pair's K, evaluated in the same order
                            K[m] = std::exp(-ai * bj / p * r2) * c->coef[c->prim_off[sha[m]] + i] *
                                   c->coef[c->prim_off[shb[m]] + j] * kSqrt2Pi25 / p;
                            sig = sig || std::fabs(K[m]) * mq >= sig_cut;
                        }
                        double P[3];
                        for (int d = 0; d < 3; ++d) P[d] = (ai * A[d] + bj * B[d]) / p;
                        (sig ? e.p0 : rest0).push_back(make_double4(p, P[0], P[1], P[2]));
                        (sig ? e.p1 : rest1).push_back(make_double4(P[0] - A[0], P[1] - A[1], P[2] - A[2], 0.5 / p));
                        (sig ? e.K : restK).push_back(make_double4(K[0], K[1], K[2], K[3]));
                    }
                e.gp.np = (int)e.p0.size();
                e.p0.insert(e.p0.end(), rest0.begin(), rest0.end());
                e.p1.insert(e.p1.end(), rest1.begin(), rest1.end());
                e.K.insert(e.K.end(), restK.begin(), restK.end());
                rest0.clear();
                rest1.clear();
                restK.clear();
                ents.push_back(std::move(e));
            }
        std::stable_sort(ents.begin(), ents.end(), [](const Entry& x, const Entry& y) { return x.key > y.key; });
        // primitive-major: every group pair of the list has nprim_X x nprim_Y primitive pairs
        const size_t base = gp0.size(), cnt = ents.size();
        const size_t nfull = ents.empty() ? 0 : ents[0].p0.size();
        gp0.resize(base + nfull * cnt);
        gp1.resize(gp0.size());
        gK.resize(gp0.size());
        for (size_t j = 0; j < cnt; ++j) {
            Entry& e = ents[j];
            for (size_t k = 0; k < nfull; ++k) {
                gp0[base + k * cnt + j] = e.p0[k];
                gp1[base + k * cnt + j] = e.p1[k];
                gK[base + k * cnt + j] = e.K[k];
            }
            e.gp.off = (int)(base + j);
            e.gp.pstride = (int)cnt;
            c->gpairs.push_back(e.gp);
            c->gkey.push_back(e.key);
        }
        c->gcnt[g] = (int)c->gpairs.size() - c->goff[g];
        c->gc_nfull[g] = (int)nfull;
    }
    return true;
}

void destroy_graphs(hfg_ctx* c) {
    for (auto& g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

int setup_geometry(hfg_ctx* c) {
    destroy_graphs(c);
    build_blocks(c);
    // block pairs (X >= Y) of every molecule
    c->bp_X.clear();
    c->bp_Y.clear();
    for (auto& M : c->mols) {
        M.bp0 = (int)c->bp_X.size();
        for (int X = M.blk0; X < M.blk0 + M.nblk; ++X)
            for (int Y = M.blk0; Y <= X; ++Y) {
                c->bp_X.push_back(X);
                c->bp_Y.push_back(Y);
            }
        M.nbp = (int)c->bp_X.size() - M.bp0;
    }
    const int nbp = (int)c->bp_X.size();
    // candidate pairs per block pair, with their Cartesian diagonal slots
    std::vector<std::vector<CandPair>> cand(nbp);
    std::vector<double4> pp0, pp1, pp2;
    long long ndiag = 0;
    for (int g = 0; g < nbp; ++g) {
        const int X = c->bp_X[g], Y = c->bp_Y[g];
        auto& v = cand[g];
        for (size_t i = 0; i < c->blk_shells[X].size(); ++i)
            for (size_t j = 0; j < c->blk_shells[Y].size(); ++j) {
                if (X == Y && j > i) continue;
                const int a = c->blk_shells[X][i], b = c->blk_shells[Y][j];
                CandPair cp{make_pair(c, a, b, pp0, pp1, pp2), a, b, ndiag};
                ndiag += (long long)fm_ncart(c->L[a]) * fm_ncart(c->L[b]);
                v.push_back(cp);
            }
    }
    // Primitive pairs that cannot reach the primitive cutoff in any quartet: every kernel skips a
    // primitive quartet when |K_ab K_cd| / sqrt(p + q) is below its cutoff (the J/K and gradient
    // cut integral_tol/1e4, the Schwarz diagonal 1e-16), and |K_cd| / sqrt(p + q) <= Mq =
    // max |K| / sqrt(q) over all primitive pairs, so a primitive pair with |K_ab| Mq below the
    // smaller cutoff contributes to no quartet. Such primitive pairs are moved behind the others
    // (which keep their basis order: the FP32 kernels' primitive sums are sensitive to the order,
    // and a largest-first order made them 30 times less accurate) and the loop count np stops
    // before them; the storage keeps all of them.
    double mq = 0.0;
    for (size_t k = 0; k < pp0.size(); ++k) mq = std::max(mq, std::fabs(pp1[k].w) / std::sqrt(pp0[k].x));
    const double sig_cut = std::min(1e-16, c->opt.integral_tol / 1e4);
    {
        const double cut = sig_cut;
        std::vector<int> perm;
        std::vector<double4> t0, t1, t2;
        for (auto& v : cand)
            for (auto& cp : v) {
                const int n = c->nprim[cp.ga] * c->nprim[cp.gb], o = cp.p.off;
                perm.resize(n);
                std::iota(perm.begin(), perm.end(), 0);
                auto sig = [&](int x) { return std::fabs(pp1[o + x].w) * mq >= cut; };
                const int nsig = (int)(std::stable_partition(perm.begin(), perm.end(), sig) - perm.begin());
                t0.assign(pp0.begin() + o, pp0.begin() + o + n);
                t1.assign(pp1.begin() + o, pp1.begin() + o + n);
                t2.assign(pp2.begin() + o, pp2.begin() + o + n);
                for (int k = 0; k < n; ++k) {
                    pp0[o + k] = t0[perm[k]];
                    pp1[o + k] = t1[perm[k]];
                    pp2[o + k] = t2[perm[k]];
                }
                cp.p.np = nsig;
            }
    }
    HFG_CHECK(c, c->d_pp0.upload(pp0, c->stream));
    HFG_CHECK(c, c->d_pp1.upload(pp1, c->stream));
    HFG_CHECK(c, c->d_pp2.upload(pp2, c->stream));
    // Cartesian Schwarz diagonals (pq|pq) of every candidate pair on the GPU, grouped by class
    std::vector<double> diag((size_t)std::max<long long>(ndiag, 1), 0.0);
    {
        std::vector<PairDev> sp;
        std::vector<long long> off;
        std::vector<std::pair<int, int>> cls_range;  // (la*3+lb, begin)
        for (int la = 0; la <= 2; ++la)
            for (int lb = 0; lb <= la; ++lb) {
                const int begin = (int)sp.size();
                for (auto& v : cand)
                    for (auto& cp : v)
                        if (c->L[cp.ga] == la && c->L[cp.gb] == lb) {
                            sp.push_back(cp.p);
                            off.push_back(cp.diag_off);
                        }
                cls_range.push_back({la * 3 + lb, begin});
            }
        cls_range.push_back({-1, (int)sp.size()});
        DevBuf<PairDev> dsp;
        DevBuf<long long> doff;
        DevBuf<double> ddiag;
        HFG_CHECK(c, dsp.upload(sp, c->stream));
        HFG_CHECK(c, doff.upload(off, c->stream));
        HFG_CHECK(c, ddiag.ensure(diag.size()));
        // the screen kernels run on the shared scratch stream: the primitive data, pair records
        // and output offsets uploaded on the context's stream must have landed first
        HFG_CHECK(c, hipStreamSynchronize(c->stream));
        for (size_t r = 0; r + 1 < cls_range.size(); ++r) {
            ScreenArgs a{};
            a.pairs = dsp.p + cls_range[r].second;
            a.out_off = doff.p + cls_range[r].second;
            a.count = cls_range[r + 1].second - cls_range[r].second;
            a.pp0 = c->d_pp0.p;
            a.pp1 = c->d_pp1.p;
            a.pp2 = c->d_pp2.p;
            a.boys = c->d_boys.p;
            a.out = ddiag.p;
            HFG_CHECK(c, launch_screen(cls_range[r].first / 3, cls_range[r].first % 3, a, c->heavy));
        }
        HFG_CHECK(c, hipMemcpyAsync(diag.data(), ddiag.p, diag.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    c->heavy));
        HFG_CHECK(c, hipStreamSynchronize(c->heavy));
        dsp.release();
        doff.release();
        ddiag.release();
    }
    // the ncart x ncart Schwarz matrix of a single molecule (hfg_schwarz)
    c->screen.clear();
    if (c->mols.size() == 1) {
        const int nc = c->mols[0].ncart;
        c->screen.assign((size_t)nc * nc, 0.0);
        for (auto& v : cand)
            for (auto& cp : v) {
                const int na = fm_ncart(c->L[cp.ga]), nb = fm_ncart(c->L[cp.gb]);
                const int ca = c->cart_off[cp.ga], cb = c->cart_off[cp.gb];
                for (int x = 0; x < na; ++x)
                    for (int y = 0; y < nb; ++y) {
                        const double val = diag[cp.diag_off + x * nb + y];
                        c->screen[(size_t)(ca + x) * nc + cb + y] = val;
                        c->screen[(size_t)(cb + y) * nc + ca + x] = val;
                    }
            }
    }
    // Sort key s_ab of every pair and the threshold on products s_ab s_cd below which a quartet is
    // skipped. With screening: the mean Schwarz value (hfscf's shell skip compares the mean of
    // sqrt|S_pq S_rs|, which factorises into a product of per-pair means; hfscf_oseri.cpp:36-46)
    // against 100 x integral_tol. Without (hfscf's eri_screen off, and its gradient, which visits
    // every quartet): a bound on the primitive prefactors against the primitive cutoff. A
    // primitive quartet is skipped when |K_ab K_cd| / sqrt(p + q) < integral_tol / 1e4
    // (hfscf_oseri.cpp:73,98-101, hfscf_oserideriv1.cpp:107,133-136), and
    // |K_ab K_cd| / sqrt(p + q) <= s_ab s_cd with s_ab = max|K_ab| / (2 min p)^(1/4) over the pair's
    // primitives (sqrt(p + q) >= sqrt(2) (p q)^(1/4)), so a quartet with s_ab s_cd below the cutoff
    // has no primitive quartet left and is exactly zero.
    c->tol = c->opt.schwarz ? 100.0 * c->opt.integral_tol : c->opt.integral_tol / 1e4;
    c->elem_tol = c->opt.schwarz ? c->tol : 0.0;
    std::vector<std::vector<std::pair<double, const CandPair*>>> lists(nbp);
    std::vector<double> smax(c->mols.size(), 0.0);
    for (int g = 0; g < nbp; ++g)
        for (auto& cp : cand[g]) {
            const int na = fm_ncart(c->L[cp.ga]), nb = fm_ncart(c->L[cp.gb]);
            double s = 0;
            if (c->opt.schwarz) {
                for (int k = 0; k < na * nb; ++k) s += std::sqrt(std::fabs(diag[cp.diag_off + k]));
                s /= (double)(na * nb);
            } else {
                double kb = 0.0, pmin = 1e300;
                for (int k = 0; k < cp.p.np; ++k) {  // the pair's significant primitives (stored first)
                    kb = std::max(kb, std::fabs(pp1[cp.p.off + k].w));
                    pmin = std::min(pmin, pp0[cp.p.off + k].x);
                }
                // widened by 1e-12 against the rounding of the kernels' prefactor products
                s = cp.p.np ? kb / std::pow(2.0 * pmin, 0.25) * (1.0 + 1e-12) : 0.0;
            }
            const int m = c->shell_mol[cp.ga];
            smax[m] = std::max(smax[m], s);
            lists[g].push_back({s, &cp});
        }
    c->pairs.clear();
    std::vector<int> pair_nfull;  // stored primitive pairs of each pair (np counts the significant ones)
    c->psmean.clear();
    c->psq.clear();
    c->bp_off.assign(nbp, 0);
    c->bp_cnt.assign(nbp, 0);
    for (int g = 0; g < nbp; ++g) {
        auto& v = lists[g];
        const int m = c->blk_mol[c->bp_X[g]];
        std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        c->bp_off[g] = (int)c->pairs.size();
        for (auto& e : v) {
            // a pair whose mean times the molecule's largest mean is below the threshold has no
            // surviving quartet
            if (e.first * smax[m] < c->tol) continue;
            PairDev pd = e.second->p;
            const int na = fm_ncart(c->L[e.second->ga]), nb = fm_ncart(c->L[e.second->gb]);
            pd.sq_off = (int)c->psq.size();
            pd.sq_min = 1e300;
            pd.sq_max = 0.0;
            for (int k = 0; k < na * nb; ++k) {
                const double q = std::sqrt(std::fabs(diag[e.second->diag_off + k]));
                c->psq.push_back(q);
                pd.sq_min = std::min(pd.sq_min, q);
                pd.sq_max = std::max(pd.sq_max, q);
            }
            c->pairs.push_back(pd);
            pair_nfull.push_back(c->nprim[e.second->ga] * c->nprim[e.second->gb]);
            c->psmean.push_back(e.first);
        }
        c->bp_cnt[g] = (int)c->pairs.size() - c->bp_off[g];
    }
    // primitive data of the surviving pairs, primitive-major within each block pair's list
    {
        size_t tot = 0;
        for (int nf : pair_nfull) tot += nf;
        std::vector<double4> q0(std::max<size_t>(tot, 1)), q1(q0.size()), q2(q0.size());
        size_t pos = 0;
        for (int g = 0; g < nbp; ++g) {
            const int cnt = c->bp_cnt[g];
            if (!cnt) continue;
            PairDev* lst = &c->pairs[c->bp_off[g]];
            const int* nfl = &pair_nfull[c->bp_off[g]];
            bool uniform = true;
            for (int j = 1; j < cnt; ++j) uniform = uniform && nfl[j] == nfl[0];
            if (uniform) {
                const int np = nfl[0];
                for (int j = 0; j < cnt; ++j) {
                    for (int k = 0; k < np; ++k) {
                        q0[pos + (size_t)k * cnt + j] = pp0[lst[j].off + k];
                        q1[pos + (size_t)k * cnt + j] = pp1[lst[j].off + k];
                        q2[pos + (size_t)k * cnt + j] = pp2[lst[j].off + k];
                    }
                    lst[j].off = (int)(pos + j);
                    lst[j].pstride = cnt;
                }
                pos += (size_t)np * cnt;
            } else {
                for (int j = 0; j < cnt; ++j) {
                    for (int k = 0; k < nfl[j]; ++k) {
                        q0[pos + k] = pp0[lst[j].off + k];
                        q1[pos + k] = pp1[lst[j].off + k];
                        q2[pos + k] = pp2[lst[j].off + k];
                    }
                    lst[j].off = (int)pos;
                    lst[j].pstride = 1;
                    pos += nfl[j];
                }
            }
        }
        HFG_CHECK(c, c->d_pp0.upload(q0, c->stream));
        HFG_CHECK(c, c->d_pp1.upload(q1, c->stream));
        HFG_CHECK(c, c->d_pp2.upload(q2, c->stream));
    }
    // shell groups of the s blocks and their group-pair lists
    std::vector<double4> gp0, gp1, gK;
    const bool gc_on = build_groups(c, mq, sig_cut, gp0, gp1, gK);
    std::vector<double> gcum_np, gcum_npnm, gcum_nm;  // per block pair prefixes over its group list
    if (gc_on) {
        HFG_CHECK(c, c->d_gpairs.upload(c->gpairs, c->stream));
        HFG_CHECK(c, c->d_gkey.upload(c->gkey, c->stream));
        HFG_CHECK(c, c->d_goff.upload(c->goff, c->stream));
        HFG_CHECK(c, c->d_gcnt.upload(c->gcnt, c->stream));
        HFG_CHECK(c, c->d_gc_p0.upload(gp0, c->stream));
        HFG_CHECK(c, c->d_gc_p1.upload(gp1, c->stream));
        HFG_CHECK(c, c->d_gc_K.upload(gK, c->stream));
        c->gc_args =
            GcArgs{c->d_gpairs.p, c->d_gkey.p, c->d_goff.p, c->d_gcnt.p, c->d_gc_p0.p, c->d_gc_p1.p, c->d_gc_K.p};
        gcum_np.assign(c->gpairs.size() + nbp, 0.0);
        gcum_npnm.assign(gcum_np.size(), 0.0);
        gcum_nm.assign(gcum_np.size(), 0.0);
        for (int P = 0; P < nbp; ++P) {
            const int k0 = c->goff[P] + P;  // list P's prefixes start at goff[P] + P
            for (int i = 0; i < c->gcnt[P]; ++i) {
                const GcPairDev& gp = c->gpairs[c->goff[P] + i];
                gcum_np[k0 + i + 1] = gcum_np[k0 + i] + gp.np;
                gcum_npnm[k0 + i + 1] = gcum_npnm[k0 + i] + (double)gp.np * gp.nmem;
                gcum_nm[k0 + i + 1] = gcum_nm[k0 + i] + gp.nmem;
            }
        }
    } else {
        c->gc_args = GcArgs{};
    }
    // work items per class
    std::vector<double> cum_np(c->pairs.size() + nbp, 0.0);  // per block pair prefix of np
    std::vector<int> cum_off(nbp, 0);
    {
        int k = 0;
        for (int P = 0; P < nbp; ++P) {
            cum_off[P] = k;
            cum_np[k++] = 0;
            for (int i = 0; i < c->bp_cnt[P]; ++i, ++k) cum_np[k] = cum_np[k - 1] + c->pairs[c->bp_off[P] + i].np;
        }
    }
    for (auto& cw : c->classes) {
        if (cw.ev_start) (void)hipEventDestroy(cw.ev_start);
        if (cw.ev_stop) (void)hipEventDestroy(cw.ev_stop);
    }
    c->classes.clear();
    c->class_timed = false;
    auto class_index = [&](int la, int lb, int lc, int ld) {
        for (size_t i = 0; i < c->classes.size(); ++i) {
            auto& w = c->classes[i];
            if (w.la == la && w.lb == lb && w.lc == lc && w.ld == ld) return (int)i;
        }
        ClassWork w;
        w.la = la;
        w.lb = lb;
        w.lc = lc;
        w.ld = ld;
        c->classes.push_back(w);
        return (int)c->classes.size() - 1;
    };
    // pass 1: total cost and cost per class
    double total_cost = 0;
    std::vector<double> class_cost;
    for (const auto& M : c->mols)
        for (int P = M.bp0; P < M.bp0 + M.nbp; ++P)
            for (int Q = M.bp0; Q <= P; ++Q) {
                if (!c->bp_cnt[P] || !c->bp_cnt[Q]) continue;
                const int A = c->bp_X[P], B = c->bp_Y[P], C = c->bp_X[Q], D = c->bp_Y[Q];
                const FlopParts fp = model_flops(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D], c->pure, false);
                const int ci = class_index(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D]);
                if ((int)class_cost.size() <= ci) class_cost.resize(ci + 1, 0.0);
                for (int i = 0; i < c->bp_cnt[P]; ++i) {
                    int cnt =
                        ket_count_host(&c->psmean[c->bp_off[Q]], c->bp_cnt[Q], c->psmean[c->bp_off[P] + i], c->tol);
                    if (P == Q) cnt = std::min(cnt, i + 1);
                    if (!cnt) continue;
                    const double cost =
                        c->pairs[c->bp_off[P] + i].np * cum_np[cum_off[Q] + cnt] * fp.per_prim + cnt * fp.per_quartet;
                    total_cost += cost;
                    class_cost[ci] += cost;
                }
            }
    // A class on the group lists: ket count and cost of group row i of list P against list Q
    // (model: per group primitive quartet the prefactor, Boys values and vertical recursion of one
    // primitive quartet, plus per member pair its weight, cutoff and a multiply-add per
    // accumulated value; per member shell quartet the horizontal recursion and digestion)
    auto gc_row = [&](int P, int Q, int i, const ClassWork& cw, const FlopParts& fp) {
        int cnt = ket_count_host(&c->gkey[c->goff[Q]], c->gcnt[Q], c->gkey[c->goff[P] + i], c->tol);
        if (P == Q) cnt = std::min(cnt, i + 1);
        const GcPairDev& gb = c->gpairs[c->goff[P] + i];
        const int k0 = c->goff[Q] + Q;
        auto ncum = [](int l) { return (l + 1) * (l + 2) * (l + 3) / 6; };
        const double vals = (double)(ncum(cw.la + cw.lb) - ncum(cw.la - 1)) * (ncum(cw.lc + cw.ld) - ncum(cw.lc - 1));
        const double cost = gb.np * (fp.per_prim * gcum_np[k0 + cnt] + (3.0 + 2.0 * vals) * gb.nmem * gcum_npnm[k0 + cnt]) +
                            fp.per_quartet * gb.nmem * gcum_nm[k0 + cnt];
        return std::make_pair(cnt, cost);
    };
    // each class's cost on the group lists (the classes with a group kernel)
    std::vector<double> gc_cost(class_cost.size(), 0.0);
    if (gc_on)
        for (const auto& M : c->mols)
            for (int P = M.bp0; P < M.bp0 + M.nbp; ++P)
                for (int Q = M.bp0; Q <= P; ++Q) {
                    if (!c->bp_cnt[P] || !c->bp_cnt[Q] || !c->gcnt[P] || !c->gcnt[Q]) continue;
                    const int A = c->bp_X[P], B = c->bp_Y[P], C = c->bp_X[Q], D = c->bp_Y[Q];
                    const int ci = class_index(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D]);
                    const ClassWork& cw = c->classes[ci];
                    if (!jk_gc_available(cw.la, cw.lb, cw.lc, cw.ld, c->pure)) continue;
                    const FlopParts fp = model_flops(cw.la, cw.lb, cw.lc, cw.ld, c->pure, false);
                    for (int i = 0; i < c->gcnt[P]; ++i) gc_cost[ci] += gc_row(P, Q, i, cw, fp).second;
                }
    const double split = c->opt.work_split > 0 ? (double)c->opt.work_split : kDefaultWorkSplit;
    const double target = std::max(total_cost / (256.0 * 6.0 * split), 1.0);
    c->slab_doubles = 0;
    c->grad_slab_doubles = 0;
    c->total_quartets = 0;
    c->total_flops = 0;
    c->total_grad_flops = 0;
    // Work items of the rows [0, nrows) of a bra list against a ket list: consecutive rows until
    // their cost reaches tgt (or kMaxRows rows), an expensive row as several items over chunks of
    // its ket columns. row(i) gives {kets of row i, cost of row i}; push(r0, r1, c0, c1) adds one.
    auto split_rows = [](int nrows, double tgt, auto&& row, auto&& push) {
        double acc = 0;
        long long q_acc = 0;
        int r0 = -1;
        auto flush = [&](int r1) {
            if (r0 >= 0 && q_acc > 0) push(r0, r1, 0, 0x7fffffff);
            r0 = -1;
            acc = 0;
            q_acc = 0;
        };
        for (int i = 0; i < nrows; ++i) {
            const std::pair<int, double> rc = row(i);
            const int cnt = rc.first;
            const double row_cost = rc.second;
            if (row_cost > tgt && cnt > 64) {
                flush(i);
                const int chunks = std::min((int)std::ceil(row_cost / tgt), (cnt + 31) / 32);
                const int width = (cnt + chunks - 1) / chunks;
                for (int j0 = 0; j0 < cnt; j0 += width) push(i, i + 1, j0, std::min(cnt, j0 + width));
                continue;
            }
            if (r0 < 0) r0 = i;
            acc += row_cost;
            q_acc += cnt;
            if (acc >= tgt || i + 1 - r0 >= kMaxRows) flush(i + 1);
        }
        flush(nrows);
    };
    for (int m = 0; m < (int)c->mols.size(); ++m) {
        const MolHost& M = c->mols[m];
        for (int P = M.bp0; P < M.bp0 + M.nbp; ++P)
            for (int Q = M.bp0; Q <= P; ++Q) {
                if (!c->bp_cnt[P] || !c->bp_cnt[Q]) continue;
                const int A = c->bp_X[P], B = c->bp_Y[P], C = c->bp_X[Q], D = c->bp_Y[Q];
                const int ci = class_index(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D]);
                ClassWork& cw = c->classes[ci];
                const FlopParts fp = model_flops(cw.la, cw.lb, cw.lc, cw.ld, c->pure, false);
                const FlopParts gp = model_flops(cw.la, cw.lb, cw.lc, cw.ld, c->pure, true);
                const int tile = c->blk_nf[A] * c->blk_nf[B] + c->blk_nf[C] * c->blk_nf[D] +
                                 c->blk_nf[A] * c->blk_nf[C] + c->blk_nf[A] * c->blk_nf[D] +
                                 c->blk_nf[B] * c->blk_nf[C] + c->blk_nf[B] * c->blk_nf[D];
                // per-class item target: a class launch has >= 256 x split items when it has the work
                const double tgt = std::max(std::min(target, class_cost[ci] / (256.0 * split)), 1.0);
                // (ss|ss) on the group lists: the J/K items come from those, the per-shell items
                // below serve the gradient only
                const bool on_groups =
                    gc_on && jk_gc_available(cw.la, cw.lb, cw.lc, cw.ld, c->pure) && c->gcnt[P] && c->gcnt[Q];
                WorkItem w{};
                w.split = 1;
                w.P = P;
                w.Q = Q;
                w.A = A;
                w.B = B;
                w.C = C;
                w.D = D;
                w.mol = m;
                auto push = [&](int r0, int r1, int c0, int c1) {
                    WorkItem v = w;
                    v.r0 = r0;
                    v.r1 = r1;
                    v.c0 = c0;
                    v.c1 = c1;
                    v.slab = on_groups ? -1 : c->slab_doubles;
                    if (!on_groups) c->slab_doubles += tile;
                    v.gslab = c->grad_slab_doubles;
                    c->grad_slab_doubles += 3 * M.natom;
                    cw.items.push_back(v);
                };
                split_rows(c->bp_cnt[P], tgt, [&](int i) {
                    int cnt = ket_count_host(&c->psmean[c->bp_off[Q]], c->bp_cnt[Q], c->psmean[c->bp_off[P] + i], c->tol);
                    if (P == Q) cnt = std::min(cnt, i + 1);
                    const double np_i = c->pairs[c->bp_off[P] + i].np;
                    const double kq = np_i * cum_np[cum_off[Q] + cnt];
                    cw.quartets += cnt;
                    cw.flops += kq * fp.per_prim + cnt * fp.per_quartet;
                    cw.grad_flops += kq * gp.per_prim + cnt * gp.per_quartet;
                    return std::make_pair(cnt, kq * fp.per_prim + cnt * fp.per_quartet);
                }, push);
                if (!on_groups) continue;
                cw.gc = true;
                // lanes per group quartet: enough that one lane evaluates at most about
                // kGcLanePrims primitive quartets (the lists' full primitive counts)
                const double pq = (double)c->gc_nfull[P] * c->gc_nfull[Q];
                const int lanes = std::max(1, std::min(kGcMaxSplit, (int)std::ceil(pq / kGcLanePrims)));
                auto push_gc = [&](int r0, int r1, int c0, int c1) {
                    WorkItem v = w;
                    v.split = lanes;
                    v.r0 = r0;
                    v.r1 = r1;
                    v.c0 = c0;
                    v.c1 = c1;
                    v.slab = c->slab_doubles;
                    c->slab_doubles += tile;
                    v.gslab = -1;
                    cw.gc_items.push_back(v);
                };
                // as many items as the per-shell lists would give the class (its items are cheaper)
                const double tgt_gc = std::max(tgt * gc_cost[ci] / std::max(class_cost[ci], 1.0), 1.0);
                split_rows(c->gcnt[P], tgt_gc, [&](int i) {
                    const std::pair<int, double> rc = gc_row(P, Q, i, cw, fp);
                    cw.gc_flops += rc.second;
                    return rc;
                }, push_gc);
            }
    }
    for (auto& cw : c->classes) {
        c->total_quartets += cw.quartets;
        c->total_flops += cw.jk_flops();
        c->total_grad_flops += cw.grad_flops;
    }
    c->classes.erase(std::remove_if(c->classes.begin(), c->classes.end(), [](const ClassWork& w) { return w.items.empty(); }),
                     c->classes.end());
    // classes heaviest first (the launch order), then onto the streams by the FLOP model
    {
        std::vector<size_t> ord(c->classes.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(),
                  [&](size_t x, size_t y) { return c->classes[x].jk_flops() > c->classes[y].jk_flops(); });
        std::vector<ClassWork> sorted;
        for (size_t k : ord) sorted.push_back(c->classes[k]);
        c->classes.swap(sorted);
        // cooperative programs of every class from L = kCoopTuneMinL (when their LDS fits a
        // workgroup); the classes from coop_min_l up start on the cooperative kernels, autotuning
        // may move the others there
        const int coop_min = c->opt.coop_min_l > 0 ? c->opt.coop_min_l : kCoopRequiredL;
        auto class_l = [](const ClassWork& cw) { return cw.la + cw.lb + cw.lc + cw.ld; };
        std::vector<int4> words;
        std::vector<std::pair<size_t, CoopHostProgram>> progs;
        for (size_t k = 0; k < c->classes.size(); ++k) {
            ClassWork& cw = c->classes[k];
            cw.coop = cw.coop_ok = false;
            if (class_l(cw) < kCoopTuneMinL) continue;
            CoopHostProgram hp = make_coop_program(cw.la, cw.lb, cw.lc, cw.ld, c->pure);
            cw.coop_lds[0] = hp.lds_bytes(false);
            cw.coop_lds[1] = hp.lds_bytes(true);
            if (cw.coop_lds[0] + kCoopStaticLds > (size_t)c->max_lds) continue;
            const size_t base = words.size();
            words.insert(words.end(), hp.words.begin(), hp.words.end());
            progs.emplace_back(base, std::move(hp));
            cw.coop_ok = true;
            cw.coop = class_l(cw) >= coop_min || jk_uses_coop(cw.la, cw.lb, cw.lc, cw.ld);
        }
        // gradient programs of the same classes for the cooperative gradient kernel
        std::vector<int4> gwords;
        std::vector<std::pair<size_t, CoopGradHostProgram>> gprogs;
        for (size_t k = 0; k < c->classes.size(); ++k) {
            ClassWork& cw = c->classes[k];
            cw.gcoop = cw.gcoop_ok = false;
            if (class_l(cw) < kCoopTuneMinL) continue;
            const size_t limit = (size_t)c->max_lds - kCoopGradStaticLds;
            CoopGradHostProgram hp;
            try {
                hp = make_coop_grad_program(cw.la, cw.lb, cw.lc, cw.ld, c->pure, limit);
            } catch (const std::exception&) {
                continue;  // the class has no gradient kernel on this device (hfg_energy_backward refuses it)
            }
            cw.gcoop_lds[0] = hp.lds_bytes(false);
            cw.gcoop_lds[1] = hp.lds_bytes(true);
            const size_t base = gwords.size();
            gwords.insert(gwords.end(), hp.words.begin(), hp.words.end());
            gprogs.emplace_back(base, std::move(hp));
            cw.gcoop_ok = true;
            cw.gcoop = class_l(cw) >= coop_min || grad_uses_coop(cw.la, cw.lb, cw.lc, cw.ld);
        }
        if (!gwords.empty()) {
            HFG_CHECK(c, c->d_gcoop.upload(gwords, c->stream));
            size_t i = 0;
            for (size_t k = 0; k < c->classes.size(); ++k)
                if (c->classes[k].gcoop_ok) {
                    c->classes[k].gcoop_args = coop_grad_args_on_device(gprogs[i].second, c->d_gcoop.p + gprogs[i].first);
                    ++i;
                }
        }
        // the largest dynamic LDS any context of the process may request (the attribute is per
        // kernel, so every context sets the same value)
        HFG_CHECK(c, coop_set_lds_limit((size_t)c->max_lds - kCoopStaticLds, (size_t)c->max_lds - kCoopGradStaticLds));
        if (!words.empty()) {
            HFG_CHECK(c, c->d_coop.upload(words, c->stream));
            size_t i = 0;
            for (size_t k = 0; k < c->classes.size(); ++k)
                if (c->classes[k].coop_ok) {
                    c->classes[k].coop_args = coop_args_on_device(progs[i].second, c->d_coop.p + progs[i].first);
                    ++i;
                }
        }
        std::vector<double> wj, wg;
        for (const auto& cw : c->classes) {
            wj.push_back(cw.jk_flops());
            wg.push_back(cw.grad_flops);
        }
        assign_streams(c, false, wj);
        assign_streams(c, true, wg);
    }
    // gather lists (one output tile per block pair) and per-molecule item lists
    std::vector<std::vector<GatherEntry>> lj(nbp), lk(nbp);
    std::vector<WorkItem> all;
    std::vector<std::vector<int>> mitems(c->mols.size());
    c->total_items = 0;
    for (auto& cw : c->classes) {
        cw.dev_off = all.size();
        for (auto& w : cw.items) {
            mitems[w.mol].push_back((int)all.size());
            all.push_back(w);
        }
        cw.gc_dev_off = all.size();
        all.insert(all.end(), cw.gc_items.begin(), cw.gc_items.end());
        for (auto& w : cw.jk_items()) {
            const int f[4] = {c->blk_nf[w.A], c->blk_nf[w.B], c->blk_nf[w.C], c->blk_nf[w.D]};
            const int blk[4] = {w.A, w.B, w.C, w.D};
            // (row block, col block) of Jab Jcd Kac Kad Kbc Kbd
            const int rr[6] = {0, 2, 0, 0, 1, 1}, cc[6] = {1, 3, 2, 3, 2, 3};
            long long off = w.slab;
            for (int r = 0; r < 6; ++r) {
                const int R = blk[rr[r]], C = blk[cc[r]];
                auto& Lst = (r < 2) ? lj : lk;
                if (R > C) {
                    Lst[bp_index(c, R, C)].push_back({off, f[cc[r]], 0});
                } else if (R < C) {
                    Lst[bp_index(c, C, R)].push_back({off, f[cc[r]], 1});
                } else {
                    Lst[bp_index(c, R, R)].push_back({off, f[cc[r]], 0});
                    Lst[bp_index(c, R, R)].push_back({off, f[cc[r]], 1});
                }
                off += (long long)f[rr[r]] * f[cc[r]];
            }
        }
        c->total_items += (long long)cw.jk_items().size();
    }
    c->tiles.clear();
    c->gJ_off.assign(1, 0);
    c->gK_off.assign(1, 0);
    c->gJ.clear();
    c->gK.clear();
    for (int g = 0; g < nbp; ++g) {
        c->tiles.push_back(make_int4(c->bp_X[g], c->bp_Y[g], c->blk_mol[c->bp_X[g]], 0));
        c->gJ.insert(c->gJ.end(), lj[g].begin(), lj[g].end());
        c->gK.insert(c->gK.end(), lk[g].begin(), lk[g].end());
        c->gJ_off.push_back((int)c->gJ.size());
        c->gK_off.push_back((int)c->gK.size());
    }
    // first-pass chunks of kGatherChunk entries: the J chunks of every tile, then the K chunks
    c->gchunks.clear();
    c->cJ_off.assign(1, 0);
    c->cK_off.clear();
    c->gpart_doubles = 0;
    for (int list = 0; list < 2; ++list) {
        std::vector<int>& coff = list ? c->cK_off : c->cJ_off;
        if (list) coff.push_back((int)c->gchunks.size());
        const std::vector<int>& goff = list ? c->gK_off : c->gJ_off;
        for (int t = 0; t < nbp; ++t) {
            const int elems = c->blk_nf[c->bp_X[t]] * c->blk_nf[c->bp_Y[t]];
            for (int g0 = goff[t]; g0 < goff[t + 1]; g0 += kGatherChunk) {
                c->gchunks.push_back(GatherChunk{t, list, g0, std::min(g0 + kGatherChunk, goff[t + 1]), c->gpart_doubles});
                c->gpart_doubles += elems;
            }
            coff.push_back((int)c->gchunks.size());
        }
    }
    std::vector<int> mol_items, mol_items_off(1, 0), out_mol, out_k;
    for (size_t m = 0; m < c->mols.size(); ++m) {
        mol_items.insert(mol_items.end(), mitems[m].begin(), mitems[m].end());
        mol_items_off.push_back((int)mol_items.size());
        for (int k = 0; k < 3 * c->mols[m].natom; ++k) {
            out_mol.push_back((int)m);
            out_k.push_back(k);
        }
    }
    std::vector<MolDev> md;
    for (auto& M : c->mols)
        md.push_back(MolDev{M.moff, M.nbf, M.natom, M.shell0, M.nshell, M.dm_off, 0});
    // upload
    HFG_CHECK(c, c->d_pairs.upload(c->pairs, c->stream));
    HFG_CHECK(c, c->d_psmean.upload(c->psmean, c->stream));
    HFG_CHECK(c, c->d_psq.upload(c->psq, c->stream));
    HFG_CHECK(c, c->d_bp_off.upload(c->bp_off, c->stream));
    HFG_CHECK(c, c->d_bp_cnt.upload(c->bp_cnt, c->stream));
    HFG_CHECK(c, c->d_blk_nf.upload(c->blk_nf, c->stream));
    HFG_CHECK(c, c->d_blk_func_off.upload(c->blk_func_off, c->stream));
    HFG_CHECK(c, c->d_blk_func_ao.upload(c->blk_func_ao, c->stream));
    HFG_CHECK(c, c->d_items.upload(all, c->stream));
    HFG_CHECK(c, c->d_tiles.upload(c->tiles, c->stream));
    HFG_CHECK(c, c->d_gJ_off.upload(c->gJ_off, c->stream));
    HFG_CHECK(c, c->d_gK_off.upload(c->gK_off, c->stream));
    HFG_CHECK(c, c->d_gJ.upload(c->gJ, c->stream));
    HFG_CHECK(c, c->d_gK.upload(c->gK, c->stream));
    HFG_CHECK(c, c->d_gchunks.upload(c->gchunks, c->stream));
    HFG_CHECK(c, c->d_cJ_off.upload(c->cJ_off, c->stream));
    HFG_CHECK(c, c->d_cK_off.upload(c->cK_off, c->stream));
    HFG_CHECK(c, c->d_gpart.ensure((size_t)std::max<long long>(c->gpart_doubles, 1)));
    HFG_CHECK(c, c->d_shell_atom.upload(c->atom_local, c->stream));
    {
        std::vector<int> ao(c->nshell), nf(c->nshell);
        for (int s = 0; s < c->nshell; ++s) {
            ao[s] = c->pure ? c->pure_off[s] : c->cart_off[s];
            nf[s] = nfun(c->L[s], c->pure);
        }
        HFG_CHECK(c, c->d_shell_ao.upload(ao, c->stream));
        HFG_CHECK(c, c->d_shell_nf.upload(nf, c->stream));
        HFG_CHECK(c, c->d_shell_mol.upload(c->shell_mol, c->stream));
    }
    HFG_CHECK(c, c->d_mols.upload(md, c->stream));
    HFG_CHECK(c, c->d_mol_items.upload(mol_items, c->stream));
    HFG_CHECK(c, c->d_mol_items_off.upload(mol_items_off, c->stream));
    HFG_CHECK(c, c->d_out_mol.upload(out_mol, c->stream));
    HFG_CHECK(c, c->d_out_k.upload(out_k, c->stream));
    HFG_CHECK(c, c->d_queue.ensure(2 * std::max<size_t>(c->classes.size(), 1)));
    HFG_CHECK(c, c->d_slab.ensure((size_t)std::max<long long>(c->slab_doubles, 1)));
    const size_t nn = (size_t)c->nn_total;
    HFG_CHECK(c, c->d_D.ensure(nn));
    HFG_CHECK(c, c->d_J.ensure(nn));
    HFG_CHECK(c, c->d_K.ensure(nn));
    HFG_CHECK(c, c->d_F.ensure(nn));
    HFG_CHECK(c, c->d_H.ensure(nn));
    HFG_CHECK(c, c->d_G.ensure(nn));
    size_t ndm = 0;
    for (auto& M : c->mols) ndm += (size_t)M.nshell * M.nshell;
    HFG_CHECK(c, c->d_dmax.ensure(ndm));  // allocated once here, so graph arguments stay valid
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_prev = false;
    return HFG_OK;
}

// Largest |D| of every shell block of every molecule, uploaded for the density screen.
int upload_dmax(hfg_ctx* c, const double* D) {
    size_t ndm = 0;
    for (auto& M : c->mols) ndm += (size_t)M.nshell * M.nshell;
    c->dmax.assign(ndm, 0.0f);
    for (const auto& M : c->mols) {
        const int ns = M.nshell, n = M.nbf;
        const double* Dm = D + M.moff;
        float* out = c->dmax.data() + M.dm_off;
        for (int a = 0; a < ns; ++a) {
            const int ga = M.shell0 + a;
            const int oa = c->pure ? c->pure_off[ga] : c->cart_off[ga], na = nfun(c->L[ga], c->pure);
            for (int b = 0; b < ns; ++b) {
                const int gb = M.shell0 + b;
                const int ob = c->pure ? c->pure_off[gb] : c->cart_off[gb], nb = nfun(c->L[gb], c->pure);
                double mx = 0.0;
                for (int p = 0; p < na; ++p)
                    for (int q = 0; q < nb; ++q) mx = std::max(mx, std::fabs(Dm[(size_t)(oa + p) * n + ob + q]));
                // rounded up so the single-precision table never understates the bound
                out[(size_t)a * ns + b] = mx > 0.0 ? std::nextafter((float)mx, INFINITY) : 0.0f;
            }
        }
    }
    if (c->d_dmax.n < ndm) return fail(c, HFG_ERR_ARG, "density table larger than at set-up");
    HFG_CHECK(c, hipMemcpyAsync(c->d_dmax.p, c->dmax.data(), ndm * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return HFG_OK;
}

// Device buffers one J/K build reads and writes: the density, J and K, the work items' tile
// slab, the class launches' queue counters and the gather's chunk sums (the context's own, or
// caller-provided ones).
struct JkBuffers {
    const double* D;
    double* J;
    double* K;
    double* slab;
    int* queue;
    double* part;  // the gather's chunk sums
};

JkBuffers own_buffers(hfg_ctx* c) {
    return JkBuffers{c->d_D.p, c->d_J.p, c->d_K.p, c->d_slab.p, c->d_queue.p, c->d_gpart.p};
}

// The class launches of the set `set` on `main` (forked to the class streams and joined back),
// each preceded by the reset of its queue counter: the shared-stream classes' counters
// [0, nshared) on the shared stream, the side classes' [nshared, ncls) on main before the fork.
// With `timed`, every class launch is bracketed by events on its stream.
int enqueue_classes(hfg_ctx* c, hipStream_t main, bool dscreen, bool timed, ClassSet set, const JkBuffers& buf,
                    bool join_shared = true) {
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
    a.D = buf.D;
    a.mols = c->d_mols.p;
    a.tol = c->tol;
    a.psq = c->d_psq.p;
    a.elem_tol = c->elem_tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = buf.slab;
    a.dmax = c->d_dmax.p;
    a.dtol = dscreen ? c->opt.density_tol : 0.0;
    const int ncls = (int)c->classes.size(), nshared = shared_class_count(c, false);
    if (set != ClassSet::Shared) HFG_CHECK(c, launch_zero_ints(buf.queue + nshared, ncls - nshared, main));
    bool shared_zeroed = false;
    HFG_CHECK(c, for_each_class_concurrent(c, main, [&](ClassWork& cw, hipStream_t st, int qi) {
        hipError_t e = hipSuccess;
        if (st == c->heavy && !shared_zeroed) {
            if ((e = launch_zero_ints(buf.queue, nshared, st)) != hipSuccess) return e;
            shared_zeroed = true;
        }
        JkArgs b = a;
        b.items = c->d_items.p + cw.jk_dev_off();
        b.queue = buf.queue + qi;
        if (timed) {
            if (!cw.ev_start && (e = hipEventCreate(&cw.ev_start)) != hipSuccess) return e;
            if (!cw.ev_stop && (e = hipEventCreate(&cw.ev_stop)) != hipSuccess) return e;
            if ((e = hipEventRecord(cw.ev_start, st)) != hipSuccess) return e;
        }
        e = launch_jk_class(c, cw, b, cw.jk_shape, st);
        if (e == hipSuccess && timed) e = hipEventRecord(cw.ev_stop, st);
        return e;
    }, false, set, join_shared));
    c->class_timed = timed;
    return HFG_OK;
}

// The gather of every work item's tiles into J and K, on `main`.
int enqueue_gather(hfg_ctx* c, hipStream_t main, const JkBuffers& buf) {
    GatherArgs g{};
    g.tiles = c->d_tiles.p;
    g.gJ_off = c->d_gJ_off.p;
    g.gJ = c->d_gJ.p;
    g.gK_off = c->d_gK_off.p;
    g.gK = c->d_gK.p;
    g.slab = buf.slab;
    g.blk_nf = c->d_blk_nf.p;
    g.blk_func_off = c->d_blk_func_off.p;
    g.blk_func_ao = c->d_blk_func_ao.p;
    g.mols = c->d_mols.p;
    g.J = buf.J;
    g.K = buf.K;
    HFG_CHECK(c, launch_gather_chunks(g, c->d_gchunks.p, c->d_cJ_off.p, c->d_cK_off.p, (int)c->gchunks.size(),
                                      (int)c->tiles.size(), buf.part, main));
    return HFG_OK;
}

// J and K of the densities already in d_D, into d_J / d_K, directly or by replaying the
// captured graph of this configuration. ev0 / ev1 bracket the build (the class kernels only,
// when launched directly). A graph holds only the side-stream classes: the classes of the
// device's shared stream are launched directly on it at every replay, so no capture ever
// records work of that stream (other contexts may enqueue there concurrently) and its kernels
// keep running on the one queue that holds their scratch.
int run_jk(hfg_ctx* c, bool dscreen = false) {
    if (int r = check_private_segments(c, false)) return r;
    const JkBuffers buf = own_buffers(c);
    if (!c->opt.graph) {
        HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
        int r = enqueue_classes(c, c->stream, dscreen, true, ClassSet::All, buf);
        if (r) return r;
        HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
        return enqueue_gather(c, c->stream, buf);
    }
    const int mixed = c->opt.precision == HFG_MIXED ? 1 : 0;
    const double dtol = dscreen ? c->opt.density_tol : 0.0;
    JkGraph* gr = nullptr;
    for (auto& g : c->graphs)
        if (g.valid && g.mixed == mixed && g.dscreen == dscreen && g.dtol == dtol) gr = &g;
    if (!gr) {
        hipGraph_t graph = nullptr;
        HFG_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        const int r = enqueue_classes(c, c->stream, dscreen, false, ClassSet::Side, buf);
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r != HFG_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return r != HFG_OK ? r : fail(c, HFG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
        JkGraph ng;
        const hipError_t ei = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(c, HFG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        ng.valid = true;
        ng.mixed = mixed;
        ng.dscreen = dscreen;
        ng.dtol = dtol;
        c->graphs.push_back(ng);
        gr = &c->graphs.back();
    }
    HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
    const bool shared = shared_class_count(c, false) > 0;
    if (shared) {  // concurrent with the graph: main waits for the shared stream after the launch
        const int r = enqueue_classes(c, c->stream, dscreen, false, ClassSet::Shared, buf, false);
        if (r) return r;
    }
    HFG_CHECK(c, hipGraphLaunch(gr->exec, c->stream));
    if (shared) HFG_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_heavy_join, 0));
    HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
    return enqueue_gather(c, c->stream, buf);
}

int finish_stats(hfg_ctx* c, std::chrono::steady_clock::time_point t0) {
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HFG_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.quartets = c->total_quartets;
    c->stats.model_flops = c->total_flops;
    c->stats.kernel_seconds = ms * 1e-3;
    c->stats.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->stats.work_items = c->total_items;
    c->stats.slab_bytes = c->slab_doubles * 8;
    return HFG_OK;
}

int check_ready(hfg_ctx* c) {
    if (!c) return HFG_ERR_ARG;
    if (!c->have_basis || !c->have_geom) return fail(c, HFG_ERR_ARG, "basis and geometry must be set first");
    return HFG_OK;
}

// Append one molecule's basis (offsets within the molecule) to the concatenated tables.
int add_basis(hfg_ctx* c, int nshell, const int* L, const int* atom_of_shell, const int* nprim, const double* alpha,
              const double* coef, int pure, const int* cart_offset, const int* pure_offset, int atom0) {
    if (nshell <= 0 || !L || !atom_of_shell || !nprim || !alpha || !coef || !cart_offset || !pure_offset)
        return fail(c, HFG_ERR_ARG, "basis: bad arguments");
    if (!c->mols.empty() && c->pure != (pure != 0))
        return fail(c, HFG_ERR_ARG, "basis: every molecule of a batch must use the same pure/Cartesian choice");
    MolHost M;
    M.shell0 = c->nshell;
    M.nshell = nshell;
    int tot = (int)c->alpha.size();
    int ncart = 0, nsph = 0;
    for (int s = 0; s < nshell; ++s) {
        if (L[s] < 0 || L[s] > 2) return fail(c, HFG_ERR_UNSUPPORTED, "basis: only s, p and d shells are supported");
        if (nprim[s] <= 0) return fail(c, HFG_ERR_ARG, "basis: shell without primitives");
        ncart = std::max(ncart, cart_offset[s] + (L[s] + 1) * (L[s] + 2) / 2);
        nsph = std::max(nsph, pure_offset[s] + 2 * L[s] + 1);
    }
    int nprim_tot = 0;
    for (int s = 0; s < nshell; ++s) nprim_tot += nprim[s];
    for (int s = 0; s < nshell; ++s) {
        c->L.push_back(L[s]);
        c->atom_local.push_back(atom_of_shell[s]);
        c->atom.push_back(atom0 + atom_of_shell[s]);
        c->nprim.push_back(nprim[s]);
        c->cart_off.push_back(cart_offset[s]);
        c->pure_off.push_back(pure_offset[s]);
        c->shell_mol.push_back((int)c->mols.size());
        c->prim_off.push_back(tot);
        tot += nprim[s];
    }
    c->alpha.insert(c->alpha.end(), alpha, alpha + nprim_tot);
    c->coef.insert(c->coef.end(), coef, coef + nprim_tot);
    c->pure = pure != 0;
    M.ncart = ncart;
    M.nbf = c->pure ? nsph : ncart;
    c->nshell += nshell;
    c->mols.push_back(M);
    return HFG_OK;
}

void reset_molecules(hfg_ctx* c) {
    c->mols.clear();
    c->nshell = 0;
    c->L.clear();
    c->atom.clear();
    c->atom_local.clear();
    c->nprim.clear();
    c->prim_off.clear();
    c->cart_off.clear();
    c->pure_off.clear();
    c->shell_mol.clear();
    c->alpha.clear();
    c->coef.clear();
    c->xyz.clear();
    c->natom = 0;
    c->have_basis = c->have_geom = false;
}

// Matrix and gradient offsets of every molecule.
void layout_molecules(hfg_ctx* c) {
    long long moff = 0;
    int dm = 0, gout = 0;
    for (auto& M : c->mols) {
        M.moff = moff;
        moff += (long long)M.nbf * M.nbf;
        M.dm_off = dm;
        dm += M.nshell * M.nshell;
        M.gout0 = gout;
        gout += 3 * M.natom;
    }
    c->nn_total = moff;
    c->nout_total = gout;
    c->nbf = c->mols.empty() ? 0 : c->mols[0].nbf;
    c->ncart = c->mols.empty() ? 0 : c->mols[0].ncart;
}

}  // namespace

hipError_t hfg::launch_jk_class(hfg_ctx* c, const ClassWork& cw, const JkArgs& a, LaunchShape sh, hipStream_t st) {
    const int mixed = c->opt.precision == HFG_MIXED ? 1 : 0;
    if (cw.gc) return launch_jk_gc(cw.la, cw.lb, cw.lc, cw.ld, c->pure, mixed, a, c->gc_args, (int)cw.gc_items.size(), sh, st);
    if (cw.coop) return launch_jk_coop(a, cw.coop_args, cw.coop_lds[mixed], mixed, (int)cw.items.size(), sh, st);
    return launch_jk(cw.la, cw.lb, cw.lc, cw.ld, c->pure, mixed, a, (int)cw.items.size(), sh, st);
}

int hfg::check_private_segments(hfg_ctx* c, bool grad) {
    const bool mixed = c->opt.precision == HFG_MIXED;
    for (const ClassWork& cw : c->classes) {
        if (grad ? cw.gcoop : cw.coop) continue;  // the cooperative kernels hold their tables in LDS
        if ((grad ? grad_uses_coop(cw.la, cw.lb, cw.lc, cw.ld) : jk_uses_coop(cw.la, cw.lb, cw.lc, cw.ld)))
            return fail(c, HFG_ERR_UNSUPPORTED, "the class needs the cooperative kernel, whose LDS does not fit a "
                                                "workgroup on this device");
        const int b = private_segment_bytes(grad, mixed, cw.la, cw.lb, cw.lc, cw.ld, c->pure);
        if (b < 0 || b > kMaxPrivateBytes) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "%s kernel of class (%d%d|%d%d) needs %d bytes of scratch per lane, above the %d this "
                          "library dispatches",
                          grad ? "gradient" : "J/K", cw.la, cw.lb, cw.lc, cw.ld, b, kMaxPrivateBytes);
            return fail(c, HFG_ERR_UNSUPPORTED, b < 0 ? std::string("no private-segment entry for a class") : msg);
        }
    }
    return HFG_OK;
}

extern "C" {

void hfg_default_options(hfg_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->precision = HFG_FP64;
    opt->integral_tol = 1e-14;
    opt->schwarz = 1;
    opt->density_tol = 0.0;
    opt->graph = 0;
}

int hfg_create(hfg_ctx** out, int device, const hfg_options* opt) {
    if (!out) return HFG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return HFG_ERR_UNSUPPORTED;
    hfg_ctx* c = new hfg_ctx();
    c->device = device;
    if (opt)
        c->opt = *opt;
    else
        hfg_default_options(&c->opt);
    if ((c->opt.precision != HFG_FP64 && c->opt.precision != HFG_MIXED) || c->opt.side_streams < 0 ||
        c->opt.group_shells < -1 || c->opt.group_shells > 1 ||
        (c->opt.coop_min_l != 0 && (c->opt.coop_min_l < kCoopTuneMinL || c->opt.coop_min_l > kCoopRequiredL))) {
        delete c;
        return HFG_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_heavy_join, hipEventDisableTiming) != hipSuccess ||
        (c->heavy = shared_scratch_stream(device)) == nullptr) {
        hfg_destroy(c);
        return HFG_ERR_HIP;
    }
    if (hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        c->max_lds = 65536;
    if (noalloc_scratch(device, &c->scratch) != hipSuccess) {  // hfg_build_jk_noalloc's fixed scratch
        hfg_destroy(c);
        return HFG_ERR_HIP;
    }
    c->nside = c->opt.side_streams > 0 ? std::min(c->opt.side_streams, kSideStreams) : kDefaultSideStreams;
    for (int k = 0; k < c->nside; ++k)
        if (hipStreamCreate(&c->side[k]) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) != hipSuccess) {
            hfg_destroy(c);
            return HFG_ERR_HIP;
        }
    {
        // which hardware queue serves each side stream (HIP deals streams to a few queues per
        // process, and a queue runs its kernels in order)
        DevBuf<unsigned long long> q;
        std::vector<unsigned long long> addr(c->nside + 1, 0);
        bool ok = q.ensure(c->nside + 1) == hipSuccess;
        for (int k = 0; ok && k < c->nside; ++k) ok = launch_queue_probe(q.p, k, c->side[k]) == hipSuccess;
        ok = ok && launch_queue_probe(q.p, c->nside, c->heavy) == hipSuccess;
        for (int k = 0; ok && k < c->nside; ++k) ok = hipStreamSynchronize(c->side[k]) == hipSuccess;
        ok = ok && hipStreamSynchronize(c->heavy) == hipSuccess &&
             hipMemcpy(addr.data(), q.p, addr.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) == hipSuccess;
        q.release();
        if (!ok) {
            hfg_destroy(c);
            return HFG_ERR_HIP;
        }
        for (int k = 0; k <= c->nside; ++k) {
            int id = k;
            for (int j = 0; j < k; ++j)
                if (addr[j] == addr[k]) {
                    id = (j < c->nside) ? c->side_queue[j] : j;
                    break;
                }
            if (k < c->nside)
                c->side_queue[k] = id;
            else
                c->heavy_queue = id < c->nside ? id : -1;
        }
    }
    std::vector<double> tab = boys_table();
    if (c->d_boys.upload(tab, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        hfg_destroy(c);
        return HFG_ERR_HIP;
    }
    *out = c;
    return HFG_OK;
}

int hfg_set_basis(hfg_ctx* c, int nshell, const int* L, const int* atom_of_shell, const int* nprim, const double* alpha,
                  const double* coef, int pure, const int* cart_offset, const int* pure_offset) {
    if (!c) return HFG_ERR_ARG;
    if (c->batch) return fail(c, HFG_ERR_ARG, "hfg_set_basis: the context holds a batch of molecules");
    reset_molecules(c);
    const int r = add_basis(c, nshell, L, atom_of_shell, nprim, alpha, coef, pure, cart_offset, pure_offset, 0);
    if (r != HFG_OK) {
        reset_molecules(c);
        return r;
    }
    layout_molecules(c);
    c->have_basis = true;
    return HFG_OK;
}

int hfg_set_geometry(hfg_ctx* c, int natom, const double* xyz_bohr, const int* Z) {
    if (!c || !c->have_basis || natom <= 0 || !xyz_bohr) return fail(c, HFG_ERR_ARG, "hfg_set_geometry: bad arguments");
    if (c->batch) return fail(c, HFG_ERR_ARG, "hfg_set_geometry: the context holds a batch of molecules");
    for (int s = 0; s < c->nshell; ++s)
        if (c->atom[s] < 0 || c->atom[s] >= natom) return fail(c, HFG_ERR_ARG, "hfg_set_geometry: shell atom out of range");
    HFG_CHECK(c, hipSetDevice(c->device));
    c->natom = natom;
    c->mols[0].atom0 = 0;
    c->mols[0].natom = natom;
    c->xyz.assign(xyz_bohr, xyz_bohr + 3 * natom);
    c->Zn.clear();
    if (Z)
        for (int a = 0; a < natom; ++a) c->Zn.push_back((double)Z[a]);
    layout_molecules(c);
    c->have_geom = false;
    int r = setup_geometry(c);
    if (r == HFG_OK) c->have_geom = true;
    return r;
}

int hfg_batch_add_molecule(hfg_ctx* c, int nshell, const int* L, const int* atom_of_shell, const int* nprim,
                           const double* alpha, const double* coef, int pure, const int* cart_offset,
                           const int* pure_offset, int natom, const double* xyz_bohr) {
    if (!c) return HFG_ERR_ARG;
    if (!c->batch && !c->mols.empty())
        return fail(c, HFG_ERR_ARG, "hfg_batch_add_molecule: the context already holds a single molecule");
    if (c->have_geom) return fail(c, HFG_ERR_ARG, "hfg_batch_add_molecule: the batch is already committed");
    if (natom <= 0 || natom > kMaxAtomsPerMolecule || !xyz_bohr)
        return fail(c, HFG_ERR_ARG, "hfg_batch_add_molecule: bad atom count or coordinates");
    for (int s = 0; s < nshell; ++s)
        if (atom_of_shell && (atom_of_shell[s] < 0 || atom_of_shell[s] >= natom))
            return fail(c, HFG_ERR_ARG, "hfg_batch_add_molecule: shell atom out of range");
    c->batch = true;
    const int r = add_basis(c, nshell, L, atom_of_shell, nprim, alpha, coef, pure, cart_offset, pure_offset, c->natom);
    if (r != HFG_OK) return r;
    MolHost& M = c->mols.back();
    M.atom0 = c->natom;
    M.natom = natom;
    c->natom += natom;
    c->xyz.insert(c->xyz.end(), xyz_bohr, xyz_bohr + 3 * natom);
    c->have_basis = true;
    return (int)c->mols.size() - 1;
}

int hfg_batch_commit(hfg_ctx* c) {
    if (!c || !c->batch || c->mols.empty()) return fail(c, HFG_ERR_ARG, "hfg_batch_commit: no molecules added");
    if (c->have_geom) return fail(c, HFG_ERR_ARG, "hfg_batch_commit: already committed");
    HFG_CHECK(c, hipSetDevice(c->device));
    layout_molecules(c);
    int r = setup_geometry(c);
    if (r == HFG_OK) c->have_geom = true;
    return r;
}

int hfg_batch_layout(const hfg_ctx* c, int* nmol, int* nbf, long long* moff, int* natom) {
    if (!c) return HFG_ERR_ARG;
    if (nmol) *nmol = (int)c->mols.size();
    for (size_t m = 0; m < c->mols.size(); ++m) {
        if (nbf) nbf[m] = c->mols[m].nbf;
        if (moff) moff[m] = c->mols[m].moff;
        if (natom) natom[m] = c->mols[m].natom;
    }
    return HFG_OK;
}

int hfg_build_jk(hfg_ctx* c, const double* D, double* J, double* K, int flags) {
    (void)flags;
    int r = check_ready(c);
    if (r) return r;
    if (!D) return fail(c, HFG_ERR_ARG, "hfg_build_jk: D is NULL");
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)c->nn_total;
    HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((r = run_jk(c))) return r;
    if (J) HFG_CHECK(c, hipMemcpyAsync(J, c->d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (K) HFG_CHECK(c, hipMemcpyAsync(K, c->d_K.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return finish_stats(c, t0);
}

int hfg_build_jk_noalloc(hfg_ctx* c, const double* D, double* J, double* K) {
    int r = check_ready(c);
    if (r) return r;
    if (!D || !J || !K) return fail(c, HFG_ERR_ARG, "hfg_build_jk_noalloc: NULL argument");
    const size_t ncls = c->classes.size();
    const size_t need = (size_t)c->slab_doubles + (size_t)c->gpart_doubles;
    if (ncls > 2 * kNoallocCounterDoubles || need > kNoallocScratchDoubles - kNoallocCounterDoubles) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "hfg_build_jk_noalloc: the build needs %.1f MiB of scratch, the fixed maximum is %.1f MiB "
                      "(use hfg_build_jk)",
                      (need + kNoallocCounterDoubles) * 8.0 / 1048576.0, kNoallocScratchDoubles * 8.0 / 1048576.0);
        return fail(c, HFG_ERR_ARG, msg);
    }
    if ((r = check_private_segments(c, false))) return r;
    auto t0 = std::chrono::steady_clock::now();
    double* const slab = c->scratch + kNoallocCounterDoubles;
    const JkBuffers buf{D, J, K, slab, reinterpret_cast<int*>(c->scratch), slab + c->slab_doubles};
    HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
    if ((r = enqueue_classes(c, c->stream, false, false, ClassSet::All, buf))) return r;
    HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
    if ((r = enqueue_gather(c, c->stream, buf))) return r;
    return finish_stats(c, t0);
}

int hfg_fock(hfg_ctx* c, const double* H, const double* D, int incremental, double* F) {
    int r = check_ready(c);
    if (r) return r;
    if (!H || !D || !F) return fail(c, HFG_ERR_ARG, "hfg_fock: NULL argument");
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)c->nn_total;
    const bool inc = incremental && c->have_prev;
    if (inc && c->prev_on_device) {
        c->Dprev.resize(nn);
        HFG_CHECK(c, hipMemcpyAsync(c->Dprev.data(), c->d_Dprev.p, nn * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
        HFG_CHECK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<double> Dd;
    const double* Duse = D;
    if (inc) {
        Dd.resize(nn);
        for (size_t k = 0; k < nn; ++k) Dd[k] = D[k] - c->Dprev[k];
        Duse = Dd.data();
    }
    const bool dscreen = inc && c->opt.density_tol > 0.0;
    if (dscreen && (r = upload_dmax(c, Duse))) return r;
    HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, Duse, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(c->d_H.p, H, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((r = run_jk(c, dscreen))) return r;
    HFG_CHECK(c, launch_fock(c->d_H.p, c->d_J.p, c->d_K.p, c->d_F.p, c->d_G.p, (int)nn, inc ? 1 : 0, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(F, c->d_F.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->Dprev.assign(D, D + nn);
    c->have_prev = true;
    c->prev_on_device = false;
    return finish_stats(c, t0);
}

int hfg_fock_device(hfg_ctx* c, const double* H, const double* D, int incremental, double* F) {
    int r = check_ready(c);
    if (r) return r;
    if (!H || !D || !F) return fail(c, HFG_ERR_ARG, "hfg_fock_device: NULL argument");
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)c->nn_total;
    const bool inc = incremental && c->have_prev;
    HFG_CHECK(c, c->d_Dprev.ensure(nn));
    if (inc && !c->prev_on_device)
        HFG_CHECK(c, hipMemcpyAsync(c->d_Dprev.p, c->Dprev.data(), nn * sizeof(double), hipMemcpyHostToDevice,
                                    c->stream));
    if (inc)
        HFG_CHECK(c, launch_axpby(1.0, D, -1.0, c->d_Dprev.p, c->d_D.p, nn, c->stream));  // D - D_prev
    else
        HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, D, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    const bool dscreen = inc && c->opt.density_tol > 0.0;
    if (dscreen)
        HFG_CHECK(c, launch_block_absmax(c->d_D.p, c->d_mols.p, c->d_shell_ao.p, c->d_shell_nf.p, c->d_shell_mol.p,
                                         c->nshell, c->d_dmax.p, c->stream));
    if ((r = run_jk(c, dscreen))) return r;
    HFG_CHECK(c, launch_fock(H, c->d_J.p, c->d_K.p, F, c->d_G.p, (int)nn, inc ? 1 : 0, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(c->d_Dprev.p, D, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->have_prev = true;
    c->prev_on_device = true;
    return finish_stats(c, t0);
}

int hfg_reset_incremental(hfg_ctx* c) {
    if (!c) return HFG_ERR_ARG;
    c->have_prev = false;
    return HFG_OK;
}

int hfg_set_precision(hfg_ctx* c, int precision) {
    if (!c) return HFG_ERR_ARG;
    if (precision != HFG_FP64 && precision != HFG_MIXED) return fail(c, HFG_ERR_ARG, "hfg_set_precision: unknown mode");
    c->opt.precision = precision;
    return HFG_OK;
}

int hfg_set_density_tol(hfg_ctx* c, double density_tol) {
    if (!c) return HFG_ERR_ARG;
    if (!(density_tol >= 0.0)) return fail(c, HFG_ERR_ARG, "hfg_set_density_tol: the threshold must be >= 0");
    c->opt.density_tol = density_tol;
    return HFG_OK;
}

int hfg_autotune(hfg_ctx* c, int what) {
    int r = check_ready(c);
    if (r) return r;
    if (what < 1 || what > 3) return fail(c, HFG_ERR_ARG, "hfg_autotune: what must be 1, 2 or 3");
    // candidates: the register kernel at threads per workgroup x persistent workgroups (0 = one
    // per work item), and for the classes from L = kCoopTrialMinL that have a cooperative program
    // that kernel at each grid (its workgroup size is fixed; block 0 marks it in the shapes)
    const int blocks[2] = {256, 128};
    const int grids[3] = {0, 512, 2048};
    destroy_graphs(c);
    for (int kind = 1; kind <= 2; ++kind) {
        if (!(what & kind)) continue;
        const bool grad = kind == 2;
        std::vector<double> measured(c->classes.size(), 0.0);
        for (size_t k = 0; k < c->classes.size(); ++k) {
            ClassWork& cw = c->classes[k];
            bool& use_coop = grad ? cw.gcoop : cw.coop;
            const bool coop_ok = grad ? cw.gcoop_ok : cw.coop_ok;
            const bool reg_ok = cw.la + cw.lb + cw.lc + cw.ld < kCoopRequiredL;
            LaunchShape& shape = grad ? cw.grad_shape : cw.jk_shape;
            bool best_coop = use_coop;
            LaunchShape best = shape;
            float best_ms = 1e30f;
            auto trial = [&](bool coop, LaunchShape sh) -> int {
                use_coop = coop;
                float total = 0.0f;
                for (int rep = 0; rep < 2; ++rep) {
                    if (int e = hfg::run_one_class(c, k, grad, sh)) return e;
                    float ms = 0.0f;
                    HFG_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
                    if (rep > 0) total += ms;  // the first run warms caches and code
                }
                if (total < best_ms) {
                    best_ms = total;
                    best = sh;
                    best_coop = coop;
                }
                return HFG_OK;
            };
            if (reg_ok)
                for (int b : blocks)
                    for (int g : grids)
                        if ((r = trial(false, LaunchShape{g, b}))) return r;
            if (coop_ok && (!reg_ok || cw.la + cw.lb + cw.lc + cw.ld >= kCoopTrialMinL))
                for (int g : grids)
                    if ((r = trial(true, LaunchShape{g, 0}))) return r;
            use_coop = best_coop;
            shape = best;
            measured[k] = best_ms;
        }
        // rebalance the side streams on the measured times of the chosen shapes
        assign_streams(c, kind == 2, measured);
    }
    return HFG_OK;
}

int hfg_launch_shapes(const hfg_ctx* c, int cap, int* la_lb_lc_ld, int* jk_grid, int* jk_block, int* grad_grid,
                      int* grad_block) {
    if (!c) return HFG_ERR_ARG;
    const int n = (int)c->classes.size();
    for (int k = 0; k < n && k < cap; ++k) {
        const ClassWork& cw = c->classes[k];
        if (la_lb_lc_ld) la_lb_lc_ld[k] = cw.la * 1000 + cw.lb * 100 + cw.lc * 10 + cw.ld;
        if (jk_grid) jk_grid[k] = cw.jk_shape.grid;
        if (jk_block) jk_block[k] = cw.gc ? -2 : cw.coop ? -1 : cw.jk_shape.block;
        if (grad_grid) grad_grid[k] = cw.grad_shape.grid;
        if (grad_block) grad_block[k] = cw.gcoop ? -1 : cw.grad_shape.block;
    }
    return n;
}

int hfg_class_stats(hfg_ctx* c, int cap, int* la_lb_lc_ld, long long* quartets, double* model_flops,
                    double* kernel_ms) {
    if (!c) return HFG_ERR_ARG;
    const int n = (int)c->classes.size();
    if (cap > 0) HFG_CHECK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < n && k < cap; ++k) {
        ClassWork& cw = c->classes[k];
        if (la_lb_lc_ld) la_lb_lc_ld[k] = cw.la * 1000 + cw.lb * 100 + cw.lc * 10 + cw.ld;
        if (quartets) quartets[k] = cw.quartets;
        if (model_flops) model_flops[k] = cw.jk_flops();
        if (kernel_ms) {
            float ms = -1.0f;
            if (c->class_timed && cw.ev_start && cw.ev_stop) HFG_CHECK(c, hipEventElapsedTime(&ms, cw.ev_start, cw.ev_stop));
            kernel_ms[k] = ms;
        }
    }
    return n;
}

int hfg_energy(hfg_ctx* c, const double* D, double* E2) {
    int r = check_ready(c);
    if (r) return r;
    if (!D || !E2) return fail(c, HFG_ERR_ARG, "hfg_energy: NULL argument");
    const size_t nn = (size_t)c->nn_total;
    std::vector<double> J(nn), K(nn);
    if ((r = hfg_build_jk(c, D, J.data(), K.data(), 0))) return r;
    double e = 0;
    for (size_t k = 0; k < nn; ++k) e += D[k] * (2.0 * J[k] - K[k]);
    *E2 = e;
    return HFG_OK;
}

int hfg_energy_backward(hfg_ctx* c, const double* D, double* grad) {
    int r = check_ready(c);
    if (r) return r;
    if (!D || !grad) return fail(c, HFG_ERR_ARG, "hfg_energy_backward: NULL argument");
    return run_gradient(c, D, grad);
}

int hfg_eri_shell_quartet(hfg_ctx* c, int s1, int s2, int s3, int s4, double* block) {
    int r = check_ready(c);
    if (r) return r;
    if (c->mols.size() != 1) return fail(c, HFG_ERR_ARG, "hfg_eri_shell_quartet: single-molecule contexts only");
    if (!block || s1 < 0 || s2 < 0 || s3 < 0 || s4 < 0 || s1 >= c->nshell || s2 >= c->nshell || s3 >= c->nshell ||
        s4 >= c->nshell)
        return fail(c, HFG_ERR_ARG, "hfg_eri_shell_quartet: bad shell index");
    // (pq|rs) = J_pq / 2 for the unit density D = E_rs + E_sr (J_pq for r == s), so the block
    // is read out of the production J/K kernels.
    const int n = c->nbf;
    const size_t nn = (size_t)n * n;
    auto off = [&](int s) { return c->pure ? c->pure_off[s] : c->cart_off[s]; };
    auto cnt = [&](int s) { return nfun(c->L[s], c->pure); };
    const int n1 = cnt(s1), n2 = cnt(s2), n3 = cnt(s3), n4 = cnt(s4);
    std::vector<double> D(nn, 0.0), J(nn);
    for (int rr = 0; rr < n3; ++rr)
        for (int ss = 0; ss < n4; ++ss) {
            const int R = off(s3) + rr, S = off(s4) + ss;
            D[(size_t)R * n + S] += 1.0;
            if (R != S) D[(size_t)S * n + R] += 1.0;
            if ((r = hfg_build_jk(c, D.data(), J.data(), nullptr, 0))) return r;
            D[(size_t)R * n + S] = 0.0;
            D[(size_t)S * n + R] = 0.0;
            const double f = (R == S) ? 1.0 : 0.5;
            for (int p = 0; p < n1; ++p)
                for (int q = 0; q < n2; ++q)
                    block[((p * n2 + q) * n3 + rr) * n4 + ss] = f * J[(size_t)(off(s1) + p) * n + off(s2) + q];
        }
    return HFG_OK;
}

int hfg_eri_vector(hfg_ctx* c, double* out, long long count) {
    int r = check_ready(c);
    if (r) return r;
    if (c->mols.size() != 1) return fail(c, HFG_ERR_ARG, "hfg_eri_vector: single-molecule contexts only");
    const long long npair = (long long)c->nbf * (c->nbf + 1) / 2, n = npair * (npair + 1) / 2;
    if (!out || count != n) return fail(c, HFG_ERR_ARG, "hfg_eri_vector: the vector must hold npair (npair + 1) / 2 values");
    HFG_CHECK(c, hipSetDevice(c->device));
    if (int r = check_private_segments(c, false)) return r;
    HFG_CHECK(c, c->d_eri_out.ensure((size_t)n));
    HFG_CHECK(c, hipMemsetAsync(c->d_eri_out.p, 0, (size_t)n * sizeof(double), c->stream));
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
    a.D = c->d_D.p;  // read by the kernel's density-block loads, unused by the export
    a.mols = c->d_mols.p;
    a.tol = c->tol;
    a.psq = c->d_psq.p;
    a.elem_tol = c->elem_tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = c->d_slab.p;
    a.dmax = c->d_dmax.p;
    a.dtol = 0.0;
    a.eri_out = c->d_eri_out.p;
    // the class kernels may spill: they run on the device's shared scratch stream, after the
    // vector's zeroing on the context's stream
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < c->classes.size(); ++k) {
        // the FP64 per-lane kernel of the class, or its cooperative kernel where the J/K builds run
        // that one (the per-shell items: a class on the shell-group lists has them too)
        const ClassWork& cw = c->classes[k];
        JkArgs b = a;
        b.items = c->d_items.p + cw.dev_off;
        b.queue = c->d_queue.p + k;
        const int n_items = (int)cw.items.size();
        HFG_CHECK(c, launch_zero_ints(b.queue, 1, c->heavy));
        if (cw.coop)
            HFG_CHECK(c, launch_jk_coop(b, cw.coop_args, cw.coop_lds[0], 0, n_items, LaunchShape{0, 0}, c->heavy));
        else
            HFG_CHECK(c, launch_jk(cw.la, cw.lb, cw.lc, cw.ld, c->pure, 0, b, n_items, LaunchShape{0, 0}, c->heavy));
    }
    HFG_CHECK(c, hipStreamSynchronize(c->heavy));
    HFG_CHECK(c, hipMemcpyAsync(out, c->d_eri_out.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    c->d_eri_out.release();  // the vector grows as N^4; the context keeps no copy
    return HFG_OK;
}

int hfg_boys_probe(hfg_ctx* c, int m, int n, const double* T, double* F, int closed_form) {
    if (!c) return HFG_ERR_ARG;
    if (!T || !F || n < 0 || m < 0 || m > kBoysMaxM || (closed_form && m > 1))
        return fail(c, HFG_ERR_ARG, "hfg_boys_probe: bad arguments");
    for (int i = 0; i < n; ++i)
        if (!(T[i] >= 0.0)) return fail(c, HFG_ERR_ARG, "hfg_boys_probe: T must be >= 0");
    if (n == 0) return HFG_OK;
    DevBuf<double> dT, dF;
    struct Release {  // DevBuf does not free itself
        DevBuf<double>& a;
        DevBuf<double>& b;
        ~Release() {
            a.release();
            b.release();
        }
    } release{dT, dF};
    HFG_CHECK(c, dT.ensure((size_t)n));
    HFG_CHECK(c, dF.ensure((size_t)n * (m + 1)));
    HFG_CHECK(c, hipMemcpyAsync(dT.p, T, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, launch_boys_probe(m, dT.p, c->d_boys.p, dF.p, n, closed_form, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(F, dF.p, (size_t)n * (m + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    return HFG_OK;
}

int hfg_copy_calibration(int device, long long bytes, int reps, double* ms_per_copy) {
    if (bytes < 16 || reps < 1) return HFG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return HFG_ERR_UNSUPPORTED;
    if (hipSetDevice(device) != hipSuccess) return HFG_ERR_HIP;
    const size_t n = (size_t)bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HFG_OK;
    if (hipMalloc((void**)&a, n * sizeof(double2)) != hipSuccess || hipMalloc((void**)&b, n * sizeof(double2)) != hipSuccess)
        rc = HFG_ERR_NOMEM;
    if (rc == HFG_OK && (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                         hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, n * sizeof(double2), st) != hipSuccess))
        rc = HFG_ERR_HIP;
    if (rc == HFG_OK && launch_copy(a, b, n, st) != hipSuccess) rc = HFG_ERR_HIP;  // warm-up
    if (rc == HFG_OK && hipEventRecord(e0, st) != hipSuccess) rc = HFG_ERR_HIP;
    for (int k = 0; rc == HFG_OK && k < reps; ++k)
        if (launch_copy(a, b, n, st) != hipSuccess) rc = HFG_ERR_HIP;
    if (rc == HFG_OK && (hipEventRecord(e1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)) rc = HFG_ERR_HIP;
    float ms = 0.0f;
    if (rc == HFG_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = HFG_ERR_HIP;
    if (rc == HFG_OK && ms_per_copy) *ms_per_copy = ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int hfg_schwarz(hfg_ctx* c, double* screen) {
    int r = check_ready(c);
    if (r) return r;
    if (!screen) return fail(c, HFG_ERR_ARG, "hfg_schwarz: NULL argument");
    if (c->mols.size() != 1) return fail(c, HFG_ERR_ARG, "hfg_schwarz: single-molecule contexts only");
    std::memcpy(screen, c->screen.data(), c->screen.size() * sizeof(double));
    return HFG_OK;
}

int hfg_last_stats(const hfg_ctx* c, hfg_stats* st) {
    if (!c || !st) return HFG_ERR_ARG;
    *st = c->stats;
    return HFG_OK;
}

int hfg_nbf(const hfg_ctx* c) { return c ? c->nbf : -1; }

const char* hfg_last_error(const hfg_ctx* c) { return c ? c->err.c_str() : "null context"; }

void hfg_destroy(hfg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    for (auto& cw : c->classes) {
        if (cw.ev_start) (void)hipEventDestroy(cw.ev_start);
        if (cw.ev_stop) (void)hipEventDestroy(cw.ev_stop);
    }
    c->d_boys.release();
    c->d_D.release();
    c->d_J.release();
    c->d_K.release();
    c->d_F.release();
    c->d_H.release();
    c->d_G.release();
    c->d_slab.release();
    c->d_psmean.release();
    c->d_psq.release();
    c->d_dmax.release();
    c->d_grad_slab.release();
    c->d_grad.release();
    c->d_gcoop.release();
    c->d_pp0.release();
    c->d_pp1.release();
    c->d_pp2.release();
    c->d_pairs.release();
    c->d_mols.release();
    c->d_bp_off.release();
    c->d_bp_cnt.release();
    c->d_blk_nf.release();
    c->d_blk_func_off.release();
    c->d_blk_func_ao.release();
    c->d_gchunks.release();
    c->d_cJ_off.release();
    c->d_cK_off.release();
    c->d_gpart.release();
    c->d_gpairs.release();
    c->d_gkey.release();
    c->d_goff.release();
    c->d_gcnt.release();
    c->d_gc_p0.release();
    c->d_one_sh.release();
    c->d_eri_out.release();
    c->d_one_pairs.release();
    c->d_one_alpha.release();
    c->d_one_coef.release();
    c->d_one_xyz.release();
    c->d_one_Z.release();
    c->d_one_buf.release();
    c->d_hess_cart.release();
    c->d_hess_items.release();
    c->d_gc_p1.release();
    c->d_gc_K.release();
    c->d_gJ_off.release();
    c->d_gK_off.release();
    c->d_shell_atom.release();
    c->d_coop.release();
    c->d_shell_ao.release();
    c->d_shell_nf.release();
    c->d_shell_mol.release();
    c->d_Dprev.release();
    c->d_queue.release();
    c->d_mol_items.release();
    c->d_mol_items_off.release();
    c->d_out_mol.release();
    c->d_out_k.release();
    c->d_items.release();
    c->d_tiles.release();
    c->d_gJ.release();
    c->d_gK.release();
    for (int k = 0; k < kSideStreams; ++k) {
        if (c->side[k]) (void)hipStreamSynchronize(c->side[k]);
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    }
    if (c->heavy) (void)hipStreamSynchronize(c->heavy);
    if (c->ev_heavy_join) (void)hipEventDestroy(c->ev_heavy_join);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
