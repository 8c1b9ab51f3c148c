// Host side of the hfgpu C ABI: set-up of blocks, primitive pairs, Schwarz screening,
// work decomposition and gather lists; J/K, Fock, energy and gradient drivers.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "boys.hpp"
#include "context.hpp"
#include "flopmodel.hpp"

using namespace hfg;

namespace {

int fail(hfg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
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

void build_blocks(hfg_ctx* c) {
    std::vector<int> order(c->nshell);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (c->L[x] != c->L[y]) return c->L[x] < c->L[y];
        return c->nprim[x] > c->nprim[y];
    });
    c->blk_L.clear();
    c->blk_nf.clear();
    c->blk_shells.clear();
    c->shell_blk.assign(c->nshell, -1);
    c->shell_slot.assign(c->nshell, -1);
    size_t k = 0;
    while (k < order.size()) {
        // one angular momentum and one primitive count per block: every shell pair of a block
        // pair then has the same number of primitive pairs, so the lanes of a wavefront run
        // primitive loops of equal length
        const int Lb = c->L[order[k]], np = c->nprim[order[k]];
        const int nf = nfun(Lb, c->pure);
        const int per = std::max(1, kBlockFuncs / nf);
        std::vector<int> sh;
        while (k < order.size() && c->L[order[k]] == Lb && c->nprim[order[k]] == np && (int)sh.size() < per)
            sh.push_back(order[k++]);
        const int b = (int)c->blk_L.size();
        for (size_t s = 0; s < sh.size(); ++s) {
            c->shell_blk[sh[s]] = b;
            c->shell_slot[sh[s]] = (int)s;
        }
        c->blk_L.push_back(Lb);
        c->blk_nf.push_back((int)sh.size() * nf);
        c->blk_shells.push_back(sh);
    }
    c->nblk = (int)c->blk_L.size();
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

// Primitive-pair data of shell pair (a, b) appended to pp0/pp1/pp2.
PairDev make_pair(hfg_ctx* c, int a, int b, std::vector<double4>& pp0, std::vector<double4>& pp1,
                  std::vector<double4>& pp2) {
    static const double kSqrt2Pi25 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    PairDev p{};
    p.sa = a;
    p.sb = b;
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

int setup_geometry(hfg_ctx* c) {
    build_blocks(c);
    const int nb = c->nblk;
    const int nbp = nb * (nb + 1) / 2;
    // candidate pairs per block pair
    std::vector<std::vector<PairDev>> cand(nbp);
    std::vector<double4> pp0, pp1, pp2;
    for (int X = 0; X < nb; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            auto& v = cand[X * (X + 1) / 2 + Y];
            for (size_t i = 0; i < c->blk_shells[X].size(); ++i)
                for (size_t j = 0; j < c->blk_shells[Y].size(); ++j) {
                    if (X == Y && j > i) continue;
                    v.push_back(make_pair(c, c->blk_shells[X][i], c->blk_shells[Y][j], pp0, pp1, pp2));
                }
        }
    HFG_CHECK(c, c->d_pp0.upload(pp0, c->stream));
    HFG_CHECK(c, c->d_pp1.upload(pp1, c->stream));
    HFG_CHECK(c, c->d_pp2.upload(pp2, c->stream));
    // Schwarz diagonal on the GPU, grouped by pair class
    {
        std::vector<PairDev> sp;
        std::vector<int> ca, cb;
        std::vector<std::pair<int, int>> cls_range;  // (la*3+lb, begin)
        std::vector<int> cls_l;
        for (int la = 0; la <= 2; ++la)
            for (int lb = 0; lb <= la; ++lb) {
                const int begin = (int)sp.size();
                for (auto& v : cand)
                    for (auto& p : v)
                        if (c->L[p.sa] == la && c->L[p.sb] == lb) {
                            sp.push_back(p);
                            ca.push_back(c->cart_off[p.sa]);
                            cb.push_back(c->cart_off[p.sb]);
                        }
                cls_range.push_back({la * 3 + lb, begin});
            }
        cls_range.push_back({-1, (int)sp.size()});
        DevBuf<PairDev> dsp;
        DevBuf<int> dca, dcb;
        HFG_CHECK(c, dsp.upload(sp, c->stream));
        HFG_CHECK(c, dca.upload(ca, c->stream));
        HFG_CHECK(c, dcb.upload(cb, c->stream));
        HFG_CHECK(c, c->d_slab.ensure((size_t)c->ncart * c->ncart));
        HFG_CHECK(c, hipMemsetAsync(c->d_slab.p, 0, (size_t)c->ncart * c->ncart * sizeof(double), c->stream));
        for (size_t r = 0; r + 1 < cls_range.size(); ++r) {
            ScreenArgs a{};
            a.pairs = dsp.p + cls_range[r].second;
            a.cart_a = dca.p + cls_range[r].second;
            a.cart_b = dcb.p + cls_range[r].second;
            a.count = cls_range[r + 1].second - cls_range[r].second;
            a.pp0 = c->d_pp0.p;
            a.pp1 = c->d_pp1.p;
            a.pp2 = c->d_pp2.p;
            a.boys = c->d_boys.p;
            a.ncart = c->ncart;
            a.screen = c->d_slab.p;
            HFG_CHECK(c, launch_screen(cls_range[r].first / 3, cls_range[r].first % 3, a, c->stream));
        }
        c->screen.assign((size_t)c->ncart * c->ncart, 0.0);
        HFG_CHECK(c, hipMemcpyAsync(c->screen.data(), c->d_slab.p, c->screen.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
        HFG_CHECK(c, hipStreamSynchronize(c->stream));
        dsp.release();
        dca.release();
        dcb.release();
    }
    // mean Schwarz value per pair (hfscf's shell skip compares the mean of sqrt|S_pq S_rs|,
    // which factorises into a product of per-pair means; hfscf_oseri.cpp:36-46)
    c->tol = c->opt.schwarz ? 100.0 * c->opt.integral_tol : 0.0;
    std::vector<std::vector<std::pair<double, PairDev>>> lists(nbp);
    double smax = 0;
    for (int P = 0; P < nbp; ++P)
        for (auto& p : cand[P]) {
            const int ma = (c->L[p.sa] + 1) * (c->L[p.sa] + 2) / 2, mb = (c->L[p.sb] + 1) * (c->L[p.sb] + 2) / 2;
            double s = 0;
            for (int x = 0; x < ma; ++x)
                for (int y = 0; y < mb; ++y)
                    s += std::sqrt(std::fabs(c->screen[(size_t)(c->cart_off[p.sa] + x) * c->ncart + c->cart_off[p.sb] + y]));
            s /= (double)(ma * mb);
            smax = std::max(smax, s);
            lists[P].push_back({s, p});
        }
    c->pairs.clear();
    c->psmean.clear();
    c->psq.clear();
    c->bp_off.assign(nbp, 0);
    c->bp_cnt.assign(nbp, 0);
    for (int P = 0; P < nbp; ++P) {
        auto& v = lists[P];
        std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        c->bp_off[P] = (int)c->pairs.size();
        for (auto& e : v) {
            if (e.first * smax < c->tol) continue;
            PairDev pd = e.second;
            const int ma = (c->L[pd.sa] + 1) * (c->L[pd.sa] + 2) / 2, mb = (c->L[pd.sb] + 1) * (c->L[pd.sb] + 2) / 2;
            pd.sq_off = (int)c->psq.size();
            pd.sq_min = 1e300;
            pd.sq_max = 0.0;
            for (int x = 0; x < ma; ++x)
                for (int y = 0; y < mb; ++y) {
                    const double q = std::sqrt(
                        std::fabs(c->screen[(size_t)(c->cart_off[pd.sa] + x) * c->ncart + c->cart_off[pd.sb] + y]));
                    c->psq.push_back(q);
                    pd.sq_min = std::min(pd.sq_min, q);
                    pd.sq_max = std::max(pd.sq_max, q);
                }
            c->pairs.push_back(pd);
            c->psmean.push_back(e.first);
        }
        c->bp_cnt[P] = (int)c->pairs.size() - c->bp_off[P];
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
    struct Row {
        int cnt;
        double cost;
    };
    c->classes.clear();
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
    auto bpX = [&](int P) {
        int X = 0;
        while ((X + 1) * (X + 2) / 2 <= P) ++X;
        return X;
    };
    for (int P = 0; P < nbp; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            if (!c->bp_cnt[P] || !c->bp_cnt[Q]) continue;
            const int A = bpX(P), B = P - A * (A + 1) / 2, C = bpX(Q), D = Q - C * (C + 1) / 2;
            const FlopParts fp = model_flops(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D], c->pure, false);
            const int ci = class_index(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D]);
            if ((int)class_cost.size() <= ci) class_cost.resize(ci + 1, 0.0);
            for (int i = 0; i < c->bp_cnt[P]; ++i) {
                int cnt = ket_count_host(&c->psmean[c->bp_off[Q]], c->bp_cnt[Q], c->psmean[c->bp_off[P] + i], c->tol);
                if (P == Q) cnt = std::min(cnt, i + 1);
                if (!cnt) continue;
                const double cost =
                    c->pairs[c->bp_off[P] + i].np * cum_np[cum_off[Q] + cnt] * fp.per_prim + cnt * fp.per_quartet;
                total_cost += cost;
                class_cost[ci] += cost;
            }
        }
    const double target = std::max(total_cost / (256.0 * 6.0), 1.0);
    c->slab_doubles = 0;
    c->total_quartets = 0;
    c->total_flops = 0;
    c->total_grad_flops = 0;
    for (int P = 0; P < nbp; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            if (!c->bp_cnt[P] || !c->bp_cnt[Q]) continue;
            const int A = bpX(P), B = P - A * (A + 1) / 2, C = bpX(Q), D = Q - C * (C + 1) / 2;
            const int ci = class_index(c->blk_L[A], c->blk_L[B], c->blk_L[C], c->blk_L[D]);
            ClassWork& cw = c->classes[ci];
            const FlopParts fp = model_flops(cw.la, cw.lb, cw.lc, cw.ld, c->pure, false);
            const FlopParts gp = model_flops(cw.la, cw.lb, cw.lc, cw.ld, c->pure, true);
            const int tile = c->blk_nf[A] * c->blk_nf[B] + c->blk_nf[C] * c->blk_nf[D] + c->blk_nf[A] * c->blk_nf[C] +
                             c->blk_nf[A] * c->blk_nf[D] + c->blk_nf[B] * c->blk_nf[C] + c->blk_nf[B] * c->blk_nf[D];
            // per-class item target: a class launch should have >= 256 items when it has the work
            const double tgt = std::max(std::min(target, class_cost[ci] / 256.0), 1.0);
            WorkItem w{};
            w.P = P;
            w.Q = Q;
            w.c0 = 0;
            w.c1 = 0x7fffffff;
            w.A = A;
            w.B = B;
            w.C = C;
            w.D = D;
            double acc = 0;
            long long q_acc = 0;
            int r0 = -1;
            auto flush = [&](int r1) {
                if (r0 < 0 || q_acc == 0) {
                    r0 = -1;
                    acc = 0;
                    q_acc = 0;
                    return;
                }
                w.r0 = r0;
                w.r1 = r1;
                w.c0 = 0;
                w.c1 = 0x7fffffff;
                w.slab = c->slab_doubles;
                c->slab_doubles += tile;
                cw.items.push_back(w);
                r0 = -1;
                acc = 0;
                q_acc = 0;
            };
            for (int i = 0; i < c->bp_cnt[P]; ++i) {
                int cnt = ket_count_host(&c->psmean[c->bp_off[Q]], c->bp_cnt[Q], c->psmean[c->bp_off[P] + i], c->tol);
                if (P == Q) cnt = std::min(cnt, i + 1);
                const double np_i = c->pairs[c->bp_off[P] + i].np;
                const double kq = np_i * cum_np[cum_off[Q] + cnt];
                cw.quartets += cnt;
                cw.flops += kq * fp.per_prim + cnt * fp.per_quartet;
                cw.grad_flops += kq * gp.per_prim + cnt * gp.per_quartet;
                const double row_cost = kq * fp.per_prim + cnt * fp.per_quartet;
                if (row_cost > tgt && cnt > 64) {
                    // an expensive row becomes several items over ket column chunks
                    flush(i);
                    const int chunks = std::min((int)std::ceil(row_cost / tgt), (cnt + 31) / 32);
                    const int width = (cnt + chunks - 1) / chunks;
                    for (int j0 = 0; j0 < cnt; j0 += width) {
                        WorkItem v = w;
                        v.r0 = i;
                        v.r1 = i + 1;
                        v.c0 = j0;
                        v.c1 = std::min(cnt, j0 + width);
                        v.slab = c->slab_doubles;
                        c->slab_doubles += tile;
                        cw.items.push_back(v);
                    }
                    continue;
                }
                if (r0 < 0) r0 = i;
                acc += row_cost;
                q_acc += cnt;
                if (acc >= tgt || i + 1 - r0 >= kMaxRows) flush(i + 1);
            }
            flush(c->bp_cnt[P]);
        }
    // heaviest items first inside each class launch
    for (auto& cw : c->classes) {
        c->total_quartets += cw.quartets;
        c->total_flops += cw.flops;
        c->total_grad_flops += cw.grad_flops;
    }
    c->classes.erase(std::remove_if(c->classes.begin(), c->classes.end(), [](const ClassWork& w) { return w.items.empty(); }),
                     c->classes.end());
    // classes to side streams, heaviest first onto the least loaded stream
    {
        std::vector<size_t> ord(c->classes.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return c->classes[x].flops > c->classes[y].flops; });
        std::vector<ClassWork> sorted;
        for (size_t k : ord) sorted.push_back(c->classes[k]);
        c->classes.swap(sorted);
        double load[kSideStreams] = {0};
        c->class_stream.assign(c->classes.size(), 0);
        for (size_t k = 0; k < c->classes.size(); ++k) {
            int best = 0;
            for (int s = 1; s < kSideStreams; ++s)
                if (load[s] < load[best]) best = s;
            c->class_stream[k] = best;
            load[best] += c->classes[k].flops;
        }
    }
    // gather lists
    const int ntile = nbp;
    std::vector<std::vector<GatherEntry>> lj(ntile), lk(ntile);
    std::vector<WorkItem> all;
    c->total_items = 0;
    for (auto& cw : c->classes) {
        cw.dev_off = all.size();
        for (auto& w : cw.items) {
            all.push_back(w);
            const int f[4] = {c->blk_nf[w.A], c->blk_nf[w.B], c->blk_nf[w.C], c->blk_nf[w.D]};
            const int blk[4] = {w.A, w.B, w.C, w.D};
            // (row block, col block) of Jab Jcd Kac Kad Kbc Kbd
            const int rr[6] = {0, 2, 0, 0, 1, 1}, cc[6] = {1, 3, 2, 3, 2, 3};
            long long off = w.slab;
            for (int r = 0; r < 6; ++r) {
                const int R = blk[rr[r]], C = blk[cc[r]];
                auto& L = (r < 2) ? lj : lk;
                if (R > C) {
                    L[R * (R + 1) / 2 + C].push_back({off, f[cc[r]], 0});
                } else if (R < C) {
                    L[C * (C + 1) / 2 + R].push_back({off, f[cc[r]], 1});
                } else {
                    L[R * (R + 1) / 2 + R].push_back({off, f[cc[r]], 0});
                    L[R * (R + 1) / 2 + R].push_back({off, f[cc[r]], 1});
                }
                off += (long long)f[rr[r]] * f[cc[r]];
            }
        }
        c->total_items += (long long)cw.items.size();
    }
    c->tiles.clear();
    c->gJ_off.assign(1, 0);
    c->gK_off.assign(1, 0);
    c->gJ.clear();
    c->gK.clear();
    for (int X = 0; X < nb; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int t = X * (X + 1) / 2 + Y;
            c->tiles.push_back(make_int2(X, Y));
            c->gJ.insert(c->gJ.end(), lj[t].begin(), lj[t].end());
            c->gK.insert(c->gK.end(), lk[t].begin(), lk[t].end());
            c->gJ_off.push_back((int)c->gJ.size());
            c->gK_off.push_back((int)c->gK.size());
        }
    // upload
    std::vector<int> shell_atom(c->atom);
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
    HFG_CHECK(c, c->d_shell_atom.upload(shell_atom, c->stream));
    HFG_CHECK(c, c->d_slab.ensure((size_t)std::max<long long>(c->slab_doubles, 1)));
    const size_t nn = (size_t)c->nbf * c->nbf;
    HFG_CHECK(c, c->d_D.ensure(nn));
    HFG_CHECK(c, c->d_J.ensure(nn));
    HFG_CHECK(c, c->d_K.ensure(nn));
    HFG_CHECK(c, c->d_F.ensure(nn));
    HFG_CHECK(c, c->d_H.ensure(nn));
    HFG_CHECK(c, c->d_G.ensure(nn));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_prev = false;
    return HFG_OK;
}

// Largest |D| of every shell block (nshell x nshell), uploaded for the density screen.
int upload_dmax(hfg_ctx* c, const double* D) {
    const int ns = c->nshell, n = c->nbf;
    c->dmax.assign((size_t)ns * ns, 0.0f);
    for (int a = 0; a < ns; ++a) {
        const int oa = c->pure ? c->pure_off[a] : c->cart_off[a], na = nfun(c->L[a], c->pure);
        for (int b = 0; b < ns; ++b) {
            const int ob = c->pure ? c->pure_off[b] : c->cart_off[b], nb = nfun(c->L[b], c->pure);
            double m = 0.0;
            for (int p = 0; p < na; ++p)
                for (int q = 0; q < nb; ++q) m = std::max(m, std::fabs(D[(size_t)(oa + p) * n + ob + q]));
            // rounded up so the single-precision table never understates the bound
            c->dmax[(size_t)a * ns + b] = std::nextafter((float)m, INFINITY);
        }
    }
    HFG_CHECK(c, c->d_dmax.upload(c->dmax, c->stream));
    return HFG_OK;
}

// J and K of the density already in d_D, into d_J / d_K. dscreen: apply the density-weighted
// screen with the table from upload_dmax.
int run_jk(hfg_ctx* c, bool dscreen = false) {
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
    a.n = c->nbf;
    a.tol = c->tol;
    a.psq = c->d_psq.p;
    a.elem_tol = c->tol;
    a.prim_cut = c->opt.integral_tol / 1e4;
    a.slab = c->d_slab.p;
    a.dmax = c->d_dmax.p;
    a.nshell = c->nshell;
    a.dtol = (dscreen && c->d_dmax.p) ? c->opt.density_tol : 0.0;
    const int mixed = c->opt.precision == HFG_MIXED ? 1 : 0;
    HFG_CHECK(c, hipEventRecord(c->ev0, c->stream));
    HFG_CHECK(c, for_each_class_concurrent(c, [&](const ClassWork& cw, hipStream_t st) {
        JkArgs b = a;
        b.items = c->d_items.p + cw.dev_off;
        return launch_jk(cw.la, cw.lb, cw.lc, cw.ld, c->pure, mixed, b, (int)cw.items.size(), st);
    }));
    HFG_CHECK(c, hipEventRecord(c->ev1, c->stream));
    GatherArgs g{};
    g.tiles = c->d_tiles.p;
    g.gJ_off = c->d_gJ_off.p;
    g.gJ = c->d_gJ.p;
    g.gK_off = c->d_gK_off.p;
    g.gK = c->d_gK.p;
    g.slab = c->d_slab.p;
    g.blk_nf = c->d_blk_nf.p;
    g.blk_func_off = c->d_blk_func_off.p;
    g.blk_func_ao = c->d_blk_func_ao.p;
    g.n = c->nbf;
    g.J = c->d_J.p;
    g.K = c->d_K.p;
    HFG_CHECK(c, launch_gather(g, (int)c->tiles.size(), c->stream));
    return HFG_OK;
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

}  // namespace

extern "C" {

void hfg_default_options(hfg_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->precision = HFG_FP64;
    opt->integral_tol = 1e-14;
    opt->schwarz = 1;
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return HFG_ERR_HIP;
    }
    for (int k = 0; k < kSideStreams; ++k)
        if (hipStreamCreate(&c->side[k]) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) != hipSuccess) {
            hfg_destroy(c);
            return HFG_ERR_HIP;
        }
    std::vector<double> tab = boys_table();
    if (c->d_boys.upload(tab, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        delete c;
        return HFG_ERR_HIP;
    }
    *out = c;
    return HFG_OK;
}

int hfg_set_basis(hfg_ctx* c, int nshell, const int* L, const int* atom_of_shell, const int* nprim, const double* alpha,
                  const double* coef, int pure, const int* cart_offset, const int* pure_offset) {
    if (!c || nshell <= 0 || !L || !atom_of_shell || !nprim || !alpha || !coef || !cart_offset || !pure_offset)
        return fail(c, HFG_ERR_ARG, "hfg_set_basis: bad arguments");
    c->nshell = nshell;
    c->L.assign(L, L + nshell);
    c->atom.assign(atom_of_shell, atom_of_shell + nshell);
    c->nprim.assign(nprim, nprim + nshell);
    c->cart_off.assign(cart_offset, cart_offset + nshell);
    c->pure_off.assign(pure_offset, pure_offset + nshell);
    c->pure = pure != 0;
    c->prim_off.assign(nshell, 0);
    int tot = 0;
    int ncart = 0, nsph = 0;
    for (int s = 0; s < nshell; ++s) {
        if (L[s] < 0 || L[s] > 2) return fail(c, HFG_ERR_UNSUPPORTED, "hfg_set_basis: only s, p and d shells are supported");
        if (nprim[s] <= 0) return fail(c, HFG_ERR_ARG, "hfg_set_basis: shell without primitives");
        c->prim_off[s] = tot;
        tot += nprim[s];
        ncart = std::max(ncart, cart_offset[s] + (L[s] + 1) * (L[s] + 2) / 2);
        nsph = std::max(nsph, pure_offset[s] + 2 * L[s] + 1);
    }
    c->alpha.assign(alpha, alpha + tot);
    c->coef.assign(coef, coef + tot);
    c->ncart = ncart;
    c->nbf = c->pure ? nsph : ncart;
    c->have_basis = true;
    c->have_geom = false;
    return HFG_OK;
}

int hfg_set_geometry(hfg_ctx* c, int natom, const double* xyz_bohr, const int* Z) {
    (void)Z;
    if (!c || !c->have_basis || natom <= 0 || !xyz_bohr) return fail(c, HFG_ERR_ARG, "hfg_set_geometry: bad arguments");
    for (int s = 0; s < c->nshell; ++s)
        if (c->atom[s] < 0 || c->atom[s] >= natom) return fail(c, HFG_ERR_ARG, "hfg_set_geometry: shell atom out of range");
    HFG_CHECK(c, hipSetDevice(c->device));
    c->natom = natom;
    c->xyz.assign(xyz_bohr, xyz_bohr + 3 * natom);
    int r = setup_geometry(c);
    if (r == HFG_OK) c->have_geom = true;
    return r;
}

int hfg_build_jk(hfg_ctx* c, const double* D, double* J, double* K, int flags) {
    (void)flags;
    int r = check_ready(c);
    if (r) return r;
    if (!D) return fail(c, HFG_ERR_ARG, "hfg_build_jk: D is NULL");
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)c->nbf * c->nbf;
    HFG_CHECK(c, hipMemcpyAsync(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((r = run_jk(c))) return r;
    if (J) HFG_CHECK(c, hipMemcpyAsync(J, c->d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (K) HFG_CHECK(c, hipMemcpyAsync(K, c->d_K.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return finish_stats(c, t0);
}

int hfg_fock(hfg_ctx* c, const double* H, const double* D, int incremental, double* F) {
    int r = check_ready(c);
    if (r) return r;
    if (!H || !D || !F) return fail(c, HFG_ERR_ARG, "hfg_fock: NULL argument");
    auto t0 = std::chrono::steady_clock::now();
    const size_t nn = (size_t)c->nbf * c->nbf;
    const bool inc = incremental && c->have_prev;
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

int hfg_energy(hfg_ctx* c, const double* D, double* E2) {
    int r = check_ready(c);
    if (r) return r;
    if (!D || !E2) return fail(c, HFG_ERR_ARG, "hfg_energy: NULL argument");
    const size_t nn = (size_t)c->nbf * c->nbf;
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

int hfg_schwarz(hfg_ctx* c, double* screen) {
    int r = check_ready(c);
    if (r) return r;
    if (!screen) return fail(c, HFG_ERR_ARG, "hfg_schwarz: NULL argument");
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
    c->d_pp0.release();
    c->d_pp1.release();
    c->d_pp2.release();
    c->d_pairs.release();
    c->d_bp_off.release();
    c->d_bp_cnt.release();
    c->d_blk_nf.release();
    c->d_blk_func_off.release();
    c->d_blk_func_ao.release();
    c->d_gJ_off.release();
    c->d_gK_off.release();
    c->d_shell_atom.release();
    c->d_items.release();
    c->d_tiles.release();
    c->d_gJ.release();
    c->d_gK.release();
    for (int k = 0; k < kSideStreams; ++k) {
        if (c->side[k]) (void)hipStreamSynchronize(c->side[k]);
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
