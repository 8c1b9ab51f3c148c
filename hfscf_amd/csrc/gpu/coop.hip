// Host side of the wavefront-cooperative J/K kernel: the per-class programs (VRR, HRR and
// contraction index lists in the layouts the kernel reads) and the launcher. The recursions are
// those of Vrr::run and Hrr::run (hgp.hpp), written out as lists instead of unrolled code.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "coop.hpp"
#include "coop_grad.hpp"

namespace hfg {

namespace {

// Compact VRR table over (e, f, m): |e| <= lab, |f| <= lcd, m <= lt - |e| - |f|, ordered by e,
// then f, then m; entry (0, 0, m) is m.
struct VrrIndex {
    int lab, lcd, lt;
    std::vector<int> eoff;  // per e: first entry of (e, 0, 0)
    int size = 0;
    VrrIndex(int lab_, int lcd_, int lt_) : lab(lab_), lcd(lcd_), lt(lt_) {
        for (int e = 0; e < ncum(lab); ++e) {
            eoff.push_back(size);
            for (int f = 0; f < ncum(lcd); ++f) size += std::max(0, lt - glvl(e) - glvl(f) + 1);
        }
    }
    int off(int e, int f) const {
        int o = eoff[e];
        for (int g = 0; g < f; ++g) o += std::max(0, lt - glvl(e) - glvl(g) + 1);
        return o;
    }
};

// Compact HRR table of one side: entries (a, b) for |b| <= lb and la <= |a| <= la + lb - |b|.
struct HrrIndex {
    int la, lb, a0;
    std::vector<int> boff;  // per b: first entry, a indexed from a0
    int size = 0;
    HrrIndex(int la_, int lb_) : la(la_), lb(lb_), a0(ncum(la_ - 1)) {
        for (int b = 0; b < ncum(lb); ++b) {
            boff.push_back(size);
            size += ncum(la + lb - glvl(b)) - a0;
        }
    }
    int off(int a, int b) const { return boff[b] + a - a0; }
};

// VRR program over the table V: bra targets (e, 0, m) by level |e|, then ket targets (e, f, m)
// by level |f| (the targets of a level depend only on lower levels).
void vrr_program(const VrrIndex& V, std::vector<int4>& vops, std::vector<int>& vlev) {
    if (V.size > 0xffff) throw std::runtime_error("cooperative program: VRR table too large");
    const int lab = V.lab, lcd = V.lcd, lt = V.lt;
    vops.clear();
    vlev.assign(1, 0);
    for (int le = 1; le <= lab; ++le) {
        for (int e = ncum(le - 1); e < ncum(le); ++e) {
            const int d = gdir(e), e1 = gminus(e, d), n1 = gexp(e1, d);
            for (int m = 0; m <= lt - le; ++m) {
                const int z = n1 > 0 ? V.off(gminus(e1, d), 0) + m : -1;
                vops.push_back(int4{(V.off(e, 0) + m) | (d << 16) | (n1 << 19), V.off(e1, 0) + m, z, -1});
            }
        }
        vlev.push_back((int)vops.size());
    }
    for (int lf = 1; lf <= lcd; ++lf) {
        for (int f = ncum(lf - 1); f < ncum(lf); ++f) {
            const int d = gdir(f), f1 = gminus(f, d), n1 = gexp(f1, d);
            for (int e = 0; e < ncum(lab); ++e) {
                const int le = glvl(e);
                if (le + lf > lt) continue;
                const int ne = gexp(e, d);
                for (int m = 0; m <= lt - le - lf; ++m) {
                    const int z = n1 > 0 ? V.off(e, gminus(f1, d)) + m : -1;
                    const int wv = ne > 0 ? V.off(gminus(e, d), f1) + m : -1;
                    vops.push_back(
                        int4{(V.off(e, f) + m) | (d << 16) | (1 << 18) | (n1 << 19) | (ne << 23), V.off(e, f1) + m, z, wv});
                }
            }
        }
        vlev.push_back((int)vops.size());
    }
}

// Bra (or ket) HRR program: ops by level |b| = 1..lb, init entries of (a, 0), outputs.
void hrr_program(int la, int lb, std::vector<int4>& ops, std::vector<int>& lev, std::vector<int>& init,
                 std::vector<int2>& out, int& size) {
    const HrrIndex h(la, lb);
    size = h.size;
    lev.assign(1, 0);
    for (int k = 1; k <= lb; ++k) {
        for (int b = ncum(k - 1); b < ncum(k); ++b) {
            const int d = gdir(b), bm = gminus(b, d);
            for (int a = h.a0; a < ncum(la + lb - k); ++a)
                ops.push_back(int4{h.off(a, b) | (d << 16), h.off(gplus(a, d), bm), h.off(a, bm), 0});
        }
        lev.push_back((int)ops.size());
    }
    init.clear();
    for (int a = h.a0; a < ncum(la + lb); ++a) init.push_back(h.off(a, 0));
    out.clear();
    const int na = ncart(la), nb = ncart(lb);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) out.push_back(int2{ia * nb + ib, h.off(h.a0 + ia, ncum(lb - 1) + ib)});
}

template <typename T>
int append(std::vector<int4>& words, const std::vector<T>& v) {
    const int at = (int)words.size();
    const size_t bytes = v.size() * sizeof(T);
    words.resize(at + (bytes + 15) / 16);
    if (bytes) std::memcpy(words.data() + at, v.data(), bytes);
    return at;
}

}  // namespace

CoopHostProgram make_coop_program(int la, int lb, int lc, int ld, bool pure) {
    CoopHostProgram h{};
    CoopArgs& p = h.args;
    p.la = la;
    p.lb = lb;
    p.lc = lc;
    p.ld = ld;
    p.LT = la + lb + lc + ld;
    p.na = ncart(la);
    p.nb = ncart(lb);
    p.nc = ncart(lc);
    p.nd = ncart(ld);
    auto nout = [&](int l) { return pure && l > 0 ? 2 * l + 1 : ncart(l); };
    p.NA = nout(la);
    p.NB = nout(lb);
    p.NC = nout(lc);
    p.ND = nout(ld);
    p.pure = pure && p.LT > 0;
    const int lab = la + lb, lcd = lc + ld, lt = p.LT;
    const int e0 = ncum(la - 1), f0 = ncum(lc - 1);
    const int NE = ncum(lab) - e0, NF = ncum(lcd) - f0;
    p.NF = NF;
    p.nacc = NE * NF;
    if (p.nacc > kCoopMaxAcc * kCoopBlock || lt + 1 > kCoopBlock || lab + lcd > kCoopMaxLevels ||
        lb > kCoopMaxLevels || ld > kCoopMaxLevels)
        throw std::runtime_error("cooperative J/K program: class beyond the kernel's limits");
    // VRR (the recursion of Vrr::run)
    const VrrIndex V(lab, lcd, lt);
    std::vector<int4> vops;
    std::vector<int> vlev;
    vrr_program(V, vops, vlev);
    std::vector<int> acc_src;
    for (int e = e0; e < e0 + NE; ++e)
        for (int f = f0; f < f0 + NF; ++f) acc_src.push_back(V.off(e, f));
    // HRR programs (Hrr::run): bra per ket function, ket per bra function pair
    std::vector<int4> hbops, hkops;
    std::vector<int> hblev, hklev, hb_init, hk_init;
    std::vector<int2> hb_out, hk_out;
    hrr_program(la, lb, hbops, hblev, hb_init, hb_out, p.hbsz);
    hrr_program(lc, ld, hkops, hklev, hk_init, hk_out, p.hksz);
    p.nvlev = (int)vlev.size() - 1;
    p.nhblev = (int)hblev.size() - 1;
    p.nhklev = (int)hklev.size() - 1;
    for (int l = 0; l < (int)vlev.size(); ++l) p.vlev[l] = vlev[l];
    for (int l = 0; l < (int)hblev.size(); ++l) p.hblev[l] = hblev[l];
    for (int l = 0; l < (int)hklev.size(); ++l) p.hklev[l] = hklev[l];
    p.nhb_out = (int)hb_out.size();
    p.nhk_out = (int)hk_out.size();
    // arrays of the program in one buffer; the pointers hold word offsets until rebased
    auto at = [](int w) { return reinterpret_cast<const void*>((size_t)w); };
    p.vops = reinterpret_cast<const int4*>(at(append(h.words, vops)));
    p.acc_src = reinterpret_cast<const int*>(at(append(h.words, acc_src)));
    p.hbops = reinterpret_cast<const int4*>(at(append(h.words, hbops)));
    p.hb_init = reinterpret_cast<const int*>(at(append(h.words, hb_init)));
    p.hb_out = reinterpret_cast<const int2*>(at(append(h.words, hb_out)));
    p.hkops = reinterpret_cast<const int4*>(at(append(h.words, hkops)));
    p.hk_init = reinterpret_cast<const int*>(at(append(h.words, hk_init)));
    p.hk_out = reinterpret_cast<const int2*>(at(append(h.words, hk_out)));
    // LDS regions (R units): H = max(VRR table, bra HRR tables, ket HRR tables, transform
    // temporaries), then the bra HRR output, the Cartesian block and the density blocks
    const int nab = p.na * p.nb;
    int hsize = std::max(V.size, NF * p.hbsz);
    hsize = std::max(hsize, p.NA * p.nb * p.nc * p.nd + p.NA * p.NB * p.nc * p.nd);
    // the ket HRR tables run in chunks of bra function pairs that fit that region (at least one)
    p.hk_chunk = std::max(1, std::min(nab, hsize / p.hksz));
    hsize = std::max(hsize, p.hk_chunk * p.hksz);
    p.off_bra = hsize;
    p.off_I = p.off_bra + nab * NF;
    p.off_D = p.off_I + nab * p.nc * p.nd;
    p.lds_r = p.off_D + p.NA * p.NB + p.NC * p.ND + p.NA * p.NC + p.NA * p.ND + p.NB * p.NC + p.NB * p.ND;
    return h;
}

size_t CoopHostProgram::lds_bytes(bool fp32) const {
    return sizeof(double) * kRegions * kBlockFuncs * kBlockFuncs + (fp32 ? sizeof(float) : sizeof(double)) * args.lds_r;
}

CoopArgs coop_args_on_device(const CoopHostProgram& h, const int4* dev) {
    CoopArgs p = h.args;
    auto reb = [&](const void* off) { return reinterpret_cast<const void*>(dev + (size_t)off); };
    p.vops = reinterpret_cast<const int4*>(reb(p.vops));
    p.acc_src = reinterpret_cast<const int*>(reb(p.acc_src));
    p.hbops = reinterpret_cast<const int4*>(reb(p.hbops));
    p.hb_init = reinterpret_cast<const int*>(reb(p.hb_init));
    p.hb_out = reinterpret_cast<const int2*>(reb(p.hb_out));
    p.hkops = reinterpret_cast<const int4*>(reb(p.hkops));
    p.hk_init = reinterpret_cast<const int*>(reb(p.hk_init));
    p.hk_out = reinterpret_cast<const int2*>(reb(p.hk_out));
    return p;
}

hipError_t launch_jk_coop(const JkArgs& a, const CoopArgs& p, size_t lds_bytes, int mixed, int nitems,
                          LaunchShape sh, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    JkArgs b = a;
    b.nitems = nitems;
    const int grid = sh.grid > 0 ? std::min(sh.grid, nitems) : nitems;
    if (mixed) {
        auto k = jk_coop_kernel<float, double>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kCoopBlock), lds_bytes, s, b, p);
    } else {
        auto k = jk_coop_kernel<double, double>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kCoopBlock), lds_bytes, s, b, p);
    }
    return hipGetLastError();
}

hipError_t coop_set_lds_limit(size_t jk_bytes, size_t grad_bytes) {
    // the dynamic LDS the launches may request, set once (not while a stream is being captured)
    hipError_t e = hipSuccess;
    auto set = [&](const void* k, size_t b) {
        if (e == hipSuccess && b > 0) e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
    };
    set(reinterpret_cast<const void*>(jk_coop_kernel<double, double>), jk_bytes);
    set(reinterpret_cast<const void*>(jk_coop_kernel<float, double>), jk_bytes);
    set(reinterpret_cast<const void*>(grad_coop_kernel<double>), grad_bytes);
    set(reinterpret_cast<const void*>(grad_coop_kernel<float>), grad_bytes);
    return e;
}

// ---------------------------------------------------------------------- gradient programs

CoopGradHostProgram make_coop_grad_program(int la, int lb, int lc, int ld, bool pure, size_t lds_limit) {
    CoopGradHostProgram h{};
    CoopGradArgs& p = h.args;
    const int L4[4] = {la, lb, lc, ld};
    p.la = la;
    p.lb = lb;
    p.lc = lc;
    p.ld = ld;
    p.LT = la + lb + lc + ld;
    p.na = ncart(la);
    p.nb = ncart(lb);
    p.nc = ncart(lc);
    p.nd = ncart(ld);
    auto nout = [&](int l) { return pure && l > 0 ? 2 * l + 1 : ncart(l); };
    p.NA = nout(la);
    p.NB = nout(lb);
    p.NC = nout(lc);
    p.ND = nout(ld);
    p.pure = pure && p.LT > 0;
    if (p.LT + 2 > kCoopGradBlock || la + lb + lc + ld + 2 > kCoopMaxLevels)
        throw std::runtime_error("cooperative gradient program: class beyond the kernel's limits");
    // VRR at L + 1 (|e| <= la + lb + 1, |f| <= lc + ld + 1, |e| + |f| <= L + 1)
    const VrrIndex V(la + lb + 1, lc + ld + 1, p.LT + 1);
    std::vector<int4> vops;
    std::vector<int> vlev;
    vrr_program(V, vops, vlev);
    p.nvlev = (int)vlev.size() - 1;
    for (int l = 0; l < (int)vlev.size(); ++l) p.vlev[l] = vlev[l];
    auto at = [](int w) { return reinterpret_cast<const void*>((size_t)w); };
    p.vops = reinterpret_cast<const int4*>(at(append(h.words, vops)));
    const int nc4 = p.na * p.nb * p.nc * p.nd;
    // shifted classes: a +1, a -1, b +1, b -1, c +1, c -1 (the -1 only for l > 0)
    struct Tmp {
        std::vector<int4> hbops, hkops;
        std::vector<int> hblev, hklev, hb_init, hk_init, hb_src, cg;
        std::vector<int2> hb_out, hk_out;
    };
    std::vector<Tmp> tmp;
    p.nshift = 0;
    int bra_max = 0;
    for (int pos = 0; pos < 3; ++pos)
        for (int sh = 1; sh >= -1; sh -= 2) {
            if (sh < 0 && L4[pos] == 0) continue;
            CoopShift& S = p.sh[p.nshift++];
            Tmp t;
            int Ls[4] = {la, lb, lc, ld};
            Ls[pos] += sh;
            S.pos = pos;
            S.shift = sh;
            S.na = ncart(Ls[0]);
            S.nb = ncart(Ls[1]);
            S.nc = ncart(Ls[2]);
            S.nd = ncart(Ls[3]);
            const int e0 = ncum(Ls[0] - 1), f0 = ncum(Ls[2] - 1);
            S.NE = ncum(Ls[0] + Ls[1]) - e0;
            S.NF = ncum(Ls[2] + Ls[3]) - f0;
            hrr_program(Ls[0], Ls[1], t.hbops, t.hblev, t.hb_init, t.hb_out, S.hbsz);
            hrr_program(Ls[2], Ls[3], t.hkops, t.hklev, t.hk_init, t.hk_out, S.hksz);
            S.nhblev = (int)t.hblev.size() - 1;
            S.nhklev = (int)t.hklev.size() - 1;
            for (int l = 0; l < (int)t.hblev.size(); ++l) S.hblev[l] = t.hblev[l];
            for (int l = 0; l < (int)t.hklev.size(); ++l) S.hklev[l] = t.hklev[l];
            S.nhb_out = (int)t.hb_out.size();
            S.nhk_out = (int)t.hk_out.size();
            for (int e = 0; e < S.NE; ++e)
                for (int f = 0; f < S.NF; ++f) t.hb_src.push_back(V.off(e0 + e, f0 + f));
            // per shifted integral (ia, ib, ic, id) and direction x: the G element of the
            // unshifted class and the factor (+1: G[x - 1_x], -1: (x_x + 1) G[x + 1_x])
            const int n4[4] = {S.na, S.nb, S.nc, S.nd};
            const int u4[4] = {p.na, p.nb, p.nc, p.nd};
            for (int ia = 0; ia < n4[0]; ++ia)
                for (int ib = 0; ib < n4[1]; ++ib)
                    for (int ic = 0; ic < n4[2]; ++ic)
                        for (int id = 0; id < n4[3]; ++id) {
                            int ci[4] = {ia, ib, ic, id};
                            const int g = ncum(Ls[pos] - 1) + ci[pos];
                            for (int x = 0; x < 3; ++x) {
                                const int ex = gexp(g, x);
                                if (sh > 0 && ex == 0) {
                                    t.cg.push_back(-1);
                                    continue;
                                }
                                const int gu = sh > 0 ? gminus(g, x) : gplus(g, x);
                                int cu[4] = {ia, ib, ic, id};
                                cu[pos] = lidx(gu);
                                const int gi = ((cu[0] * u4[1] + cu[1]) * u4[2] + cu[2]) * u4[3] + cu[3];
                                const int fac = sh > 0 ? 1 : ex + 1;
                                t.cg.push_back(gi | (fac << 20));
                            }
                        }
            bra_max = std::max(bra_max, S.na * S.nb * S.NF);
            tmp.push_back(std::move(t));
        }
    // LDS layout (R units): G, V, bra, H; H gets what the limit leaves, the HRR tables run in
    // chunks that fit it
    p.off_V = nc4;
    const int vsz = std::max(V.size, 2 * nc4);
    p.off_bra = p.off_V + vsz;
    p.off_Hb = p.off_bra + bra_max;
    p.off_Hk = p.off_Hb;
    const long long budget = (long long)(lds_limit / sizeof(double)) - p.off_Hb;
    int hmax = 0;
    for (int k = 0; k < p.nshift; ++k) {
        CoopShift& S = p.sh[k];
        if (budget < std::max(S.hbsz, S.hksz))
            throw std::runtime_error("cooperative gradient program: LDS too small for the class");
        S.hb_chunk = (int)std::min<long long>(S.NF, budget / S.hbsz);
        S.hk_chunk = (int)std::min<long long>(S.na * S.nb, budget / S.hksz);
        hmax = std::max(hmax, std::max(S.hb_chunk * S.hbsz, S.hk_chunk * S.hksz));
    }
    p.lds_r = p.off_Hb + hmax;
    for (int k = 0; k < p.nshift; ++k) {
        CoopShift& S = p.sh[k];
        Tmp& t = tmp[k];
        S.hbops = reinterpret_cast<const int4*>(at(append(h.words, t.hbops)));
        S.hb_init = reinterpret_cast<const int*>(at(append(h.words, t.hb_init)));
        S.hb_src = reinterpret_cast<const int*>(at(append(h.words, t.hb_src)));
        S.hb_out = reinterpret_cast<const int2*>(at(append(h.words, t.hb_out)));
        S.hkops = reinterpret_cast<const int4*>(at(append(h.words, t.hkops)));
        S.hk_init = reinterpret_cast<const int*>(at(append(h.words, t.hk_init)));
        S.hk_out = reinterpret_cast<const int2*>(at(append(h.words, t.hk_out)));
        S.cg = reinterpret_cast<const int*>(at(append(h.words, t.cg)));
    }
    return h;
}

CoopGradArgs coop_grad_args_on_device(const CoopGradHostProgram& h, const int4* dev) {
    CoopGradArgs p = h.args;
    auto reb = [&](const void* off) { return reinterpret_cast<const void*>(dev + (size_t)off); };
    p.vops = reinterpret_cast<const int4*>(reb(p.vops));
    for (int k = 0; k < p.nshift; ++k) {
        CoopShift& S = p.sh[k];
        S.hbops = reinterpret_cast<const int4*>(reb(S.hbops));
        S.hb_init = reinterpret_cast<const int*>(reb(S.hb_init));
        S.hb_src = reinterpret_cast<const int*>(reb(S.hb_src));
        S.hb_out = reinterpret_cast<const int2*>(reb(S.hb_out));
        S.hkops = reinterpret_cast<const int4*>(reb(S.hkops));
        S.hk_init = reinterpret_cast<const int*>(reb(S.hk_init));
        S.hk_out = reinterpret_cast<const int2*>(reb(S.hk_out));
        S.cg = reinterpret_cast<const int*>(reb(S.cg));
    }
    return p;
}

hipError_t launch_grad_coop(const GradArgs& a, const CoopGradArgs& p, size_t lds_bytes, int mixed, int nitems,
                            LaunchShape sh, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    GradArgs b = a;
    b.nitems = nitems;
    const int grid = sh.grid > 0 ? std::min(sh.grid, nitems) : nitems;
    if (mixed) {
        auto k = grad_coop_kernel<float>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kCoopGradBlock), lds_bytes, s, b, p);
    } else {
        auto k = grad_coop_kernel<double>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kCoopGradBlock), lds_bytes, s, b, p);
    }
    return hipGetLastError();
}

}  // namespace hfg
