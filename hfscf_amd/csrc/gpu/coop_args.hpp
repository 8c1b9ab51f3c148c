// Program of one class for the wavefront-cooperative J/K kernel (coop.hpp): index lists of its
// recursions in device memory and the LDS layout of its launches.
#pragma once
#include <hip/hip_runtime.h>

namespace hfg {

constexpr int kCoopMaxLevels = 16;
constexpr int kCoopMaxAcc = 16;   // contracted values per lane: ceil(NE * NF / 64), (dd|dd) 961 -> 16
constexpr int kCoopBlock = 64;    // one wavefront per workgroup
constexpr int kCoopStaticLds = 2048;  // static LDS of the kernel (row prefix sums, queue slot), bounded

// Program of one class (device pointers into one buffer) and the LDS layout of its kernel.
struct CoopArgs {
    int la, lb, lc, ld, LT;
    int NA, NB, NC, ND;          // functions per shell in the output basis
    int na, nb, nc, nd;          // Cartesian functions per shell
    int NF;                      // contracted ket range (functions c with lc <= |c| <= lc + ld)
    int nacc;                    // contracted [e0|f0] values, NE * NF
    int pure;                    // spherical transform
    // VRR: ops of level l at [vlev[l], vlev[l+1]); an op is
    //   x = target | d << 16 | ket << 18 | n1 << 19 | ne << 23, y = source (m), z = source of the
    //   n1 term or -1, w = source of the ne term or -1 (ket ops); sources at m and m + 1
    const int4* vops;
    int vlev[kCoopMaxLevels + 1];
    int nvlev;
    const int* acc_src;          // table entry of contracted value k = (e - e0) * NF + (f - f0)
    // bra HRR, one table of hbsz values per ket function f: init from the contracted values
    // (hb_init[e - e0] = entry of (e, 0)), ops x = target | d << 16, y = (a + 1_d, b - 1_d),
    // z = (a, b - 1_d), output hb_out = (a * nb + b, entry)
    const int4* hbops;
    int hblev[kCoopMaxLevels + 1];
    int nhblev, hbsz;
    const int* hb_init;
    const int2* hb_out;
    int nhb_out;
    // ket HRR, one table of hksz values per bra function pair, the same forms with c, d
    const int4* hkops;
    int hklev[kCoopMaxLevels + 1];
    int nhklev, hksz;
    int hk_chunk;                // bra function pairs whose ket tables fit region H at once
    const int* hk_init;
    const int2* hk_out;
    int nhk_out;
    // LDS layout in R units after the tiles: region H (VRR table, then the HRR tables, then the
    // transform temporaries), the bra HRR output, the Cartesian block, the density blocks
    int off_bra, off_I, off_D, lds_r;
};

constexpr int kCoopGradBlock = 256;  // four wavefronts share a quartet of the gradient kernel
constexpr int kCoopGradStaticLds = 6144;  // static LDS of the gradient kernel (atom sums, rows), bounded

// One shifted class of a gradient program.
struct CoopShift {
    int pos, shift;              // shifted index (0 a, 1 b, 2 c) and +1 / -1
    int na, nb, nc, nd;          // Cartesian functions of the shifted class
    int NE, NF;                  // bra range (functions e with sa <= |e| <= sa + sb), ket range
    const int4* hbops;           // bra HRR program (the form of CoopArgs)
    int hblev[kCoopMaxLevels + 1];
    int nhblev, hbsz, hb_chunk;  // hb_chunk ket functions at a time
    const int* hb_init;          // entry of (e, 0) per bra function e
    const int* hb_src;           // VRR entry of [e0|f0]^(0), index e * NF + f
    const int2* hb_out;
    int nhb_out;
    const int4* hkops;
    int hklev[kCoopMaxLevels + 1];
    int nhklev, hksz, hk_chunk;
    const int* hk_init;
    const int2* hk_out;          // (c * nd + d, entry)
    int nhk_out;
    const int* cg;               // per shifted integral and direction: G entry | factor << 20, or -1
};

struct CoopGradArgs {
    int la, lb, lc, ld, LT;      // unshifted class; the VRR runs at LT + 1
    int NA, NB, NC, ND, na, nb, nc, nd;
    int pure;
    const int4* vops;
    int vlev[kCoopMaxLevels + 1];
    int nvlev;
    int nshift;
    CoopShift sh[6];
    // LDS regions in R units: G (Cartesian density), V (VRR table; first the back-transform
    // temporaries, 2 x the Cartesian block), H (bra HRR tables, then ket HRR tables), bra output
    int off_V, off_Hb, off_bra, off_Hk, lds_r;
};

}  // namespace hfg
