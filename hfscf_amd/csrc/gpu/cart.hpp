// Compile-time Cartesian component algebra for the Obara-Saika / Head-Gordon-Pople kernels.
//
// Components of angular momentum L are ordered as in hfscf
// (for i in 0..L: lx = L-i; for j in 0..i: ly = i-j, lz = j; hfscf: src/integrals/hfscf_shell.cpp:27-40).
// A "global" index g enumerates all components with |e| <= L level by level:
// g = ncum(L-1) + cidx(ly, lz).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace hfg {

#define HFG_INLINE __device__ __forceinline__
#define HFG_HD __host__ __device__ constexpr

HFG_HD int ncart(int L) { return (L + 1) * (L + 2) / 2; }
HFG_HD int ncum(int L) { return L < 0 ? 0 : (L + 1) * (L + 2) * (L + 3) / 6; }
HFG_HD int cidx(int ly, int lz) { return (ly + lz) * (ly + lz + 1) / 2 + lz; }
HFG_HD int gidx(int lx, int ly, int lz) { return ncum(lx + ly + lz - 1) + cidx(ly, lz); }

// level (total angular momentum) of global index g
HFG_HD int glvl(int g) {
    int L = 0;
    while (ncum(L) <= g) ++L;
    return L;
}
// exponent of direction d of global index g
HFG_HD int gexp(int g, int d) {
    const int L = glvl(g);
    const int c = g - ncum(L - 1);
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= c) ++i;
    const int lz = c - i * (i + 1) / 2;
    const int ly = i - lz;
    const int lx = L - i;
    return d == 0 ? lx : (d == 1 ? ly : lz);
}
// recursion direction of a target: its first nonzero Cartesian direction
HFG_HD int gdir(int g) { return gexp(g, 0) > 0 ? 0 : (gexp(g, 1) > 0 ? 1 : 2); }
// global index of g - 1_d (requires gexp(g, d) > 0)
HFG_HD int gminus(int g, int d) {
    return gidx(gexp(g, 0) - (d == 0), gexp(g, 1) - (d == 1), gexp(g, 2) - (d == 2));
}
HFG_HD int gplus(int g, int d) { return gidx(gexp(g, 0) + (d == 0), gexp(g, 1) + (d == 1), gexp(g, 2) + (d == 2)); }
// index of component g of level L within that level
HFG_HD int lidx(int g) { return g - ncum(glvl(g) - 1); }

template <int I>
using ic = std::integral_constant<int, I>;

template <typename F, int... I>
HFG_INLINE void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(ic<I>{}), ...);
}
// Fully unrolled loop whose induction variable is a compile-time constant.
template <int N, typename F>
HFG_INLINE void sfor(F&& f) {
    if constexpr (N > 0) sfor_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace hfg
