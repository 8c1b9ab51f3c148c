// Schwarz diagonal, gather, counter reset and Fock kernels.
#include "jk_impl.hpp"

namespace hfg {

// ------------------------------------------------------------------ Schwarz diagonal (ab|ab)
// hfscf keeps the Cartesian diagonal S_pq = (pq|pq) per shell pair (hfscf_oseri_screen.cpp:22-181).
template <int LA, int LB>
__global__ __launch_bounds__(64) void screen_kernel(ScreenArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.count) return;
    constexpr int na = ncart(LA), nb = ncart(LB);
    const PairDev p = a.pairs[t];
    double blk[na * nb * na * nb];
    // hfscf's screen matrix skips primitive quartets with |prefactor| < 1e-16
    // (hfscf_oseri_screen.cpp:78); diagonals of distant pairs can vanish entirely, which then
    // drops their quartets in the shell and element screens
    eri_cart<LA, LB, LA, LB, double>(p, p, a.pp0, a.pp1, a.pp2, a.boys, 1e-16, blk);
    double* out = a.out + a.out_off[t];
    for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) out[x * nb + y] = blk[((x * nb + y) * na + x) * nb + y];
}

hipError_t launch_screen(int la, int lb, const ScreenArgs& a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    const dim3 grid((a.count + 63) / 64), block(64);
#define HFG_SCREEN(A, B)                                                   \
    if (la == A && lb == B) {                                              \
        hipLaunchKernelGGL((screen_kernel<A, B>), grid, block, 0, s, a);   \
        return hipGetLastError();                                          \
    }
    HFG_SCREEN(0, 0)
    HFG_SCREEN(1, 0)
    HFG_SCREEN(1, 1)
    HFG_SCREEN(2, 0)
    HFG_SCREEN(2, 1)
    HFG_SCREEN(2, 2)
#undef HFG_SCREEN
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ gather
// Sum of element (u, v) over the work-item tiles [g0, g1) of a gather list, in a fixed order:
// four interleaved partial sums (entries g0 + 4i + q), so four independent loads are in flight.
__device__ __forceinline__ double gather_sum(const GatherEntry* __restrict__ list, int g0, int g1,
                                             const double* __restrict__ slab, int u, int v) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int g = g0;
    for (; g + 3 < g1; g += 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const GatherEntry en = list[g + q];
            s[q] += en.trans ? slab[en.off + (long long)v * en.cols + u] : slab[en.off + (long long)u * en.cols + v];
        }
    }
    for (; g < g1; ++g) {
        const GatherEntry en = list[g];
        s[0] += en.trans ? slab[en.off + (long long)v * en.cols + u] : slab[en.off + (long long)u * en.cols + v];
    }
    return (s[0] + s[1]) + (s[2] + s[3]);
}

// One workgroup per output tile (X >= Y): sums every work-item tile that targets it
// (transposed where the work item held it as (Y, X)) in a fixed order, then writes the
// symmetric J and K elements.
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs a) {
    const int4 xy = a.tiles[blockIdx.x];
    const MolDev md = a.mols[xy.z];
    double* __restrict__ J = a.J + md.moff;
    double* __restrict__ K = a.K + md.moff;
    const int fX = a.blk_nf[xy.x], fY = a.blk_nf[xy.y];
    const int* aoX = a.blk_func_ao + a.blk_func_off[xy.x];
    const int* aoY = a.blk_func_ao + a.blk_func_off[xy.y];
    const int j0 = a.gJ_off[blockIdx.x], j1 = a.gJ_off[blockIdx.x + 1];
    const int k0 = a.gK_off[blockIdx.x], k1 = a.gK_off[blockIdx.x + 1];
    for (int e = threadIdx.x; e < fX * fY; e += blockDim.x) {
        const int u = e / fY, v = e - u * fY;
        if (xy.x == xy.y && u < v) continue;  // diagonal tile: one thread per symmetric element pair
        const double sj = gather_sum(a.gJ, j0, j1, a.slab, u, v);
        const double sk = gather_sum(a.gK, k0, k1, a.slab, u, v);
        const size_t i1 = (size_t)aoX[u] * md.n + aoY[v];
        const size_t i2 = (size_t)aoY[v] * md.n + aoX[u];
        J[i1] = sj;
        J[i2] = sj;
        K[i1] = sk;
        K[i2] = sk;
    }
}

hipError_t launch_gather(const GatherArgs& a, int ntiles, hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(ntiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void zero_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Fock
// G = 2J - K (+ Facc when accumulating); F = base + G; Facc keeps G for incremental builds.
__global__ void fock_kernel(const double* base, const double* J, const double* K, double* F, double* Facc, int nn,
                            int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double g = 2.0 * J[i] - K[i];
    if (accumulate) g += Facc[i];
    Facc[i] = g;
    F[i] = base[i] + g;
}

hipError_t launch_fock(const double* base, const double* J, const double* K, double* F, double* Facc, int nn,
                       int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(fock_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, base, J, K, F, Facc, nn, accumulate);
    return hipGetLastError();
}

}  // namespace hfg
