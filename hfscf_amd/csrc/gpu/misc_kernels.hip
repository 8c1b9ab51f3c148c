// Schwarz diagonal, gather, counter reset and Fock kernels.
#include "jk_impl.hpp"
#include "misc.hpp"

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

// First pass: one workgroup per chunk, one thread per tile element, the chunk's entries summed
// in list order.
__global__ __launch_bounds__(256) void gather_part_kernel(GatherArgs a, const GatherChunk* __restrict__ chunks,
                                                          double* __restrict__ part) {
    const GatherChunk ch = chunks[blockIdx.x];
    const int4 xy = a.tiles[ch.tile];
    const int fX = a.blk_nf[xy.x], fY = a.blk_nf[xy.y];
    const GatherEntry* list = ch.list ? a.gK : a.gJ;
    for (int e = threadIdx.x; e < fX * fY; e += blockDim.x) {
        const int u = e / fY, v = e - u * fY;
        part[ch.part + e] = gather_sum(list, ch.g0, ch.g1, a.slab, u, v);
    }
}

// Second pass: one workgroup per output tile, the chunk sums of its lists in chunk order.
__global__ __launch_bounds__(256) void gather_final_kernel(GatherArgs a, const GatherChunk* __restrict__ chunks,
                                                           const int* __restrict__ cJ_off,
                                                           const int* __restrict__ cK_off,
                                                           const double* __restrict__ part) {
    const int t = blockIdx.x;
    const int4 xy = a.tiles[t];
    const MolDev md = a.mols[xy.z];
    double* __restrict__ J = a.J + md.moff;
    double* __restrict__ K = a.K + md.moff;
    const int fX = a.blk_nf[xy.x], fY = a.blk_nf[xy.y];
    const int* aoX = a.blk_func_ao + a.blk_func_off[xy.x];
    const int* aoY = a.blk_func_ao + a.blk_func_off[xy.y];
    for (int e = threadIdx.x; e < fX * fY; e += blockDim.x) {
        const int u = e / fY, v = e - u * fY;
        if (xy.x == xy.y && u < v) continue;  // diagonal tile: one thread per symmetric element pair
        double sj = 0.0, sk = 0.0;
        for (int k = cJ_off[t]; k < cJ_off[t + 1]; ++k) sj += part[chunks[k].part + e];
        for (int k = cK_off[t]; k < cK_off[t + 1]; ++k) sk += part[chunks[k].part + e];
        const size_t i1 = (size_t)aoX[u] * md.n + aoY[v];
        const size_t i2 = (size_t)aoY[v] * md.n + aoX[u];
        J[i1] = sj;
        J[i2] = sj;
        K[i1] = sk;
        K[i2] = sk;
    }
}

hipError_t launch_gather_chunks(const GatherArgs& a, const GatherChunk* chunks, const int* cJ_off, const int* cK_off,
                                int nchunks, int ntiles, double* part, hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    if (nchunks > 0) hipLaunchKernelGGL(gather_part_kernel, dim3(nchunks), dim3(256), 0, s, a, chunks, part);
    hipLaunchKernelGGL(gather_final_kernel, dim3(ntiles), dim3(256), 0, s, a, chunks, cJ_off, cK_off, part);
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

// The address of the hardware queue (HSA queue descriptor) the dispatch came from, into out[k]:
// streams that HIP placed on one hardware queue run their kernels one after another.
__global__ void queue_probe_kernel(unsigned long long* out, int k) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[k] = (unsigned long long)(size_t)__builtin_amdgcn_queue_ptr();
}

hipError_t launch_queue_probe(unsigned long long* out, int k, hipStream_t s) {
    hipLaunchKernelGGL(queue_probe_kernel, dim3(1), dim3(64), 0, s, out, k);
    return hipGetLastError();
}

// Plain streaming copy, 16 bytes per lane and iteration (a kernel of known traffic, to scale a
// profiler's HBM byte counters).
__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const double2* src, double2* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ device-side density helpers
// z = a x + b y (the density difference of an incremental build made on the device).
__global__ __launch_bounds__(256) void axpby_kernel(double a, const double* __restrict__ x, double b,
                                                    const double* __restrict__ y, double* __restrict__ z, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) z[i] = a * x[i] + b * y[i];
}

hipError_t launch_axpby(double a, const double* x, double b, const double* y, double* z, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(axpby_kernel, dim3(g), dim3(256), 0, s, a, x, b, y, z, n);
    return hipGetLastError();
}

// Largest |D| of every shell block of every molecule (one workgroup per shell row), rounded up
// to the next float so the single-precision table never understates the density-screening
// bound; the device form of the table the host builds for hfg_fock.
__global__ __launch_bounds__(64) void block_absmax_kernel(const double* __restrict__ D, const MolDev* __restrict__ mols,
                                                          const int* __restrict__ shell_ao,
                                                          const int* __restrict__ shell_nf,
                                                          const int* __restrict__ shell_mol, float* __restrict__ dmax) {
    const int ga = blockIdx.x;
    const MolDev md = mols[shell_mol[ga]];
    const int a = ga - md.shell0, oa = shell_ao[ga], na = shell_nf[ga], n = md.n;
    const double* Dm = D + md.moff;
    for (int b = threadIdx.x; b < md.nshell; b += blockDim.x) {
        const int gb = md.shell0 + b, ob = shell_ao[gb], nb = shell_nf[gb];
        double mx = 0.0;
        for (int p = 0; p < na; ++p)
            for (int q = 0; q < nb; ++q) mx = fmax(mx, fabs(Dm[(size_t)(oa + p) * n + ob + q]));
        dmax[md.dm_off + a * md.nshell + b] = mx > 0.0 ? nextafterf((float)mx, INFINITY) : 0.0f;
    }
}

hipError_t launch_block_absmax(const double* D, const MolDev* mols, const int* shell_ao, const int* shell_nf,
                               const int* shell_mol, int nshell, float* dmax, hipStream_t s) {
    if (nshell <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_absmax_kernel, dim3(nshell), dim3(64), 0, s, D, mols, shell_ao, shell_nf, shell_mol, dmax);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Boys probe (test hook)
// F_0..F_M at n values of T through the device functions the class kernels use: the table
// with Taylor step and recursions (boys_eval_rs), or the closed forms of F_0 and F_1 (boys_low).
template <int M>
__global__ void boys_probe_kernel(const double* __restrict__ T, const double* __restrict__ tab, double* out, int n,
                                  int closed_form) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double F[M + 1];
    if constexpr (M <= 1) {
        if (closed_form) {
            boys_low<M, double>(T[i], 1.0, F);
        } else {
            boys_eval_rs<M, double>(T[i], tab, 1.0, F);
        }
    } else {
        boys_eval_rs<M, double>(T[i], tab, 1.0, F);
    }
    for (int m = 0; m <= M; ++m) out[(size_t)i * (M + 1) + m] = F[m];
}

template <int M>
hipError_t launch_boys_probe_m(int m, const double* T, const double* tab, double* out, int n, int closed_form,
                               hipStream_t s) {
    if constexpr (M > kBoysMaxM) {
        return hipErrorInvalidValue;
    } else {
        if (m != M) return launch_boys_probe_m<M + 1>(m, T, tab, out, n, closed_form, s);
        hipLaunchKernelGGL(boys_probe_kernel<M>, dim3((n + 127) / 128), dim3(128), 0, s, T, tab, out, n,
                           closed_form);
        return hipGetLastError();
    }
}

hipError_t launch_boys_probe(int m, const double* T, const double* tab, double* out, int n, int closed_form,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (m < 0 || m > kBoysMaxM || (closed_form && m > 1)) return hipErrorInvalidValue;
    return launch_boys_probe_m<0>(m, T, tab, out, n, closed_form, s);
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
