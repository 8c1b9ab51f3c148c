// Run-time dispatch of the adjoint gradient kernels and the fixed-order slab reduction.
#include "coop_grad.hpp"
#include "launch_bounds.hpp"

namespace hfg {

#define HFG_DECL1(A, B, C, D, P)                                                                    \
    template <>                                                                                     \
    hipError_t launch_grad_class64<A, B, C, D, P>(const GradArgs&, int, LaunchShape, hipStream_t);  \
    template <>                                                                                     \
    hipError_t launch_grad_class32<A, B, C, D, P>(const GradArgs&, int, LaunchShape, hipStream_t);
#define HFG_DECL(A, B, C, D) HFG_DECL1(A, B, C, D, true) HFG_DECL1(A, B, C, D, false)
HFG_FOR_CLASSES(HFG_DECL)
#undef HFG_DECL
#undef HFG_DECL1

namespace {

// The register kernels (FP64 or FP32) of the classes below L = 7; the classes from L = 7 up run
// the cooperative kernel (coop_grad.hpp), launched by the context, and have no register kernel.
template <int A, int B, int C, int D, bool P>
hipError_t grad_class(const GradArgs& a, int nitems, LaunchShape sh, int mixed, hipStream_t s) {
    if constexpr (grad_uses_coop(A, B, C, D)) {
        return hipErrorInvalidValue;
    } else {
        const LaunchShape bs = scratch_safe_shape(A, B, C, D, true, sh, nitems);
        return mixed ? launch_grad_class32<A, B, C, D, P>(a, nitems, bs, s)
                     : launch_grad_class64<A, B, C, D, P>(a, nitems, bs, s);
    }
}

}  // namespace

hipError_t launch_grad(int la, int lb, int lc, int ld, bool pure, int mixed, const GradArgs& a, int nitems,
                       LaunchShape want, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
#define HFG_CASE(A, B, C, D)                                                                        \
    if (la == A && lb == B && lc == C && ld == D)                                                   \
        return pure ? grad_class<A, B, C, D, true>(a, nitems, want, mixed, s)                       \
                    : grad_class<A, B, C, D, false>(a, nitems, want, mixed, s);
    HFG_FOR_CLASSES(HFG_CASE)
#undef HFG_CASE
    return hipErrorInvalidValue;
}

// out[k] = sum over the work items of k's molecule of their partial k, summed in a fixed order
// (strided per thread, then a tree in LDS) so the result is reproducible.
__global__ __launch_bounds__(256) void grad_reduce_kernel(GradReduceArgs a) {
    __shared__ double part[256];
    const int o = blockIdx.x;
    const int m = a.out_mol[o], k = a.out_k[o];
    const int i0 = a.mol_items_off[m], i1 = a.mol_items_off[m + 1];
    double s = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += a.slab[a.items[a.mol_items[i]].gslab + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.out[o] = part[0];
}

hipError_t launch_grad_reduce(const GradReduceArgs& a, int nout, hipStream_t s) {
    if (nout == 0) return hipSuccess;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(nout), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace hfg
