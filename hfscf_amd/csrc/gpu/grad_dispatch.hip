// Run-time dispatch of the adjoint gradient kernels and the fixed-order slab reduction.
#include "grad_impl.hpp"

namespace hfg {

#define HFG_DECL(A, B, C, D)                                                                 \
    template <>                                                                              \
    hipError_t launch_grad_class64<A, B, C, D>(bool, const GradArgs&, int, hipStream_t);     \
    template <>                                                                              \
    hipError_t launch_grad_class32<A, B, C, D>(bool, const GradArgs&, int, hipStream_t);
HFG_FOR_CLASSES(HFG_DECL)
#undef HFG_DECL

hipError_t launch_grad(int la, int lb, int lc, int ld, bool pure, int mixed, const GradArgs& a, int nitems,
                       hipStream_t s) {
    if (nitems == 0) return hipSuccess;
#define HFG_CASE(A, B, C, D)                                                                 \
    if (la == A && lb == B && lc == C && ld == D)                                            \
        return mixed ? launch_grad_class32<A, B, C, D>(pure, a, nitems, s)                   \
                     : launch_grad_class64<A, B, C, D>(pure, a, nitems, s);
    HFG_FOR_CLASSES(HFG_CASE)
#undef HFG_CASE
    return hipErrorInvalidValue;
}

// out[k] = sum over work items of slab[item * nout + k], summed in a fixed order
// (strided per thread, then a tree in LDS) so the result is reproducible.
__global__ __launch_bounds__(256) void grad_reduce_kernel(const double* slab, int nitems, int nout, double* out) {
    __shared__ double part[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nitems; i += blockDim.x) s += slab[(size_t)i * nout + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = part[0];
}

hipError_t launch_grad_reduce(const double* slab, int nitems, int nout, double* out, hipStream_t s) {
    if (nout == 0) return hipSuccess;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(nout), dim3(256), 0, s, slab, nitems, nout, out);
    return hipGetLastError();
}

}  // namespace hfg
