// One-electron integrals of a single molecule on the GPU (S, T, V and their gradient terms).
#pragma once
#include <hip/hip_runtime.h>

struct hfg_ctx;

namespace hfg {

// A shell for the one-electron kernels: angular momentum, primitive count, offset of its
// exponents / normalised coefficients, atom, first output-basis function.
struct OneShell {
    int L, np, po, atom, ao;
};

struct OneArgs {
    const OneShell* sh;
    const double* alpha;
    const double* coef;
    const double* xyz;   // natom x 3, bohr
    const double* Z;     // nuclear charges
    const double* boys;  // the Boys table
    const int2* pairs;   // shell pairs (a, b), a >= b, of the S/T/V kernel
    int npairs;
    int natom, nshell, n, pure;
    double* S;  // n x n outputs (device)
    double* T;
    double* V;
};

// S, T and V (n x n, both triangles) of every pair in g.pairs.
hipError_t launch_oneint(const OneArgs& g, hipStream_t s);
// Device copies of the context's single molecule (shells, exponents, normalised coefficients,
// geometry, charges) into g; an HFG_ error code (message in the context) if it has none.
int one_upload(hfg_ctx* c, OneArgs& g);

// grad (natom x 3, device) += 2 D d(T+V)/dX - 2 Q dS/dX over every ordered shell pair.
hipError_t launch_oneint_grad(const OneArgs& g, const double* D, const double* Q, double* grad, hipStream_t s);

}  // namespace hfg
