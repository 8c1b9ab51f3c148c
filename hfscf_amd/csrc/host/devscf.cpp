// Device-resident SCF step of hf_solver: the loop of scf_run with every n x n matrix on the GPU.
// The Fock build goes through JKBuilder::fock_device (hfg_fock_device, no host copies); DIIS
// (error vectors, B matrix, extrapolation), F' = X^T F X, the diagonalisation, the density and
// the energy run in rocBLAS / rocSOLVER on device buffers. The host keeps the scalars (rms,
// energies, the (m+1) x (m+1) DIIS system and its QR solve) and copies D, F, C and the orbital
// energies back once at the end.
//
// Semantics are hf_solver's (hfscf: src/scfsolver/hfscf_hf_solver.cpp:184-252,389-476,
// src/scfsolver/hfscf_diis.cpp:9-123): X = S^-1/2 by symmetric orthogonalisation, CORE guess,
// energy sum D (H + F) with the previous extrapolated F (H + H at iteration 0), DIIS error
// X^T (F D S - (F D S)^T) X with rms sqrt(mean e^2), the largest-rms vector evicted before a new
// one is appended once more than diis_range are held, extrapolation from two vectors on,
// B normalised when sum |B| < 1e-14, hfscf's convergence test; the FP64 finish of a mixed run
// and the periodic full rebuilds as in the host loop.
//
// rocBLAS is column-major. F, D, S, H and X are symmetric, so their row-major arrays are also
// their column-major arrays and every product below is the same matrix; the DIIS error comes out
// transposed (an antisymmetric matrix negated), which leaves its norm and every dot product of
// two error vectors unchanged. The MO coefficients are transposed on the way back.
#include <hip/hip_runtime_api.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "scf.hpp"

namespace hfscf {

namespace {

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("device SCF step: ") + what + ": " + hipGetErrorString(e));
}

void bcheck(rocblas_status s, const char* what) {
    if (s != rocblas_status_success)
        throw std::runtime_error(std::string("device SCF step: ") + what + ": " + rocblas_status_to_string(s));
}

// Device buffers, the rocBLAS handle and its stream of one SCF run; freed on every exit path.
struct DevCtx {
    rocblas_handle h = nullptr;
    hipStream_t st = nullptr;
    std::vector<void*> bufs;
    explicit DevCtx(int device) {
        hcheck(hipSetDevice(device), "hipSetDevice");
        hcheck(hipStreamCreate(&st), "hipStreamCreate");
        bcheck(rocblas_create_handle(&h), "rocblas_create_handle");
        bcheck(rocblas_set_stream(h, st), "rocblas_set_stream");
        bcheck(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "rocblas_set_pointer_mode");
    }
    ~DevCtx() {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : bufs) (void)hipFree(p);
        if (h) (void)rocblas_destroy_handle(h);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        hcheck(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        bufs.push_back(p);
        return (T*)p;
    }
    void upload(double* d, const double* h_, size_t n) {
        hcheck(hipMemcpyAsync(d, h_, n * sizeof(double), hipMemcpyHostToDevice, st), "upload");
    }
    void download(double* h_, const double* d, size_t n) {
        hcheck(hipMemcpyAsync(h_, d, n * sizeof(double), hipMemcpyDeviceToHost, st), "download");
        hcheck(hipStreamSynchronize(st), "download sync");
    }
    void copy(double* dst, const double* src, size_t n) {
        hcheck(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    }
    void sync() { hcheck(hipStreamSynchronize(st), "sync"); }
    // C = op(A) op(B), n x n (k inner when given)
    void gemm(bool ta, bool tb, int n, const double* A, const double* B, double* C, int k = -1) {
        const double one = 1.0, zero = 0.0;
        const int kk = k < 0 ? n : k;
        bcheck(rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                             tb ? rocblas_operation_transpose : rocblas_operation_none, n, n, kk, &one, A, n, B, n,
                             &zero, C, n),
               "rocblas_dgemm");
    }
    double dot(const double* x, const double* y, size_t n) {
        double r = 0.0;
        bcheck(rocblas_ddot(h, (rocblas_int)n, x, 1, y, 1, &r), "rocblas_ddot");
        return r;
    }
    // A (symmetric) -> eigenvectors in A (columns), eigenvalues ascending in w (device)
    void eigh(int n, double* A, double* w, double* e, rocblas_int* info) {
        bcheck(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, e, info),
               "rocsolver_dsyevd");
        rocblas_int hinfo = 0;
        hcheck(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, st), "info");
        sync();
        if (hinfo != 0) throw std::runtime_error("device SCF step: rocsolver_dsyevd did not converge");
    }
};

}  // namespace

bool hf_solver::scf_run_device() {
    const int n = n_;
    const size_t nn = (size_t)n * n;
    const int range = opt_.diis_range;
    const int nslot = std::max(range, 0) + 2;
    DevCtx dv(jk_->device());
    double* S = dv.alloc<double>(nn);
    double* H = dv.alloc<double>(nn);
    double* X = dv.alloc<double>(nn);
    double* F = dv.alloc<double>(nn);
    double* D = dv.alloc<double>(nn);
    double* C = dv.alloc<double>(nn);
    double* V = dv.alloc<double>(nn);
    double* T1 = dv.alloc<double>(nn);
    double* T2 = dv.alloc<double>(nn);
    double* Enew = dv.alloc<double>(nn);
    double* W = dv.alloc<double>(n);
    double* Wk = dv.alloc<double>(n);
    double* cvec = dv.alloc<double>(nslot);
    double* eslot = dv.alloc<double>(nn * nslot);
    double* fslot = dv.alloc<double>(nn * nslot);
    rocblas_int* info = dv.alloc<rocblas_int>(1);
    // the extrapolation is one product over every slot with zero weights on the unused ones, so
    // the slots start at zero: 0 x (whatever hipMalloc's memory held, NaN included) is not 0
    hcheck(hipMemsetAsync(eslot, 0, nn * nslot * sizeof(double), dv.st), "hipMemsetAsync");
    hcheck(hipMemsetAsync(fslot, 0, nn * nslot * sizeof(double), dv.st), "hipMemsetAsync");
    dv.upload(S, S_.data(), nn);
    dv.upload(H, H_.data(), nn);

    // X = S^-1/2 = U diag(w^-1/2) U^T (hfscf_hf_solver.cpp:184-204)
    std::vector<double> w(n);
    dv.copy(V, S, nn);
    dv.eigh(n, V, W, Wk, info);
    dv.download(w.data(), W, n);
    for (int k = 0; k < n; ++k) {
        if (!(w[k] > 0.0)) throw std::runtime_error("device SCF step: overlap matrix is not positive definite");
        w[k] = 1.0 / std::sqrt(w[k]);
    }
    dv.upload(Wk, w.data(), n);
    bcheck(rocblas_ddgmm(dv.h, rocblas_side_right, n, n, V, n, Wk, 1, T1, n), "rocblas_ddgmm");
    dv.gemm(false, true, n, T1, V, X);

    // F' = X^T Fin X, its eigenvectors, C = X C', D = C_occ C_occ^T, orbital energies in W
    auto density = [&](const double* Fin) {
        dv.gemm(true, false, n, X, Fin, T1);
        dv.gemm(false, false, n, T1, X, V);
        dv.eigh(n, V, W, Wk, info);
        dv.gemm(false, false, n, X, V, C);
        dv.gemm(false, true, n, C, C, D, nocc_);
    };
    auto energy = [&](const double* Fin) { return dv.dot(D, H, nn) + dv.dot(D, Fin, nn); };

    // CORE guess; energy with H + H at iteration 0
    iteration_ = 0;
    energies_.clear();
    density(H);
    e_prev_ = e_cur_;
    e_cur_ = energy(H);
    energies_.push_back(e_cur_);

    // DIIS state: logical list of slots with their rms, dot products between slots
    std::vector<int> list;
    std::vector<double> rms_list;
    std::vector<double> dots((size_t)nslot * nslot, 0.0);
    double rms_cur = 1.0;
    auto diis = [&]() {
        // e = X^T (F D S - (F D S)^T) X
        dv.gemm(false, false, n, F, D, T1);
        dv.gemm(false, false, n, T1, S, T2);
        const double one = 1.0, mone = -1.0;
        bcheck(rocblas_dgeam(dv.h, rocblas_operation_none, rocblas_operation_transpose, n, n, &one, T2, n, &mone, T2,
                             n, T1, n),
               "rocblas_dgeam");
        dv.gemm(true, false, n, X, T1, T2);
        dv.gemm(false, false, n, T2, X, Enew);
        double nrm = 0.0;
        bcheck(rocblas_dnrm2(dv.h, (rocblas_int)nn, Enew, 1, &nrm), "rocblas_dnrm2");
        const double rms = nrm / std::sqrt((double)nn);
        if (range == 0) {
            rms_cur = 0.5 * rms;  // hfscf_diis.cpp:18-27
            return;
        }
        if ((int)rms_list.size() > range) {
            const size_t pos = (size_t)(std::max_element(rms_list.begin(), rms_list.end()) - rms_list.begin());
            list.erase(list.begin() + pos);
            rms_list.erase(rms_list.begin() + pos);
        }
        int slot = 0;
        while (std::find(list.begin(), list.end(), slot) != list.end()) ++slot;
        dv.copy(eslot + (size_t)slot * nn, Enew, nn);
        dv.copy(fslot + (size_t)slot * nn, F, nn);
        list.push_back(slot);
        rms_list.push_back(rms);
        rms_cur = rms;
        for (int s : list) {
            const double d = dv.dot(eslot + (size_t)slot * nn, eslot + (size_t)s * nn, nn);
            dots[(size_t)slot * nslot + s] = dots[(size_t)s * nslot + slot] = d;
        }
        if (list.size() < 2) return;
        const int m = (int)list.size();
        std::vector<double> B((size_t)(m + 1) * (m + 1), 0.0);
        double absum = 0.0, maxc = -1e300;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double b = dots[(size_t)list[i] * nslot + list[j]];
                B[(size_t)i * (m + 1) + j] = b;
                absum += std::fabs(b);
                maxc = std::max(maxc, b);
            }
        if (absum < 1e-14)
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) B[(size_t)i * (m + 1) + j] /= maxc;
        for (int i = 0; i <= m; ++i) {
            B[(size_t)m * (m + 1) + i] = -1.0;
            B[(size_t)i * (m + 1) + m] = -1.0;
        }
        B[(size_t)m * (m + 1) + m] = 0.0;
        std::vector<double> rhs(m + 1, 0.0);
        rhs[m] = -1.0;
        const auto c = qr_solve(m + 1, B, rhs);
        std::vector<double> cs(nslot, 0.0);
        for (int i = 0; i < m; ++i) cs[list[i]] = c[i];
        dv.upload(cvec, cs.data(), nslot);
        const double one1 = 1.0, zero = 0.0;
        bcheck(rocblas_dgemv(dv.h, rocblas_operation_none, (rocblas_int)nn, nslot, &one1, fslot, (rocblas_int)nn, cvec,
                             1, &zero, F, 1),
               "rocblas_dgemv");
    };

    bool conv = false;
    for (;;) {
        const bool inc = wants_incremental();
        ++iteration_;
        if (!inc) chain_start_ = iteration_;
        auto t0 = std::chrono::steady_clock::now();
        if (!jk_->fock_device(H, D, inc ? 1 : 0, F)) throw std::runtime_error("backend has no device Fock build");
        fock_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fock_total_ += fock_seconds_;
        diis();
        density(F);
        e_prev_ = e_cur_;
        e_cur_ = energy(F);
        energies_.push_back(e_cur_);
        if (opt_.verbose)
            std::fprintf(stderr, "  iter %3d  E = %.12f  dE = %.3e  rms = %.3e  fock %.4fs (device step)\n", iteration_,
                         e_cur_, e_cur_ - e_prev_, rms_cur, fock_seconds_);
        const bool polish_on = opt_.mixed && opt_.polish_rms > 0.0;
        if (polish_on && !polish_ && rms_cur < opt_.polish_rms) {
            // FP64 finish: fresh DIIS subspace, the next build a full FP64 build
            polish_ = true;
            chain_start_ = iteration_ + 1;
            list.clear();
            rms_list.clear();
            rms_cur = 1.0;
            jk_->begin_fp64_phase();
        }
        if ((!polish_on || polish_) && rms_cur < opt_.rms_tol && std::fabs(e_cur_ - e_prev_) < opt_.energy_tol) {
            conv = true;
            break;
        }
        if (iteration_ >= opt_.max_iter) break;
    }
    double e_final = e_cur_;
    if (opt_.final_build) {
        jk_->set_final(true);
        if (!jk_->fock_device(H, D, 0, T2)) throw std::runtime_error("backend has no device Fock build");
        e_final = energy(T2);
        jk_->set_final(false);
    }
    e_final_ = e_final;
    device_rms_ = rms_cur;
    // results to the host: D and F (symmetric), C transposed to row-major, orbital energies
    D_.resize(nn);
    F_.resize(nn);
    dv.download(D_.data(), D, nn);
    dv.download(F_.data(), F, nn);
    std::vector<double> Ct(nn);
    dv.download(Ct.data(), C, nn);
    C_.resize(nn);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) C_[(size_t)i * n + k] = Ct[(size_t)k * n + i];
    eps_.resize(n);
    dv.download(eps_.data(), W, n);
    return conv;
}

}  // namespace hfscf
