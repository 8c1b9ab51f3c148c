#include "scf.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "linalg.hpp"
#include "oneint.hpp"

namespace hfscf {

void diis_solver::diis_extrapolate(int n, double* F, const double* S, const double* D, const double* X) {
    const size_t nn = (size_t)n * n;
    // error vector e = X^T (F D S - (F D S)^T) X
    std::vector<double> FD(nn), FDS(nn), G(nn), T(nn), E(nn);
    gemm(n, F, false, D, false, FD.data());
    gemm(n, FD.data(), false, S, false, FDS.data());
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G[(size_t)i * n + j] = FDS[(size_t)i * n + j] - FDS[(size_t)j * n + i];
    gemm(n, X, true, G.data(), false, T.data());
    gemm(n, T.data(), false, X, false, E.data());
    double ss = 0.0;
    for (double v : E) ss += v * v;
    const double rms = std::sqrt(ss / (double)nn);

    if (range_ == 0) {
        rms_ = 0.5 * rms;  // hfscf: src/scfsolver/hfscf_diis.cpp:18-27
        return;
    }
    if ((int)rms_list_.size() > range_) {
        // evict the vector with the largest error before appending (hfscf_diis.cpp:29-47)
        size_t pos = 0;
        double worst = 0.0;
        for (size_t i = 0; i < rms_list_.size(); ++i)
            if (rms_list_[i] > worst) {
                worst = rms_list_[i];
                pos = i;
            }
        e_list_.erase(e_list_.begin() + pos);
        f_list_.erase(f_list_.begin() + pos);
        rms_list_.erase(rms_list_.begin() + pos);
    }
    f_list_.emplace_back(F, F + nn);
    e_list_.push_back(E);
    rms_ = rms;
    rms_list_.push_back(rms);
    if (rms_list_.size() < 2) return;

    const int m = (int)e_list_.size();
    std::vector<double> B((size_t)(m + 1) * (m + 1), 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double s = 0.0;
            const double* a = e_list_[i].data();
            const double* b = e_list_[j].data();
            for (size_t k = 0; k < nn; ++k) s += a[k] * b[k];
            B[(size_t)i * (m + 1) + j] = s;
        }
    double absum = 0.0, maxc = -1e300;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            absum += std::fabs(B[(size_t)i * (m + 1) + j]);
            maxc = std::max(maxc, B[(size_t)i * (m + 1) + j]);
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
    auto c = qr_solve(m + 1, B, rhs);
    for (size_t k = 0; k < nn; ++k) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += c[i] * f_list_[i][k];
        F[k] = s;
    }
}

hf_solver::hf_solver(const Molecule& mol, const ScfOptions& opt, JKBuilder* jk) : mol_(mol), opt_(opt), jk_(jk) {
    n_ = mol.nbf();
    const int ne = mol.nelectrons(opt.charge);
    if (ne % 2 != 0) throw std::runtime_error("RHF needs an even number of electrons");
    nocc_ = ne / 2;
}

void hf_solver::init_data() {
    const size_t nn = (size_t)n_ * n_;
    iteration_ = 0;
    energies_.clear();
    S_.resize(nn);
    T_.resize(nn);
    V_.resize(nn);
    if (!jk_ || !jk_->oneint(S_.data(), T_.data(), V_.data())) {  // the GPU backend computes them on the device
        overlap_kinetic(mol_, S_, T_);
        nuclear_attraction(mol_, V_);
    }
    H_.resize(nn);
    for (size_t k = 0; k < nn; ++k) H_[k] = T_[k] + V_[k];
    if (on_device()) return;  // X, the guess and every later step are formed on the GPU
    // symmetric orthogonalisation X = S^-1/2 (hfscf_hf_solver.cpp:184-204)
    std::vector<double> w(n_), U(nn);
    eigh(n_, S_.data(), w.data(), U.data());
    X_.assign(nn, 0.0);
    for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_; ++j) {
            double s = 0.0;
            for (int k = 0; k < n_; ++k) s += U[(size_t)i * n_ + k] * U[(size_t)j * n_ + k] / std::sqrt(w[k]);
            X_[(size_t)i * n_ + j] = s;
        }
    std::vector<double> tmp(nn);
    Fp_.resize(nn);
    gemm(n_, X_.data(), true, H_.data(), false, tmp.data());
    gemm(n_, tmp.data(), false, X_.data(), false, Fp_.data());
    F_ = H_;
    D_.assign(nn, 0.0);
    calc_density_matrix();  // CORE guess
}

void hf_solver::calc_density_matrix() {
    const size_t nn = (size_t)n_ * n_;
    std::vector<double> U(nn);
    eps_.resize(n_);
    eigh(n_, Fp_.data(), eps_.data(), U.data());
    C_.resize(nn);
    gemm(n_, X_.data(), false, U.data(), false, C_.data());
    for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_; ++j) {
            double s = 0.0;
            for (int k = 0; k < nocc_; ++k) s += C_[(size_t)i * n_ + k] * C_[(size_t)j * n_ + k];
            D_[(size_t)i * n_ + j] = s;
        }
    calc_scf_energy();
}

void hf_solver::calc_scf_energy() {
    e_prev_ = e_cur_;
    double e = 0.0;
    const size_t nn = (size_t)n_ * n_;
    // iteration 0 uses H + H, later the previous (extrapolated) Fock matrix (hfscf_hf_solver.cpp:248)
    for (size_t k = 0; k < nn; ++k) e += D_[k] * (H_[k] + (iteration_ > 0 ? F_[k] : H_[k]));
    e_cur_ = e;
    energies_.push_back(e);
}

bool hf_solver::wants_incremental() const {
    const int next = iteration_ + 1;
    if (!opt_.incremental || next <= chain_start_) return false;
    return opt_.rebuild_every <= 0 || (next - chain_start_) % opt_.rebuild_every != 0;
}

bool hf_solver::wants_polish() const {
    return opt_.mixed && opt_.polish_rms > 0.0 && !polish_ && diis_->current_rms() < opt_.polish_rms;
}

void hf_solver::begin_polish() {
    polish_ = true;
    chain_start_ = iteration_ + 1;  // the next build is a full FP64 build
    diis_ = std::make_unique<diis_solver>(opt_.diis_range);
}

void hf_solver::update_fock_matrix() {
    const bool inc = wants_incremental();
    ++iteration_;
    if (!inc) chain_start_ = iteration_;
    auto t0 = std::chrono::steady_clock::now();
    jk_->fock(H_.data(), D_.data(), inc ? 1 : 0, F_.data());
    auto t1 = std::chrono::steady_clock::now();
    fock_seconds_ = std::chrono::duration<double>(t1 - t0).count();
    fock_total_ += fock_seconds_;
    after_fock();
}

void hf_solver::after_fock() {
    diis_->diis_extrapolate(n_, F_.data(), S_.data(), D_.data(), X_.data());
    std::vector<double> tmp((size_t)n_ * n_);
    gemm(n_, X_.data(), true, F_.data(), false, tmp.data());
    gemm(n_, tmp.data(), false, X_.data(), false, Fp_.data());
    calc_density_matrix();
}

void hf_solver::begin() { diis_ = std::make_unique<diis_solver>(opt_.diis_range); }

void hf_solver::step(const double* F, double fock_seconds) {
    if (!wants_incremental()) chain_start_ = iteration_ + 1;
    ++iteration_;
    std::copy(F, F + (size_t)n_ * n_, F_.begin());
    fock_seconds_ = fock_seconds;
    fock_total_ += fock_seconds;
    after_fock();
}

bool hf_solver::converged() const {
    // hfscf_hf_solver.cpp:415; a mixed run with polishing converges only in its FP64 phase
    if (opt_.mixed && opt_.polish_rms > 0.0 && !polish_) return false;
    return diis_->current_rms() < opt_.rms_tol && std::fabs(e_cur_ - e_prev_) < opt_.energy_tol;
}

void hf_solver::finish(double e_final_elec) {
    e_final_ = e_final_elec;
    // final coefficients from the last orthogonalised Fock matrix (hfscf_hf_solver.cpp:469-473)
    std::vector<double> U((size_t)n_ * n_);
    eps_.resize(n_);
    eigh(n_, Fp_.data(), eps_.data(), U.data());
    gemm(n_, X_.data(), false, U.data(), false, C_.data());
}

bool hf_solver::scf_run() {
    if (on_device()) return scf_run_device();
    begin();
    bool conv = false;
    for (;;) {
        update_fock_matrix();
        if (opt_.verbose)
            std::fprintf(stderr, "  iter %3d  E = %.12f  dE = %.3e  rms = %.3e  fock %.4fs\n", iteration_, e_cur_,
                         e_cur_ - e_prev_, diis_->current_rms(), fock_seconds_);
        if (wants_polish()) {
            begin_polish();
            jk_->begin_fp64_phase();
        }
        if (converged()) {
            conv = true;
            break;
        }
        if (iteration_ >= opt_.max_iter) break;
    }
    double e_final = e_cur_;
    if (opt_.final_build) {
        jk_->set_final(true);
        std::vector<double> Ff((size_t)n_ * n_);
        jk_->fock(H_.data(), D_.data(), 0, Ff.data());
        e_final = 0.0;
        for (size_t k = 0; k < Ff.size(); ++k) e_final += D_[k] * (H_[k] + Ff[k]);
        jk_->set_final(false);
    }
    finish(e_final);
    return conv;
}

std::vector<double> calc_scf_gradient_rhf(const Molecule& mol, const double* D, const double* F, JKBuilder* jk) {
    const int n = mol.nbf();
    const size_t nn = (size_t)n * n;
    std::vector<double> DF(nn), Q(nn);
    gemm(n, D, false, F, false, DF.data());
    gemm(n, DF.data(), false, D, false, Q.data());
    std::vector<double> g((size_t)mol.natom() * 3, 0.0);
    jk->gradient_2e(D, g.data());
    if (!jk->oneint_gradient(D, Q.data(), g.data())) oneint_gradient(mol, D, Q.data(), g.data());
    nuclear_repulsion_gradient(mol, g.data());
    return g;
}

}  // namespace hfscf
