// Restricted Hartree-Fock SCF driver and RHF gradient, shaped like hfscf's
// Mol::hf_solver (hfscf: src/scfsolver/hfscf_hf_solver.hpp:23-121),
// DIIS_SOLVER::diis_solver (hfscf: src/scfsolver/hfscf_diis.hpp:18-63) and
// Mol::scf_gradient::calc_scf_gradient_rhf (hfscf: src/gradient/hfscf_gradient.hpp:34-37).
//
// The two-electron work (J/K Fock build and the two-electron gradient contraction) goes
// through a JKBuilder, the run-time-selectable seam that the GPU library implements.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "basis.hpp"

namespace hfscf {

// Two-electron backend. D is the half density C_occ C_occ^T (nbf x nbf, row-major).
struct JKBuilder {
    virtual ~JKBuilder() = default;
    // J_ij = sum_kl D_kl (ij|kl), K_ij = sum_kl D_kl (ik|jl)
    virtual void build_jk(const double* D, double* J, double* K) = 0;
    // F = H + 2J - K. With incremental != 0 the backend may build from D - D_prev and add
    // to its own previous Fock matrix.
    virtual void fock(const double* H, const double* D, int incremental, double* F) = 0;
    // Two-electron part of the RHF gradient, sum_ijkl D_ij D_kl (2 (ij|kl)' - (ik|jl)'), natom x 3.
    virtual void gradient_2e(const double* D, double* grad) = 0;
    // Full-precision rebuild hint for the final iteration of a mixed-precision run.
    virtual void set_final(bool final_build) { (void)final_build; }
    // End of the reduced-precision phase of a mixed-precision run: every later build is FP64,
    // and the next one a full build (the accumulated FP32 error of the incremental chain goes).
    virtual void begin_fp64_phase() {}
    // Device-resident form of fock(): H, D and F are device pointers on device(). A backend
    // without one returns false (device() < 0), and the solver keeps its host loop.
    virtual bool fock_device(const double* H, const double* D, int incremental, double* F) {
        (void)H, (void)D, (void)incremental, (void)F;
        return false;
    }
    virtual int device() const { return -1; }
    // One-electron integrals S, T, V (n x n) from the backend's device; false when it has none and
    // the host layer computes them.
    virtual bool oneint(double* S, double* T, double* V) {
        (void)S, (void)T, (void)V;
        return false;
    }
    // grad += 2 D.d(T+V) - 2 Q.dS from the backend's device; false when it has none.
    virtual bool oneint_gradient(const double* D, const double* Q, double* grad) {
        (void)D, (void)Q, (void)grad;
        return false;
    }
};

struct ScfOptions {
    int diis_range = 6;  // hfscf keeps diis_range+1 vectors (hfscf: src/scfsolver/hfscf_diis.cpp:29-47)
    double energy_tol = 1e-10;
    double rms_tol = 1e-10;
    int max_iter = 100;
    int charge = 0;
    bool incremental = false;  // incremental (delta-D) Fock builds
    bool final_build = true;   // full-precision Fock build at the converged density
    int verbose = 0;
    // incremental builds: every rebuild_every-th iteration is a full build, which removes the
    // error that density screening accumulates along the incremental chain (0 = never)
    int rebuild_every = 0;
    // mixed precision: once the DIIS rms falls below polish_rms, the run continues with FP64
    // builds from a fresh full build and a fresh DIIS subspace, and converges to rms_tol /
    // energy_tol there, so D and F are those of the FP64 SCF (0 = stay mixed to the end)
    double polish_rms = 0.0;
    bool mixed = false;
    // with a backend that builds on device memory, run the whole SCF step on the GPU
    // (scf_run_device: DIIS, diagonalisation, density and energy in rocBLAS / rocSOLVER)
    bool device_step = true;
};

class diis_solver {
   public:
    explicit diis_solver(int range) : range_(range) {}
    // hfscf: src/scfsolver/hfscf_diis.cpp:9-123
    void diis_extrapolate(int n, double* F, const double* S, const double* D, const double* X);
    double current_rms() const { return rms_; }

   private:
    int range_;
    double rms_ = 1.0;
    std::vector<std::vector<double>> e_list_, f_list_;
    std::vector<double> rms_list_;
};

class hf_solver {
   public:
    hf_solver(const Molecule& mol, const ScfOptions& opt, JKBuilder* jk);
    virtual ~hf_solver() = default;
    void init_data();  // hfscf: src/scfsolver/hfscf_hf_solver.cpp:79-182 (CORE guess)
    bool scf_run();    // hfscf: src/scfsolver/hfscf_hf_solver.cpp:389-476

    // The iteration of scf_run split for drivers that build the Fock matrices of many
    // molecules at once: begin() starts the loop; step(F) takes the Fock matrix of the current
    // density and does what update_fock_matrix does after the build (DIIS, new density,
    // energy); converged() is hfscf's test; finish(e) stores the energy of the final build
    // (sum D (H + F(D))) and forms the final coefficients.
    void begin();
    void step(const double* F, double fock_seconds);
    bool converged() const;
    void finish(double e_final_elec);
    // incremental build for the next iteration (iteration_ + 1): after the first, except every
    // rebuild_every-th and the first after the switch to FP64
    bool wants_incremental() const;
    // mixed precision: the switch to FP64 is due (DIIS rms below polish_rms); begin_polish()
    // restarts DIIS and counts the iterations of the FP64 phase
    bool wants_polish() const;
    void begin_polish();
    bool polishing() const { return polish_; }
    const ScfOptions& options() const { return opt_; }

    double energy_electronic() const { return e_cur_; }
    // E = sum D (H + F(D)) from one more full-precision Fock build at the converged density
    // (stationary in D, so free of the DIIS-extrapolation noise in hfscf's energy expression)
    double energy_final_total() const { return e_final_ + mol_.nuclear_repulsion(); }
    double energy_total() const { return e_cur_ + mol_.nuclear_repulsion(); }
    int iterations() const { return iteration_; }
    int nbf() const { return n_; }
    const std::vector<double>& density() const { return D_; }
    const std::vector<double>& fock() const { return F_; }
    const std::vector<double>& hcore() const { return H_; }
    const std::vector<double>& overlap() const { return S_; }
    const std::vector<double>& mo_energies() const { return eps_; }
    const std::vector<double>& energies() const { return energies_; }
    double last_fock_seconds() const { return fock_seconds_; }
    double total_fock_seconds() const { return fock_total_; }

   protected:
    virtual void update_fock_matrix();  // hfscf: src/scfsolver/hfscf_hf_solver.cpp:254-286
    void after_fock();                  // DIIS, F' = X^T F X, density and energy
    void calc_density_matrix();         // hfscf: src/scfsolver/hfscf_hf_solver.cpp:206-242
    void calc_scf_energy();             // hfscf: src/scfsolver/hfscf_hf_solver.cpp:244-252
    bool on_device() const { return opt_.device_step && jk_ && jk_->device() >= 0; }
    bool scf_run_device();  // the loop of scf_run with every matrix on the GPU (devscf.cpp)

    const Molecule& mol_;
    ScfOptions opt_;
    JKBuilder* jk_;
    int n_ = 0, nocc_ = 0, iteration_ = 0;
    std::vector<double> S_, T_, V_, H_, X_, F_, Fp_, D_, C_, eps_, energies_;
    double e_cur_ = 0.0, e_prev_ = 0.0, e_final_ = 0.0;
    double fock_seconds_ = 0.0, fock_total_ = 0.0;
    std::unique_ptr<diis_solver> diis_;
    bool polish_ = false;
    int chain_start_ = 1;  // iteration of the last full build
    double device_rms_ = 0.0;
};

// RHF analytic gradient (natom x 3) from the converged half density D and Fock matrix F:
// g = sum D.dG/dX + 2 sum D.(dT + dV) - 2 sum dS.(D F D) + dV_NN
// (hfscf: src/gradient/hfscf_grad_rhf.cpp:8-178).
std::vector<double> calc_scf_gradient_rhf(const Molecule& mol, const double* D, const double* F, JKBuilder* jk);

}  // namespace hfscf
