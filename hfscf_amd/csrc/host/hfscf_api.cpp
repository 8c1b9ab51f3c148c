// C ABI of the host RHF layer and the GPU two-electron backend (GpuJK) that plugs the
// hfgpu library into the hf_solver seam.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../../include/hfgpu.h"
#include "../../../include/hfscf.h"
#include "basis.hpp"
#include "input.hpp"
#include "linalg.hpp"
#include "oneint.hpp"
#include "scf.hpp"

using namespace hfscf;

struct hfscf_mol {
    Molecule mol;
};

namespace {

thread_local std::string g_err;

int set_err(const std::string& s) {
    g_err = s;
    return -1;
}

// Shell table of a molecule in the hfgpu layout.
struct BasisArrays {
    std::vector<int> L, atom, np, co, po;
    std::vector<double> alpha, coef, xyz;
    explicit BasisArrays(const Molecule& mol) {
        for (const Shell& sh : mol.shells) {
            L.push_back(sh.L);
            atom.push_back(sh.atom);
            np.push_back((int)sh.alpha.size());
            co.push_back(sh.idx);
            po.push_back(sh.ids);
            alpha.insert(alpha.end(), sh.alpha.begin(), sh.alpha.end());
            coef.insert(coef.end(), sh.c.begin(), sh.c.end());
        }
        for (auto& p : mol.xyz) xyz.insert(xyz.end(), p.begin(), p.end());
    }
};

hfg_options gpu_options(const hfscf_scf_options& o) {
    hfg_options go;
    hfg_default_options(&go);
    go.precision = o.precision;
    go.integral_tol = o.integral_tol;
    go.schwarz = o.schwarz;
    go.density_tol = o.density_tol;
    go.graph = o.graph;
    return go;
}

// Density threshold of one incremental build. The FP32 density-difference builds of a mixed run
// keep the caller's density_tol (their rounding error is the larger one). An FP64 build must not
// add more error than the energy tolerance allows, and the contributions of the quartets it
// skips add up over the build: on a 384-function water cluster a threshold t left SCF energies
// that wandered by a few hundred to a few thousand t from one iteration to the next. FP64
// builds therefore use at most 1e-5 x energy_tol.
double build_density_tol(const hfscf_scf_options& o, bool fp32_build) {
    return fp32_build ? o.density_tol : std::min(o.density_tol, 1e-5 * o.energy_tol);
}

// GPU backend: ERIOS/hf_solver/scf_gradient two-electron work through the hfgpu C ABI.
class GpuJK final : public JKBuilder {
   public:
    GpuJK(const Molecule& mol, const hfscf_scf_options& o)
        : mol_(mol), opt_(o), precision_(o.precision), device_(o.device) {
        const hfg_options go = gpu_options(o);
        int r = hfg_create(&ctx_, o.device, &go);
        if (r) throw std::runtime_error("hfg_create failed (" + std::to_string(r) + "): no usable GPU");
        BasisArrays b(mol);
        check(hfg_set_basis(ctx_, (int)b.L.size(), b.L.data(), b.atom.data(), b.np.data(), b.alpha.data(),
                            b.coef.data(), mol.pure ? 1 : 0, b.co.data(), b.po.data()));
        check(hfg_set_geometry(ctx_, mol.natom(), b.xyz.data(), mol.Z.data()));
    }
    ~GpuJK() override { hfg_destroy(ctx_); }
    void build_jk(const double* D, double* J, double* K) override { check(hfg_build_jk(ctx_, D, J, K, 0)); }
    void fock(const double* H, const double* D, int incremental, double* F) override {
        prepare(incremental);
        check(hfg_fock(ctx_, H, D, incremental, F));
        have_prev_ = true;
    }
    bool fock_device(const double* H, const double* D, int incremental, double* F) override {
        prepare(incremental);
        check(hfg_fock_device(ctx_, H, D, incremental, F));
        have_prev_ = true;
        return true;
    }
    int device() const override { return device_; }
    void begin_fp64_phase() override {
        fp64_ = true;
        have_prev_ = false;
        check(hfg_reset_incremental(ctx_));
    }
    void gradient_2e(const double* D, double* grad) override {
        std::vector<double> g((size_t)mol_.natom() * 3);
        check(hfg_set_precision(ctx_, precision_));
        check(hfg_energy_backward(ctx_, D, g.data()));
        for (size_t k = 0; k < g.size(); ++k) grad[k] += g[k];
    }
    void set_final(bool final_build) override { final_ = final_build; }
    bool oneint(double* S, double* T, double* V) override {
        check(hfg_oneint(ctx_, S, T, V));
        return true;
    }
    bool oneint_gradient(const double* D, const double* Q, double* grad) override {
        check(hfg_oneint_gradient(ctx_, D, Q, grad));
        return true;
    }

   private:
    void check(int r) {
        if (r) throw std::runtime_error(std::string("hfgpu: ") + hfg_last_error(ctx_));
    }
    // mixed precision: FP32 only for builds on a density difference, whose FP32 error shrinks
    // with the difference; full builds (the first, the final one) and the FP64 finish in FP64
    void prepare(int incremental) {
        const bool inc = incremental && have_prev_;
        const bool fp32 = precision_ == HFG_MIXED && inc && !final_ && !fp64_;
        check(hfg_set_precision(ctx_, fp32 ? HFG_MIXED : HFG_FP64));
        check(hfg_set_density_tol(ctx_, build_density_tol(opt_, fp32)));
    }
    const Molecule& mol_;
    const hfscf_scf_options opt_;
    hfg_ctx* ctx_ = nullptr;
    int precision_;
    int device_;
    bool have_prev_ = false, final_ = false, fp64_ = false;
};

// Callback backend (an external Fock builder behind the same seam).
class CallbackJK final : public JKBuilder {
   public:
    CallbackJK(const Molecule& mol, const hfscf_callbacks& cb) : mol_(mol), cb_(cb) {}
    void build_jk(const double* D, double* J, double* K) override {
        if (!cb_.build_jk || cb_.build_jk(cb_.user, D, J, K)) throw std::runtime_error("callback build_jk failed");
    }
    void fock(const double* H, const double* D, int incremental, double* F) override {
        (void)incremental;
        const size_t nn = (size_t)mol_.nbf() * mol_.nbf();
        std::vector<double> J(nn), K(nn);
        build_jk(D, J.data(), K.data());
        for (size_t k = 0; k < nn; ++k) F[k] = H[k] + 2.0 * J[k] - K[k];
    }
    void gradient_2e(const double* D, double* grad) override {
        std::vector<double> g((size_t)mol_.natom() * 3);
        if (!cb_.grad2e || cb_.grad2e(cb_.user, D, g.data())) throw std::runtime_error("callback grad2e failed");
        for (size_t k = 0; k < g.size(); ++k) grad[k] += g[k];
    }

   private:
    const Molecule& mol_;
    hfscf_callbacks cb_;
};

std::unique_ptr<JKBuilder> make_backend(const Molecule& mol, const hfscf_scf_options& o, const hfscf_callbacks* cb) {
    if (o.backend == HFSCF_BACKEND_CALLBACK) {
        if (!cb) throw std::runtime_error("callback backend selected without callbacks");
        return std::make_unique<CallbackJK>(mol, *cb);
    }
    if (o.backend != HFSCF_BACKEND_GPU) throw std::runtime_error("unknown backend");
    return std::make_unique<GpuJK>(mol, o);
}

// Lockstep SCF of many molecules with one batched GPU Fock build per iteration.
int batch_rhf(int nmol, hfscf_mol* const* mols, const hfscf_scf_options& opt, hfscf_scf_result* res, double* D_all,
              double* F_all, double* grad_all, hfscf_batch_stats* st) {
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    hfscf_batch_stats S{};
    const auto t_setup = clk::now();
    const hfg_options go = gpu_options(opt);
    hfg_ctx* ctx = nullptr;
    int r = hfg_create(&ctx, opt.device, &go);
    if (r) throw std::runtime_error("hfg_create failed (" + std::to_string(r) + "): no usable GPU");
    std::unique_ptr<hfg_ctx, void (*)(hfg_ctx*)> guard(ctx, hfg_destroy);
    auto check = [&](int rc) {
        if (rc < 0) throw std::runtime_error(std::string("hfgpu: ") + hfg_last_error(ctx));
    };
    for (int m = 0; m < nmol; ++m) {
        const Molecule& mol = mols[m]->mol;
        BasisArrays b(mol);
        check(hfg_batch_add_molecule(ctx, (int)b.L.size(), b.L.data(), b.atom.data(), b.np.data(), b.alpha.data(),
                                     b.coef.data(), mol.pure ? 1 : 0, b.co.data(), b.po.data(), mol.natom(),
                                     b.xyz.data()));
    }
    check(hfg_batch_commit(ctx));
    std::vector<long long> moff(nmol);
    std::vector<int> nbf(nmol);
    check(hfg_batch_layout(ctx, nullptr, nbf.data(), moff.data(), nullptr));
    long long nn = 0;
    for (int m = 0; m < nmol; ++m) nn = std::max(nn, moff[m] + (long long)nbf[m] * nbf[m]);
    ScfOptions so;
    so.diis_range = opt.diis_range;
    so.energy_tol = opt.energy_tol;
    so.rms_tol = opt.rms_tol;
    so.max_iter = opt.max_iter;
    so.incremental = opt.incremental != 0;
    so.final_build = opt.final_fp64 != 0;
    so.verbose = 0;
    so.rebuild_every = opt.rebuild_every;  // no FP64 polish phase in batches (so.mixed stays false)
    std::vector<std::unique_ptr<hf_solver>> sol(nmol);
    std::vector<double> H(nn), D(nn), F(nn);
    // one-electron set-up per molecule (independent, so in parallel)
    std::string err;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < nmol; ++m) {
        try {
            sol[m] = std::make_unique<hf_solver>(mols[m]->mol, so, nullptr);
            sol[m]->init_data();
            sol[m]->begin();
            std::copy(sol[m]->hcore().begin(), sol[m]->hcore().end(), H.begin() + moff[m]);
        } catch (const std::exception& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw std::runtime_error(err);
    S.setup_seconds = secs(t_setup, clk::now());
    std::vector<char> active(nmol, 1), conv(nmol, 0);
    int chain_start = 1;  // iteration of the last full build
    const auto t_scf = clk::now();
    for (int it = 1;; ++it) {
        int nact = 0;
        for (int m = 0; m < nmol; ++m) nact += active[m];
        if (!nact) break;
        for (int m = 0; m < nmol; ++m)
            std::copy(sol[m]->density().begin(), sol[m]->density().end(), D.begin() + moff[m]);
        // as GpuJK: the first (full) build in FP64, density-difference builds in the chosen
        // precision; every rebuild_every-th build is a full one (as hf_solver::wants_incremental
        // decides for a solo run, so that each molecule follows its solo trajectory)
        const int inc = (opt.incremental && it > chain_start &&
                         (opt.rebuild_every <= 0 || (it - chain_start) % opt.rebuild_every != 0)) ? 1 : 0;
        if (!inc) chain_start = it;
        const bool fp32 = opt.precision == HFG_MIXED && inc;
        check(hfg_set_precision(ctx, fp32 ? HFG_MIXED : HFG_FP64));
        check(hfg_set_density_tol(ctx, build_density_tol(opt, fp32)));
        const auto t0 = clk::now();
        check(hfg_fock(ctx, H.data(), D.data(), inc, F.data()));
        const double tf = secs(t0, clk::now());
        S.fock_seconds += tf;
        S.iterations = it;
        if (it == 1 && opt.autotune) {
            const auto ta = clk::now();
            check(hfg_autotune(ctx, 1));
            S.autotune_seconds = secs(ta, clk::now());
        }
#pragma omp parallel for schedule(dynamic)
        for (int m = 0; m < nmol; ++m) {
            if (!active[m]) continue;
            sol[m]->step(F.data() + moff[m], tf);
            if (sol[m]->converged()) {
                conv[m] = 1;
                active[m] = 0;
            } else if (sol[m]->iterations() >= so.max_iter) {
                active[m] = 0;
            }
        }
    }
    // final full FP64 build at every converged density (energy_final)
    std::vector<double> e_final(nmol);
    for (int m = 0; m < nmol; ++m) e_final[m] = sol[m]->energy_electronic();
    if (so.final_build) {
        for (int m = 0; m < nmol; ++m)
            std::copy(sol[m]->density().begin(), sol[m]->density().end(), D.begin() + moff[m]);
        check(hfg_set_precision(ctx, HFG_FP64));
        std::vector<double> Ff(nn);
        check(hfg_fock(ctx, H.data(), D.data(), 0, Ff.data()));
        for (int m = 0; m < nmol; ++m) {
            double e = 0.0;
            const size_t n2 = (size_t)nbf[m] * nbf[m];
            for (size_t k = 0; k < n2; ++k) e += D[moff[m] + k] * (H[moff[m] + k] + Ff[moff[m] + k]);
            e_final[m] = e;
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < nmol; ++m) sol[m]->finish(e_final[m]);
    S.scf_seconds = secs(t_scf, clk::now());
    hfg_stats hs{};
    hfg_last_stats(ctx, &hs);
    S.work_items = (int)hs.work_items;
    S.quartets = hs.quartets;
    for (int m = 0; m < nmol; ++m) {
        const hf_solver& s = *sol[m];
        if (res) {
            hfscf_scf_result& o = res[m];
            o.energy = s.energy_total();
            o.energy_elec = s.energy_electronic();
            o.energy_nuc = mols[m]->mol.nuclear_repulsion();
            o.iterations = s.iterations();
            o.converged = conv[m];
            o.fock_seconds = s.total_fock_seconds();
            o.last_fock_seconds = s.last_fock_seconds();
            o.fock_builds = s.iterations();
            o.energy_final = s.energy_final_total();
        }
        if (D_all) std::copy(s.density().begin(), s.density().end(), D_all + moff[m]);
        if (F_all) std::copy(s.fock().begin(), s.fock().end(), F_all + moff[m]);
    }
    if (grad_all) {
        const auto tg = clk::now();
        for (int m = 0; m < nmol; ++m)
            std::copy(sol[m]->density().begin(), sol[m]->density().end(), D.begin() + moff[m]);
        check(hfg_set_precision(ctx, opt.precision));
        int natot = 0;
        std::vector<int> g0(nmol);
        for (int m = 0; m < nmol; ++m) {
            g0[m] = natot;
            natot += 3 * mols[m]->mol.natom();
        }
        std::vector<double> g2(natot);
        // hfscf's gradient visits every shell quartet (hfscf_grad_ints.cpp:110-151): a second
        // batch context without the Schwarz screen
        hfg_ctx* gctx = ctx;
        std::unique_ptr<hfg_ctx, void (*)(hfg_ctx*)> gguard(nullptr, hfg_destroy);
        if (go.schwarz) {
            hfg_options gu = go;
            gu.schwarz = 0;
            gctx = nullptr;
            if ((r = hfg_create(&gctx, opt.device, &gu)))
                throw std::runtime_error("hfg_create failed (" + std::to_string(r) + "): no usable GPU");
            gguard.reset(gctx);
            for (int m = 0; m < nmol; ++m) {
                BasisArrays b(mols[m]->mol);
                check(hfg_batch_add_molecule(gctx, (int)b.L.size(), b.L.data(), b.atom.data(), b.np.data(),
                                             b.alpha.data(), b.coef.data(), mols[m]->mol.pure ? 1 : 0, b.co.data(),
                                             b.po.data(), mols[m]->mol.natom(), b.xyz.data()));
            }
            check(hfg_batch_commit(gctx));
            check(hfg_set_precision(gctx, opt.precision));
        }
        if (hfg_energy_backward(gctx, D.data(), g2.data()) < 0)
            throw std::runtime_error(std::string("hfgpu: ") + hfg_last_error(gctx));
#pragma omp parallel for schedule(dynamic)
        for (int m = 0; m < nmol; ++m) {
            const Molecule& mol = mols[m]->mol;
            const int n = nbf[m];
            const double* Dm = sol[m]->density().data();
            const double* Fm = sol[m]->fock().data();
            std::vector<double> DF((size_t)n * n), Q((size_t)n * n);
            gemm(n, Dm, false, Fm, false, DF.data());
            gemm(n, DF.data(), false, Dm, false, Q.data());
            double* g = grad_all + g0[m];
            std::copy(g2.begin() + g0[m], g2.begin() + g0[m] + 3 * mol.natom(), g);
            oneint_gradient(mol, Dm, Q.data(), g);
            nuclear_repulsion_gradient(mol, g);
        }
        S.gradient_seconds = secs(tg, clk::now());
    }
    if (st) *st = S;
    int nconv = 0;
    for (int m = 0; m < nmol; ++m) nconv += conv[m];
    return nconv == nmol ? 0 : -2;
}

}  // namespace

extern "C" {

void hfscf_default_scf_options(hfscf_scf_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->backend = HFSCF_BACKEND_GPU;
    o->device = 0;
    o->precision = 0;
    o->incremental = 0;
    o->diis_range = 6;
    o->max_iter = 100;
    o->schwarz = 1;
    o->energy_tol = 1e-10;
    o->rms_tol = 1e-10;
    o->integral_tol = 1e-14;
    o->final_fp64 = 1;
    o->density_tol = 1e-11;
    o->rebuild_every = 8;
    o->device_step = 1;
    o->polish_rms = 1e-6;
}

hfscf_mol* hfscf_mol_create(int natom, const int* Z, const double* xyz_bohr, const char* gbs_path, int pure_override) {
    try {
        std::vector<int> z(Z, Z + natom);
        std::vector<std::array<double, 3>> x(natom);
        for (int a = 0; a < natom; ++a) x[a] = {xyz_bohr[3 * a], xyz_bohr[3 * a + 1], xyz_bohr[3 * a + 2]};
        auto* m = new hfscf_mol();
        m->mol = build_molecule(z, x, read_gbs(gbs_path), pure_override);
        return m;
    } catch (const std::exception& e) {
        set_err(e.what());
        return nullptr;
    }
}

hfscf_mol* hfscf_mol_from_xyz(const char* xyz_path, const char* gbs_path, int pure_override) {
    try {
        std::vector<int> z;
        std::vector<std::array<double, 3>> x;
        read_xyz(xyz_path, z, x);
        auto* m = new hfscf_mol();
        m->mol = build_molecule(z, x, read_gbs(gbs_path), pure_override);
        return m;
    } catch (const std::exception& e) {
        set_err(e.what());
        return nullptr;
    }
}

void hfscf_mol_destroy(hfscf_mol* m) { delete m; }

int hfscf_mol_info(const hfscf_mol* m, int* natom, int* nshell, int* nbf, int* ncart, int* nelec, int* pure) {
    if (!m) return set_err("null molecule");
    if (natom) *natom = m->mol.natom();
    if (nshell) *nshell = (int)m->mol.shells.size();
    if (nbf) *nbf = m->mol.nbf();
    if (ncart) *ncart = m->mol.ncart_total;
    if (nelec) *nelec = m->mol.nelectrons();
    if (pure) *pure = m->mol.pure ? 1 : 0;
    return 0;
}

int hfscf_mol_nprim_total(const hfscf_mol* m) {
    int n = 0;
    for (auto& s : m->mol.shells) n += (int)s.alpha.size();
    return n;
}

int hfscf_mol_shells(const hfscf_mol* m, int* L, int* atom, int* nprim, int* cart_off, int* pure_off, double* alpha,
                     double* coef_norm, double* coef_raw) {
    if (!m) return set_err("null molecule");
    int k = 0;
    for (size_t s = 0; s < m->mol.shells.size(); ++s) {
        const Shell& sh = m->mol.shells[s];
        L[s] = sh.L;
        atom[s] = sh.atom;
        nprim[s] = (int)sh.alpha.size();
        cart_off[s] = sh.idx;
        pure_off[s] = sh.ids;
        for (size_t p = 0; p < sh.alpha.size(); ++p, ++k) {
            alpha[k] = sh.alpha[p];
            coef_norm[k] = sh.c[p];
            coef_raw[k] = sh.c_raw[p];
        }
    }
    return 0;
}

int hfscf_mol_geometry(const hfscf_mol* m, int* Z, double* xyz_bohr) {
    if (!m) return set_err("null molecule");
    for (int a = 0; a < m->mol.natom(); ++a) {
        if (Z) Z[a] = m->mol.Z[a];
        for (int d = 0; d < 3; ++d) xyz_bohr[3 * a + d] = m->mol.xyz[a][d];
    }
    return 0;
}

int hfscf_mol_set_geometry(hfscf_mol* m, const double* xyz_bohr) {
    if (!m) return set_err("null molecule");
    std::vector<std::array<double, 3>> x(m->mol.natom());
    for (int a = 0; a < m->mol.natom(); ++a) x[a] = {xyz_bohr[3 * a], xyz_bohr[3 * a + 1], xyz_bohr[3 * a + 2]};
    m->mol.set_geometry(x);
    return 0;
}

double hfscf_mol_enuc(const hfscf_mol* m) { return m ? m->mol.nuclear_repulsion() : 0.0; }

struct hfscf_input {
    hfscf::InputFile in;
    std::vector<std::pair<std::string, std::string>> kv;  // settings in keyword order
};

hfscf_input* hfscf_input_read(const char* path) {
    if (!path) {
        set_err("null path");
        return nullptr;
    }
    auto h = std::make_unique<hfscf_input>();
    std::string err;
    if (!hfscf::read_input(path, h->in, err)) {
        set_err(err);
        return nullptr;
    }
    for (const auto& e : h->in.settings) h->kv.emplace_back(e.first, e.second);
    return h.release();
}

void hfscf_input_destroy(hfscf_input* in) { delete in; }

int hfscf_input_info(const hfscf_input* in, int* natom, int* charge, int* multiplicity) {
    if (!in) return set_err("null input");
    if (natom) *natom = (int)in->in.Z.size();
    if (charge) *charge = in->in.charge;
    if (multiplicity) *multiplicity = in->in.multiplicity;
    return 0;
}

int hfscf_input_atoms(const hfscf_input* in, int* Z, double* xyz_bohr) {
    if (!in) return set_err("null input");
    if (Z) std::copy(in->in.Z.begin(), in->in.Z.end(), Z);
    if (xyz_bohr) std::copy(in->in.xyz.begin(), in->in.xyz.end(), xyz_bohr);
    return 0;
}

int hfscf_input_setting(const hfscf_input* in, int index, const char** keyword, const char** value) {
    if (!in) return set_err("null input");
    const int n = (int)in->kv.size();
    if (index >= 0 && index < n) {
        if (keyword) *keyword = in->kv[index].first.c_str();
        if (value) *value = in->kv[index].second.c_str();
    }
    return n;
}

const char* hfscf_input_basis_file(const hfscf_input* in) { return in ? in->in.basis_file.c_str() : ""; }
const char* hfscf_input_comment(const hfscf_input* in) { return in ? in->in.comment.c_str() : ""; }

int hfscf_oneint(const hfscf_mol* m, double* S, double* T, double* V) {
    if (!m) return set_err("null molecule");
    try {
        std::vector<double> s, t, v;
        overlap_kinetic(m->mol, s, t);
        nuclear_attraction(m->mol, v);
        if (S) std::memcpy(S, s.data(), s.size() * sizeof(double));
        if (T) std::memcpy(T, t.data(), t.size() * sizeof(double));
        if (V) std::memcpy(V, v.data(), v.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int hfscf_oneint_gradient(const hfscf_mol* m, const double* D, const double* Q, double* grad) {
    if (!m) return set_err("null molecule");
    std::memset(grad, 0, sizeof(double) * 3 * m->mol.natom());
    oneint_gradient(m->mol, D, Q, grad);
    nuclear_repulsion_gradient(m->mol, grad);
    return 0;
}

int hfscf_rhf_run(hfscf_mol* m, const hfscf_scf_options* o, const hfscf_callbacks* cb, hfscf_scf_result* res,
                  double* D, double* F, double* energies) {
    if (!m) return set_err("null molecule");
    try {
        hfscf_scf_options opt;
        if (o)
            opt = *o;
        else
            hfscf_default_scf_options(&opt);
        auto jk = make_backend(m->mol, opt, cb);
        ScfOptions so;
        so.diis_range = opt.diis_range;
        so.energy_tol = opt.energy_tol;
        so.rms_tol = opt.rms_tol;
        so.max_iter = opt.max_iter;
        so.incremental = opt.incremental != 0;
        so.final_build = opt.final_fp64 != 0;
        so.verbose = opt.verbose;
        so.rebuild_every = opt.rebuild_every;
        so.polish_rms = opt.polish_rms;
        so.mixed = opt.precision == HFG_MIXED;
        so.device_step = opt.device_step != 0;
        hf_solver solver(m->mol, so, jk.get());
        solver.init_data();
        bool conv = solver.scf_run();
        if (res) {
            res->energy = solver.energy_total();
            res->energy_elec = solver.energy_electronic();
            res->energy_nuc = m->mol.nuclear_repulsion();
            res->iterations = solver.iterations();
            res->converged = conv ? 1 : 0;
            res->fock_seconds = solver.total_fock_seconds();
            res->last_fock_seconds = solver.last_fock_seconds();
            res->fock_builds = solver.iterations();
            res->energy_final = solver.energy_final_total();
        }
        const size_t nn = (size_t)m->mol.nbf() * m->mol.nbf();
        if (D) std::memcpy(D, solver.density().data(), nn * sizeof(double));
        if (F) std::memcpy(F, solver.fock().data(), nn * sizeof(double));
        if (energies)
            for (size_t k = 0; k < solver.energies().size(); ++k) energies[k] = solver.energies()[k];
        return conv ? 0 : -2;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int hfscf_batch_rhf(int nmol, hfscf_mol* const* mols, const hfscf_scf_options* o, hfscf_scf_result* res,
                    double* D_all, double* F_all, double* grad_all, hfscf_batch_stats* stats) {
    if (nmol <= 0 || !mols) return set_err("hfscf_batch_rhf: no molecules");
    for (int m = 0; m < nmol; ++m)
        if (!mols[m]) return set_err("hfscf_batch_rhf: null molecule");
    try {
        hfscf_scf_options opt;
        if (o)
            opt = *o;
        else
            hfscf_default_scf_options(&opt);
        if (opt.backend != HFSCF_BACKEND_GPU) return set_err("hfscf_batch_rhf: batching needs the GPU backend");
        return batch_rhf(nmol, mols, opt, res, D_all, F_all, grad_all, stats);
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

int hfscf_rhf_gradient(hfscf_mol* m, const hfscf_scf_options* o, const hfscf_callbacks* cb, const double* D,
                       const double* F, double* grad) {
    if (!m) return set_err("null molecule");
    try {
        hfscf_scf_options opt;
        if (o)
            opt = *o;
        else
            hfscf_default_scf_options(&opt);
        // hfscf's gradient visits every shell quartet (hfscf_grad_ints.cpp:110-151): no Schwarz
        // screen for the two-electron part
        opt.schwarz = 0;
        auto jk = make_backend(m->mol, opt, cb);
        auto g = calc_scf_gradient_rhf(m->mol, D, F, jk.get());
        std::memcpy(grad, g.data(), g.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        return set_err(e.what());
    }
}

const char* hfscf_last_error(void) { return g_err.c_str(); }

}  // extern "C"
