/*
 * hfgpu — C ABI of the MI355X (gfx950) RHF two-electron engine.
 *
 * The library replaces hfscf's two-electron hot path behind three C++ seams:
 *   - ERIOS::Erios, the shell-quartet ERI engine with Schwarz screening
 *     (hfscf: src/integrals/hfscf_oseri.hpp:27-123, hfscf_oseri.cpp:22-246,
 *      hfscf_oseri_screen.cpp:17-181),
 *   - Mol::hf_solver::update_fock_matrix / update_fock_matrix_scf_direct, the J/K Fock build
 *     (hfscf: src/scfsolver/hfscf_hf_solver.hpp:105-107, hfscf_hf_solver.cpp:254-373),
 *   - the two-electron part of Mol::scf_gradient::calc_scf_gradient_rhf
 *     (hfscf: src/gradient/hfscf_grad_rhf.cpp:76-141, hfscf_grad_ints.cpp:79-151).
 *
 * Conventions
 *   - All host pointers are caller-owned and row-major; calls are synchronous on return.
 *   - D is hfscf's half density C_occ C_occ^T; F = H + sum_kl D_kl (2 (ij|kl) - (ik|jl)),
 *     J_ij = sum_kl D_kl (ij|kl), K_ij = sum_kl D_kl (ik|jl) (hfscf_hf_solver.cpp:221-231,261-278).
 *   - Matrices use the output basis: spherical (m = -l..l) when pure != 0, else Cartesian,
 *     functions ordered as hfscf orders them (shell order, Cartesian lx-major order).
 *   - Status 0 = success, negative = error; hfg_last_error() gives the message. The library
 *     never exits the process and never falls back to a CPU path.
 *   - A context owns its main stream, its side streams (six by default) and all its device
 *     memory; it is not thread-safe. The class kernels that spill (private segments of several
 *     KB per lane) run on one more stream that every context on a device shares, one launch at
 *     a time, so that a single queue holds their scratch (hfg_eri_vector and hfg_hessian_2e
 *     launch there too); contexts used from several threads on
 *     one device therefore serialise those launches, and a HIP graph (hfg_options.graph) never
 *     captures them (they are launched directly at each replay).
 */
#ifndef HFGPU_H
#define HFGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfg_ctx hfg_ctx;

enum hfg_status {
    HFG_OK = 0,
    HFG_ERR_ARG = -1,      /* invalid argument or call order */
    HFG_ERR_HIP = -2,      /* HIP runtime error (message in hfg_last_error) */
    HFG_ERR_NOMEM = -3,    /* device allocation failed */
    HFG_ERR_UNSUPPORTED = -4 /* angular momentum above d, or no GPU */
};

enum hfg_precision {
    HFG_FP64 = 0,  /* every ERI and digestion in FP64 */
    HFG_MIXED = 1  /* FP32 recursions/digestion (and the whole primitive loop up to L = 3), FP64
                      Boys values above L = 3, Schwarz, tile sums and Fock sum */
};

typedef struct hfg_options {
    int precision;        /* hfg_precision */
    double integral_tol;  /* hfscf integral_tol (default 1e-14): Schwarz threshold is 100x this,
                             primitive cutoff integral_tol/1e4 (hfscf_oseri.hpp:31-32, hfscf_oseri.cpp:73) */
    int schwarz;          /* 1 = screen shell quartets on the mean Schwarz bound (hfscf eri_screen) */
    double density_tol;   /* incremental builds only: skip a quartet when its density-weighted Schwarz
                             bound sqrt(ab|ab) sqrt(cd|cd) max(4|dD_ab|, 4|dD_cd|, |dD_ac|, |dD_ad|,
                             |dD_bc|, |dD_bd|) is below this (the bound of hfscf's direct digest,
                             hfscf_hf_solver.cpp:317-324); 0 = off (default) */
    int graph;            /* 1 = capture each J/K build configuration once as a HIP graph and replay it */
    int work_split;       /* work items per class launch relative to the default split (1, 2, 4, ...;
                             0 = the default). More items keep more wavefronts in flight at the cost of
                             more tiles to gather. */
    int coop_min_l;       /* classes with la+lb+lc+ld >= coop_min_l run on the wavefront-cooperative
                             kernels (one workgroup per quartet, recursion tables in LDS), the others
                             on the per-lane register kernels; 4..7, 0 = 7 (the default). From L = 4 up,
                             hfg_autotune keeps whichever form it measures faster for each class. */
    int side_streams;     /* streams the class launches of one build are spread over (1..8, 0 = 6) */
    int group_shells;     /* 0 or 1 (the default): s shells of one atom with identical exponents (the
                             contracted s shells of cc-pVXZ, which share one set of primitives) form
                             groups, and the J/K (ss|ss) class evaluates each primitive quartet of a
                             group quartet once for all of its member shell quartets; -1 = off. The
                             results are those of the per-shell kernel to rounding. */
    int reserved[1];
} hfg_options;

/* Fill opt with defaults (FP64, integral_tol 1e-14, screening on, no density screening). */
void hfg_default_options(hfg_options* opt);

/* Create a context on HIP device `device`. Replaces: ERIOS::Erios::Erios(is_pure, cut_off). */
int hfg_create(hfg_ctx** out, int device, const hfg_options* opt);

/* Basis description, shells in hfscf order (hfscf: src/basis/hfscf_basis.cpp:245-275).
 * L[s] <= 2; nprim[s] primitives; alpha / coef concatenated over shells, coef normalised as
 * BASIS::Shell does (hfscf: src/integrals/hfscf_shell.cpp:44-54); cart_offset / pure_offset as
 * Shell::get_idx / get_ids. atom_of_shell indexes the geometry passed to hfg_set_geometry. */
int hfg_set_basis(hfg_ctx* ctx, int nshell, const int* L, const int* atom_of_shell, const int* nprim,
                  const double* alpha, const double* coef, int pure, const int* cart_offset,
                  const int* pure_offset);

/* Atom positions in bohr (natom x 3) and charges. Recomputes the primitive-pair data, the
 * Schwarz diagonal (ab|ab) on the GPU, the pair lists and the work decomposition.
 * Replaces: ShellPair::set_params + Erios::init_screen/compute_screen_matrix
 * (hfscf_shellpair.cpp:33-60, hfscf_oseri_screen.cpp:17-181). */
int hfg_set_geometry(hfg_ctx* ctx, int natom, const double* xyz_bohr, const int* Z);

/* J and K (nbf x nbf) for the half density D. Either output may be NULL. flags reserved (0). */
int hfg_build_jk(hfg_ctx* ctx, const double* D, double* J, double* K, int flags);

/* hfg_build_jk on device memory without any allocation. No counterpart in hfscf: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). D, J and K are device pointers (nbf x nbf, for a batch the
 * concatenated matrices) on the context's device; the kernels' queue counters and tile slab live
 * in a fixed 32 MiB array per GPU, so a build whose tiles need more is refused (HFG_ERR_ARG,
 * the message gives the size; larger molecules take hfg_build_jk). Every call
 * on a GPU shares that array: calls that overlap on one GPU (threads, contexts) race. The caller
 * makes D ready before the call; the call is synchronous on return. */
int hfg_build_jk_noalloc(hfg_ctx* ctx, const double* D, double* J, double* K);

/* Fock matrix F = H + 2J - K. When incremental != 0 and a previous build exists, J and K are
 * built from D - D_prev and added to the previous two-electron matrix (the FP32 error of the
 * mixed mode then shrinks with the density change). Replaces: hf_solver::update_fock_matrix's
 * digestion (hfscf_hf_solver.cpp:261-278). */
int hfg_fock(hfg_ctx* ctx, const double* H, const double* D, int incremental, double* F);

/* hfg_fock on device memory: H, D and F are device pointers (n x n on the context's device,
 * for a batch the concatenated matrices), so a device-resident SCF step (DIIS, diagonalisation and
 * density on the GPU) hands its density to the build and takes the Fock matrix back without
 * host copies. Synchronous on return like every other call; incremental builds work as in
 * hfg_fock, and the two calls may be mixed. Replaces the same digestion as hfg_fock. */
int hfg_fock_device(hfg_ctx* ctx, const double* H, const double* D, int incremental, double* F);

/* Forget the previous density (next hfg_fock is a full build). */
int hfg_reset_incremental(hfg_ctx* ctx);

/* Precision of the following builds and gradients (hfg_precision). A mixed-precision SCF runs
 * its first (full) build in FP64, the incremental builds in HFG_MIXED and a final FP64 build
 * on the converged density; the previous Fock matrix is kept across the switch. */
int hfg_set_precision(hfg_ctx* ctx, int precision);

/* Density threshold of the following incremental builds (hfg_options.density_tol; >= 0, 0 turns
 * density screening off). No counterpart in hfscf, which has no incremental builds: the SCF
 * driver tightens it for builds whose errors must stay below the energy tolerance. */
int hfg_set_density_tol(hfg_ctx* ctx, double density_tol);

/* E_2e = sum_ij D_ij (2J - K)_ij for the half density D. */
int hfg_energy(hfg_ctx* ctx, const double* D, double* E2);

/* Backward of hfg_energy: two-electron gradient sum_pqrs [2 D_pq D_rs - D_pr D_qs] (pq|rs)'
 * (natom x 3, overwritten), contracted by the adjoint derivative-ERI kernel without storing
 * derivative integrals. Replaces: calc_coulomb_exchange_integral_centers + the dGdX loop of
 * calc_scf_gradient_rhf (hfscf_grad_ints.cpp:110-151, hfscf_grad_rhf.cpp:121-141). */
int hfg_energy_backward(hfg_ctx* ctx, const double* D, double* grad);

/* Test hook: output-basis block (s1 s2|s3 s4), laid out [a][b][c][d], read out of the production
 * J/K kernels with unit densities, so it follows the context's screening (create the context with
 * schwarz = 0 for the unscreened block) and costs n3 * n4 full J/K builds.
 * Replaces: Erios::compute_contracted_shell_quartet for one quartet (hfscf_oseri.cpp:22-246). */
int hfg_eri_shell_quartet(hfg_ctx* ctx, int s1, int s2, int s3, int s4, double* block);

/* Cartesian Schwarz diagonal (pq|pq), ncart x ncart (hfscf's Erios::screen). */
int hfg_schwarz(hfg_ctx* ctx, double* screen);

/* Test hook: F_0..F_m(T) (n x (m+1), row-major) for n values T >= 0, evaluated on the device by
 * the functions the class kernels use: the tabulated Taylor step with down/upward recursion
 * (closed_form = 0, m <= 17) or the closed forms of F_0 and F_1 (closed_form = 1, m <= 1).
 * Replaces: extramath::F_nu and the downward recursion at its call site (hfscf: src/math/gamma.cpp:18-24,
 * src/integrals/hfscf_oseri.cpp:103-118), without hfscf's T >= 1e-8 floor, which the kernels apply
 * where hfscf does (L = 0 quartets). */
int hfg_boys_probe(hfg_ctx* ctx, int m, int n, const double* T, double* F, int closed_form);

/* Measurement hook, no hfscf counterpart: `reps` launches of a plain copy kernel of `bytes` bytes
 * (reads and writes `bytes` each) on `device`, average milliseconds per copy in *ms_per_copy.
 * A kernel of known traffic against which a profiler's HBM byte counters are scaled. */
int hfg_copy_calibration(int device, long long bytes, int reps, double* ms_per_copy);

/* ---- Batching mode: many independent molecules in one context ----------------------------
 * Molecules are added one by one (basis as for hfg_set_basis, offsets within the molecule, and
 * the molecule's own geometry), then committed. Every shell quartet of every molecule then
 * goes through one persistent launch per angular-momentum class: work items of all molecules
 * of that class sit in one device queue. hfg_build_jk, hfg_fock, hfg_energy and
 * hfg_energy_backward take and return the per-molecule matrices concatenated in molecule order
 * (n_m x n_m each, offsets from hfg_batch_layout) and gradients concatenated (natom_m x 3).
 * A single-molecule context is the case of one molecule. No hfscf counterpart: hfscf solves
 * one molecule per process. */

/* Returns the molecule index (>= 0) or a negative status. natom <= 128 per molecule. */
int hfg_batch_add_molecule(hfg_ctx* ctx, int nshell, const int* L, const int* atom_of_shell, const int* nprim,
                           const double* alpha, const double* coef, int pure, const int* cart_offset,
                           const int* pure_offset, int natom, const double* xyz_bohr);
/* Pair data, Schwarz screening and work decomposition of every added molecule. */
int hfg_batch_commit(hfg_ctx* ctx);
/* Molecule count, and per molecule its basis size, matrix offset and atom count (any may be NULL). */
int hfg_batch_layout(const hfg_ctx* ctx, int* nmol, int* nbf, long long* moff, int* natom);

/* One-electron integrals of the context's single molecule on the GPU: S, T and V (nbf x nbf
 * row-major host arrays in the output basis; any may be NULL), with the nuclear charges given to
 * hfg_set_geometry. Replaces: OSOverlap / OSNuclear compute_contracted_shell as called by
 * hf_solver::create_one_electron_hamiltonian (hfscf: src/scfsolver/hfscf_hf_hamiltonian.cpp:14-38;
 * src/integrals/hfscf_osoverlap.cpp:10, hfscf_osnuclear.cpp:17-83). */
int hfg_oneint(hfg_ctx* ctx, double* S, double* T, double* V);
/* grad (natom x 3, host, added into) += 2 sum_ij D_ij d(T+V)_ij/dX - 2 sum_ij Q_ij dS_ij/dX on the
 * GPU, D the half density and Q = D F D (both symmetric, as in RHF). Replaces: the dTdX, dVndX and dSdX terms of
 * calc_scf_gradient_rhf (hfscf: src/gradient/hfscf_grad_rhf.cpp:143-167; hfscf_osoverlapderiv1.cpp:8,
 * hfscf_oskineticderiv1.cpp:8, hfscf_osnuclearderiv1.cpp:15). */
int hfg_oneint_gradient(hfg_ctx* ctx, const double* D, const double* Q, double* grad);

/* The packed two-electron integral vector of the context's single molecule on the GPU: out[count]
 * receives (ij|kl) at hfscf's canonical index index_ijkl = t(t(i,j), t(k,l)), t(x,y) =
 * max(x,y)(max(x,y)+1)/2 + min(x,y), count = npair (npair + 1) / 2 with npair = nbf (nbf + 1) / 2,
 * in the output basis, with the context's screening (quartets and elements hfscf's screen drops
 * stay 0; its Cartesian diagonal), as hfscf stores them for post-SCF methods and the AO -> MO
 * transform. The J/K kernels compute it (FP64) and write each element instead of digesting it.
 * Replaces: the stored repulsion vector of hf_solver (hfscf: src/hfscf/hfscf.cpp:76,
 * src/basis/hfscf_trans_basis.cpp:7, src/math/hfscf_math.hpp:90; built by
 * src/integrals/hfscf_oseri.cpp:22-246). */
int hfg_eri_vector(hfg_ctx* ctx, double* out, long long count);

/* Two-electron part of the RHF Hessian of the context's single molecule at a fixed density, on
 * the GPU: HC = sum D_ij D_kl 2 d2(ij|kl)/dX1 dX2 and HX = -sum D_ij D_kl d2(ik|jl)/dX1 dX2, each
 * (3 natom) x (3 natom) row-major over (atom, x y z), overwritten. D: nbf x nbf (hfscf's d_mat, the
 * output basis). Second-derivative integrals of every shell quartet (no Schwarz screen; primitive
 * quartets below integral_tol / 1e4 skipped), contracted as they are made; shells up to d.
 * Replaces: Mol::scf_hessian::calc_scf_hessian_tei, its m_hessianC and m_hessianEx (hfscf:
 * src/hessian/hfscf_hessian_ints.cpp:20-313; integrals src/integrals/hfscf_oserideriv2qq.cpp:28-749,
 * hfscf_oserideriv2pq.cpp:23-799). */
int hfg_hessian_2e(hfg_ctx* ctx, const double* D, double* HC, double* HX);

/* Autotuner: times every class launch alone over candidate launch shapes (threads per
 * workgroup x persistent workgroups) on the density of the most recent build and keeps the
 * fastest per class. what: 1 = J/K, 2 = gradient, 3 = both. */
int hfg_autotune(hfg_ctx* ctx, int what);
/* Current launch shape per class (la*1000+lb*100+lc*10+ld, grid, block for J/K and gradient;
 * grid 0 = one workgroup per work item, block 0 = the default, block -1 = the class runs on the
 * cooperative kernel, whose workgroup size is fixed, J/K block -2 = the (ss|ss) class runs on the
 * shell-group kernel, hfg_options.group_shells). Returns the number of classes. */
int hfg_launch_shapes(const hfg_ctx* ctx, int cap, int* la_lb_lc_ld, int* jk_grid, int* jk_block, int* grad_grid,
                      int* grad_block);

/* Per class of the last J/K build with direct launches (graph replays are not bracketed): class
 * la*1000+lb*100+lc*10+ld, surviving quartets, model FLOPs and the class launch's duration
 * (HIP events on its stream; -1 when not timed). Returns the number of classes. */
int hfg_class_stats(hfg_ctx* ctx, int cap, int* la_lb_lc_ld, long long* quartets, double* model_flops,
                    double* kernel_ms);

/* Statistics of the last hfg_build_jk / hfg_fock: surviving shell quartets, model FLOPs
 * (the algorithmic FLOP model over surviving quartets), J/K kernel seconds (HIP events). */
typedef struct hfg_stats {
    long long quartets;
    double model_flops;
    double kernel_seconds;
    double total_seconds;
    long long work_items;
    long long slab_bytes;
} hfg_stats;
int hfg_last_stats(const hfg_ctx* ctx, hfg_stats* st);

int hfg_nbf(const hfg_ctx* ctx);
const char* hfg_last_error(const hfg_ctx* ctx);
void hfg_destroy(hfg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* HFGPU_H */
