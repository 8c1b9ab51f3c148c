/*
 * hfscf host layer — C ABI over the C++ RHF driver (basis, one-electron integrals, DIIS,
 * SCF loop and gradient assembly) whose two-electron work runs on the GPU through hfgpu.h.
 *
 * Mirrors hfscf's run-time seams:
 *   hfscf_mol_*        MOLEC::Molecule + BASIS::Basisset (hfscf: src/molecule/hfscf_molecule.cpp:21-336,
 *                      src/basis/hfscf_basis.cpp:19-275)
 *   hfscf_rhf_run      Mol::hf_solver::init_data + scf_run (hfscf: src/scfsolver/hfscf_hf_solver.cpp:79-476)
 *   hfscf_rhf_gradient Mol::scf_gradient::calc_scf_gradient_rhf (hfscf: src/gradient/hfscf_grad_rhf.cpp:8-178)
 * Status 0 = success, negative = error (message from hfscf_last_error()).
 */
#ifndef HFSCF_H
#define HFSCF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfscf_mol hfscf_mol;

/* Two-electron backend. backend = HFSCF_BACKEND_GPU uses libhfgpu; HFSCF_BACKEND_CALLBACK uses
 * the user functions below (the run-time seam an external Fock builder plugs into). */
enum { HFSCF_BACKEND_GPU = 0, HFSCF_BACKEND_CALLBACK = 1 };

typedef struct hfscf_callbacks {
    void* user;
    /* J and K (nbf x nbf) of half density D */
    int (*build_jk)(void* user, const double* D, double* J, double* K);
    /* natom x 3 two-electron gradient, overwritten */
    int (*grad2e)(void* user, const double* D, double* grad);
} hfscf_callbacks;

typedef struct hfscf_scf_options {
    int backend;          /* HFSCF_BACKEND_* */
    int device;           /* HIP device for the GPU backend */
    int precision;        /* hfg_precision; HFG_MIXED: first build FP64, incremental builds FP32,
                             final build FP64 */
    int incremental;      /* delta-D Fock builds */
    int diis_range;       /* hfscf diis_size (vectors kept = range + 1) */
    int max_iter;
    int schwarz;          /* Schwarz screening on/off */
    int verbose;
    double energy_tol;
    double rms_tol;
    double integral_tol;
    int final_fp64;       /* one more full FP64 Fock build at the converged density (energy_final) */
    double density_tol;   /* density-weighted quartet screening of incremental builds (hfg_options) */
    int graph;            /* replay each J/K build configuration from a captured HIP graph */
    int autotune;         /* batched runs: autotune the class launch shapes after the first build */
    int rebuild_every;    /* incremental builds: every rebuild_every-th build is a full one, which clears
                             what density screening accumulates along the chain (0 = never; default 8) */
    int device_step;      /* GPU backend, solo runs: 1 = the whole SCF step on the GPU (DIIS, F' = X^T F X,
                             diagonalisation, density and energy in rocBLAS / rocSOLVER, the Fock build on
                             device memory); 0 = the host loop (default 1) */
    double polish_rms;    /* mixed precision (solo runs): once the DIIS rms is below this, continue from a
                             full FP64 build with a fresh DIIS subspace and converge in FP64, so D, F,
                             the energy and the gradient are those of the FP64 SCF (0 = mixed to the
                             end; default 1e-6) */
} hfscf_scf_options;

typedef struct hfscf_scf_result {
    double energy;        /* total RHF energy (Eh) */
    double energy_elec;
    double energy_nuc;
    int iterations;
    int converged;
    double fock_seconds;  /* wall time inside Fock builds */
    double last_fock_seconds;
    int fock_builds;
    double energy_final;  /* sum D (H + F(D)) + E_nuc from a full-precision build at the final D */
} hfscf_scf_result;

void hfscf_default_scf_options(hfscf_scf_options* o);

/* Molecule from atomic numbers and bohr coordinates, with a Psi4 .gbs basis file.
 * pure_override: -1 basis-file header, 0 Cartesian, 1 spherical. */
hfscf_mol* hfscf_mol_create(int natom, const int* Z, const double* xyz_bohr, const char* gbs_path, int pure_override);
/* Same from an .xyz file in angstrom. */
hfscf_mol* hfscf_mol_from_xyz(const char* xyz_path, const char* gbs_path, int pure_override);
void hfscf_mol_destroy(hfscf_mol* m);
int hfscf_mol_info(const hfscf_mol* m, int* natom, int* nshell, int* nbf, int* ncart, int* nelec, int* pure);
/* Shell table (arrays sized by nshell, primitive arrays by the total primitive count). */
int hfscf_mol_nprim_total(const hfscf_mol* m);
int hfscf_mol_shells(const hfscf_mol* m, int* L, int* atom, int* nprim, int* cart_off, int* pure_off, double* alpha,
                     double* coef_norm, double* coef_raw);
int hfscf_mol_geometry(const hfscf_mol* m, int* Z, double* xyz_bohr);
int hfscf_mol_set_geometry(hfscf_mol* m, const double* xyz_bohr);
double hfscf_mol_enuc(const hfscf_mol* m);

/* One-electron matrices (nbf x nbf). */
int hfscf_oneint(const hfscf_mol* m, double* S, double* T, double* V);
/* 2 D.d(T+V) - 2 Q.dS + nuclear repulsion gradient (natom x 3, overwritten). */
int hfscf_oneint_gradient(const hfscf_mol* m, const double* D, const double* Q, double* grad);

/* RHF SCF (CORE guess). D and F (nbf x nbf) receive the converged half density and the last
 * extrapolated Fock matrix; either may be NULL. energies (max_iter) may be NULL. */
int hfscf_rhf_run(hfscf_mol* m, const hfscf_scf_options* o, const hfscf_callbacks* cb, hfscf_scf_result* res,
                  double* D, double* F, double* energies);

/* RHF gradient (natom x 3) from converged D and F. */
int hfscf_rhf_gradient(hfscf_mol* m, const hfscf_scf_options* o, const hfscf_callbacks* cb, const double* D,
                       const double* F, double* grad);

/* Batching mode (GPU backend): the SCF of every molecule runs in lockstep with one batched
 * GPU Fock build per iteration for all molecules (hfgpu batching mode); a molecule that has
 * converged keeps its density, so its incremental builds have nothing left to add. Each
 * molecule follows exactly the iteration of hfscf_rhf_run with the same options.
 * res[nmol] receives per-molecule results. D_all / F_all (concatenated n_m x n_m, may be NULL)
 * receive densities and last Fock matrices; grad_all (concatenated natom_m x 3, may be NULL)
 * the RHF gradients from one batched adjoint launch. No hfscf counterpart (hfscf runs one
 * molecule per process). */
typedef struct hfscf_batch_stats {
    double setup_seconds;     /* context, pair data, screening, work decomposition */
    double scf_seconds;       /* SCF loop including the final builds */
    double fock_seconds;      /* inside batched Fock builds */
    double gradient_seconds;  /* batched gradient (two-electron launch + host one-electron terms) */
    double autotune_seconds;
    int iterations;           /* batched builds in the SCF loop */
    int work_items;
    long long quartets;       /* surviving shell quartets of one full build of all molecules */
} hfscf_batch_stats;
int hfscf_batch_rhf(int nmol, hfscf_mol* const* mols, const hfscf_scf_options* o, hfscf_scf_result* res,
                    double* D_all, double* F_all, double* grad_all, hfscf_batch_stats* stats);

/* hfscf .dat input files: a comment line starting with '#', "keyword = value" lines checked
 * against hfscf's keyword table (names, integer / real / free-text kinds, accepted values), and a
 * { } block of atoms "symbol-or-Z x y z" in bohr, or angstrom with units = angstrom (dummy atoms X
 * dropped, at most 100 lines). charge and multiplicity are read as hfscf reads them. A z-matrix
 * block is refused. NULL on an error, with hfscf's message in hfscf_last_error().
 * Replaces: MOLEC::Molecule::init_molecule and MOLPARSE (hfscf: src/molecule/hfscf_molecule.cpp:21-200,
 * src/molecule/hfscf_parser.hpp:48-151, src/molecule/hfscf_keywords.hpp:14-72). */
typedef struct hfscf_input hfscf_input;
hfscf_input* hfscf_input_read(const char* path);
void hfscf_input_destroy(hfscf_input* in);
int hfscf_input_info(const hfscf_input* in, int* natom, int* charge, int* multiplicity);
int hfscf_input_atoms(const hfscf_input* in, int* Z, double* xyz_bohr);
/* Number of keywords given; with 0 <= index < that number, *keyword and *value receive the
 * index-th (alphabetical order; the strings live as long as the input). */
int hfscf_input_setting(const hfscf_input* in, int index, const char** keyword, const char** value);
/* basis_set as hfscf's file name: lowercase, '*' -> "_st_", ".gbs" (hfscf_parser.hpp:134-151); "" if absent. */
const char* hfscf_input_basis_file(const hfscf_input* in);
const char* hfscf_input_comment(const hfscf_input* in);

const char* hfscf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HFSCF_H */
