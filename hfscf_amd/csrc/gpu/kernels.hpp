// Kernel declarations and per-class launch dispatch.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace hfg {

struct JkArgs {
    const WorkItem* items;
    int nitems;
    const PairDev* pairs;
    const double* psmean;  // mean Schwarz value per pair (parallel to pairs)
    const int* bp_off;     // block pair -> first pair
    const int* bp_cnt;
    const int* blk_nf;     // functions per block
    const double4* pp0;
    const double4* pp1;
    const double4* pp2;
    const double* boys;
    const double* D;       // concatenated n x n densities (MolDev::moff, MolDev::n)
    const MolDev* mols;
    const double* psq;  // sqrt|(pq|pq)| per pair component (element screen)
    double elem_tol;    // element screen threshold (0 = off)
    double tol;        // shell-quartet Schwarz threshold on mean values
    double prim_cut;   // primitive-quartet prefactor cutoff
    double* slab;      // FP64 tiles per work item
    // density-weighted screening (incremental builds): a quartet is skipped when
    // sqrt(ab|ab) sqrt(cd|cd) max(4|D_ab|, 4|D_cd|, |D_ac|, |D_ad|, |D_bc|, |D_bd|) < dtol,
    // with |D_xy| the largest density element of shell block (x, y); dtol = 0 disables it
    const float* dmax;  // per molecule nshell x nshell at MolDev::dm_off
    double dtol;
    int* queue;         // work-item counter of this launch (zeroed before the launch)
    // ERI export (hfg_eri_vector): when set, every quartet's output-basis block is written to
    // this packed vector at hfscf's canonical index instead of being digested (no tiles)
    double* eri_out;
};

struct GatherArgs {
    const int4* tiles;     // (X, Y, molecule, 0) per output tile, X >= Y
    const int* gJ_off;     // CSR offsets per tile
    const GatherEntry* gJ;
    const int* gK_off;
    const GatherEntry* gK;
    const double* slab;
    const int* blk_nf;
    const int* blk_func_off;
    const int* blk_func_ao;
    const MolDev* mols;
    double* J;
    double* K;
};

struct ScreenArgs {  // Cartesian (ab|ab) diagonal per pair
    const PairDev* pairs;
    const long long* out_off;  // per pair: offset of its ncart(a) x ncart(b) diagonal in out
    int count;
    const double4* pp0;
    const double4* pp1;
    const double4* pp2;
    const double* boys;
    double* out;
};

// Launch shape of one class launch: persistent workgroups (0 = one per work item) and threads
// per workgroup (<= 256).
struct LaunchShape {
    int grid;
    int block;
};

// Launchers (one specialised kernel per class; la >= lb, la >= lc >= ld). mixed != 0 selects
// the FP32 evaluation and digestion of the class (FP64 Boys values and prefactors except in the
// classes with L <= 2, FP64 tiles and sums).
hipError_t launch_jk(int la, int lb, int lc, int ld, bool pure, int mixed, const JkArgs& a, int nitems,
                     LaunchShape sh, hipStream_t s);
hipError_t launch_screen(int la, int lb, const ScreenArgs& a, hipStream_t s);
// Sets n ints to zero (a kernel, so it may sit in a captured graph).
hipError_t launch_zero_ints(int* p, int n, hipStream_t s);
hipError_t launch_fock(const double* base, const double* J, const double* K, double* F, double* Facc, int nn,
                       int accumulate, hipStream_t s);

}  // namespace hfg
