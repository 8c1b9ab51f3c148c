// Device data layout of the two-electron engine (shared by host set-up and kernels).
//
// Shells are grouped into "blocks": shells of one angular momentum, sorted by primitive count,
// at most kBlockFuncs output functions per block. A block pair (X >= Y) owns the list of
// significant shell pairs (a in X, b in Y), sorted by descending mean Schwarz value, and a
// work item is a contiguous row range of one bra block pair against one ket block pair.
// Every shell quartet of a work item adds into six small J/K tiles (blocks A x B, C x D,
// A x C, A x D, B x C, B x D) held in LDS; the tiles are flushed to the work item's private
// slab in HBM and summed per output tile by a gather kernel, so no global atomics are used.
#pragma once
#include <hip/hip_runtime.h>

namespace hfg {

constexpr int kBlockFuncs = 16;   // output functions per shell block
constexpr int kMaxRows = 256;     // bra rows per work item
constexpr int kRegions = 6;       // Jab Jcd Kac Kad Kbc Kbd
constexpr int kMaxAtomsPerMolecule = 128;  // gradient partials of one molecule live in LDS

// One shell pair; primitive pairs live in pp0/pp1/pp2 at [off, off + np).
struct PairDev {
    int sa, sb;        // shell indices within the molecule (input order)
    int off, np;       // primitive-pair range
    int slot_a, slot_b;  // positions of sa / sb inside their blocks
    int ao_a, ao_b;    // first output-basis function of sa / sb
    int sq_off;        // offset of sqrt|(pq|pq)| per Cartesian component pair (p in sa, q in sb)
    int pstride;       // primitive pair k of this shell pair is at off + k * pstride
    double sq_min;     // smallest of those values
    double sq_max;     // largest of those values (density-weighted screening bound)
    double AB[3];      // A - B
};

// Primitive-pair data (three double4 per primitive pair), stored primitive-major within each
// block pair's sorted list (every pair of a block pair has the same primitive count), so that
// the lanes of a wavefront, which hold consecutive pairs of a list, read consecutive records:
//   pp0 = {p, Px, Py, Pz}
//   pp1 = {PAx, PAy, PAz, K} with K = exp(-ab/p |AB|^2) c_a c_b sqrt(2 pi^2.5) / p
//   pp2 = {1/(2p), alpha_a, alpha_b, 0}
// so that the quartet prefactor is K_ab K_cd / sqrt(p + q).

struct WorkItem {
    int P, Q;          // bra / ket block pair
    int r0, r1;        // bra rows [r0, r1) of P's pair list
    int c0, c1;        // ket columns [c0, c1) of each row (clipped to the row's surviving range)
    int A, B, C, D;    // blocks
    int mol;           // molecule (a context may hold many independent molecules)
    int split;         // lanes per quartet: the shell-group kernel splits each group quartet's bra
                       // primitive pairs over this many lanes (1 in every other kernel)
    long long slab;    // offset (doubles) of this item's six tiles
    long long gslab;   // offset (doubles) of this item's gradient partials (3 per atom of mol)
};

// Per-molecule tables of a context holding several molecules (one entry for a single molecule).
struct MolDev {
    long long moff;  // offset of the molecule's n x n matrices in the concatenated D/J/K/F
    int n;           // basis functions
    int natom;
    int shell0;      // first shell (global index) of the molecule
    int nshell;
    int dm_off;      // offset of its nshell x nshell block-maximum density table
    int pad_;
};

// Largest scratch (private segment) one dispatch may need; launches of spill-heavy class
// kernels are given fewer workgroups (each loops over several work items) to stay below it.
constexpr double kScratchBudget = 1024.0 * 1024.0 * 1024.0;
// Bytes per lane that registers hold (256 VGPRs + 256 AGPRs of 4 bytes, less working space).
constexpr size_t kRegisterBytes = 1536;

// Gather list entry for one output tile: a tile of `rows x cols` at `off` (row-major),
// read transposed when trans != 0.
struct GatherEntry {
    long long off;
    int cols;
    int trans;
};

}  // namespace hfg
