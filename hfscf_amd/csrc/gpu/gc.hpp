// Shells that share exponents: pair groups and the J/K kernels that evaluate each group
// quartet's primitive quartets once for every member shell quartet.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "layout.hpp"

namespace hfg {

// Most s shells of one atom with identical exponents that form one group (the shells of higher
// angular momentum stay single-shell groups).
constexpr int kGcMaxShells = 2;
constexpr int kGcMaxMembers = kGcMaxShells * kGcMaxShells;
// Most member shell pairs of a group pair of blocks with angular momenta (l1, l2).
__host__ __device__ constexpr int gc_members(int l1, int l2) {
    return (l1 == 0 ? kGcMaxShells : 1) * (l2 == 0 ? kGcMaxShells : 1);
}
// Lanes per group quartet (WorkItem::split): enough that one lane evaluates about kGcLanePrims
// primitive quartets, at most kGcMaxSplit.
constexpr double kGcLanePrims = 512.0;
constexpr int kGcMaxSplit = 16;

// A pair of shell groups (a group of a block, a group of a block): up to four member shell
// pairs (indices into the per-shell pair array, -1 where the pair did not survive the screen),
// the group's primitive pairs (shared exponents and centres; one K per member in gK) and A - B.
// Primitive pair k of a group pair is at off + k * pstride: the group pairs of one block pair's
// list have the same primitive count and are stored primitive-major, so the lanes of a wavefront,
// which hold consecutive group pairs of a list, read consecutive records; the np significant
// primitive pairs come first.
struct GcPairDev {
    int nmem;
    int mem[kGcMaxMembers];
    int off, np;
    int pstride;
    double AB[3];
};

// Device tables of the group lists, beside JkArgs (whose items index them for the classes run
// on the group kernel).
struct GcArgs {
    const GcPairDev* gpairs;  // the group lists of every block pair
    const double* gkey;       // sort key of each group (the largest of its members' keys)
    const int* goff;          // first group of each block pair's list
    const int* gcnt;          // groups in each block pair's list
    const double4* gp0;       // per group primitive pair: {p, P}
    const double4* gp1;       // per group primitive pair: {P - A, 1/(2p)}
    const double4* gK;        // per group primitive pair: K of members 0..3
};

// Whether class (la lb|lc ld) has a kernel on the group lists for this basis type.
bool jk_gc_available(int la, int lb, int lc, int ld, bool pure);

// The J/K launch of a class on the group lists (mixed != 0: the FP32 primitive loops of the mixed
// builds); the class must be jk_gc_available.
hipError_t launch_jk_gc(int la, int lb, int lc, int ld, bool pure, int mixed, const JkArgs& a, const GcArgs& g,
                        int nitems, LaunchShape sh, hipStream_t s);

}  // namespace hfg
