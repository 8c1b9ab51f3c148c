// Two-pass gather of the work items' J/K tiles into the output matrices.
#pragma once
#include "layout.hpp"

namespace hfg {

// Entries per chunk of a tile's gather list in the first pass.
constexpr int kGatherChunk = 16;

// One chunk of a tile's J (list 0) or K (list 1) gather list: entries [g0, g1), whose sums per
// tile element go to part[part .. part + tile elements).
struct GatherChunk {
    int tile;
    int list;
    int g0, g1;
    long long part;
};

}  // namespace hfg
