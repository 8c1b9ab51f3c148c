// Private-segment (scratch) size per lane of every class kernel, from the code-object metadata
// of this build (a table generated at build time by csrc/private_segments.py), and the largest
// size the library dispatches.
//
// A dispatch whose kernels need more scratch per lane than the device grants fails inside the
// runtime (HSA_STATUS_ERROR_OUT_OF_RESOURCES, queue aborted): register-form gradient kernels of
// 27.9 KB and 61 KB per lane failed that way at every grid tried, while the J/K (dd|dd) kernel of
// 13,744 bytes per lane, the largest that has run, runs. The library refuses (HFG_ERR_UNSUPPORTED)
// any class kernel above that size instead of dispatching it.
#pragma once

namespace hfg {

struct PrivateSegment {
    int grad;   // 0 = J/K kernel, 1 = gradient kernel
    int mixed;  // 0 = FP64, 1 = mixed precision
    int la, lb, lc, ld;
    int pure;
    int bytes;          // .private_segment_fixed_size
    int dynamic_stack;  // .uses_dynamic_stack
};

extern const PrivateSegment kPrivateSegments[];
extern const int kNumPrivateSegments;

constexpr int kMaxPrivateBytes = 13744;

// Bytes per lane of the kernel a class launch runs (the mixed (dd|dd) gradient runs the FP64
// workspace kernel); -1 when the table has no entry.
inline int private_segment_bytes(bool grad, bool mixed, int la, int lb, int lc, int ld, bool pure) {
    for (int pass = 0; pass < 2; ++pass) {
        const int m = (pass == 0 && mixed) ? 1 : 0;
        if (pass == 1 && !mixed) break;
        for (int i = 0; i < kNumPrivateSegments; ++i) {
            const PrivateSegment& p = kPrivateSegments[i];
            if (p.grad == (grad ? 1 : 0) && p.mixed == m && p.la == la && p.lb == lb && p.lc == lc && p.ld == ld &&
                p.pure == (pure ? 1 : 0))
                return p.dynamic_stack ? 1 << 30 : p.bytes;
        }
    }
    return -1;
}

}  // namespace hfg
