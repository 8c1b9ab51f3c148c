// Run-time dispatch from (la lb|lc ld), basis type and precision to the specialised J/K kernel.
// The classes from L = 7 up run the wavefront-cooperative kernel (coop.hpp), launched by the
// context, and have no register kernel.
#include "coop_args.hpp"
#include "jk_impl.hpp"
#include "launch_bounds.hpp"

namespace hfg {

#define HFG_DECL1(A, B, C, D, P)                                                                  \
    template <>                                                                                   \
    hipError_t launch_jk_class64<A, B, C, D, P>(const JkArgs&, int, LaunchShape, hipStream_t);    \
    template <>                                                                                   \
    hipError_t launch_jk_class32<A, B, C, D, P>(const JkArgs&, int, LaunchShape, hipStream_t);
#define HFG_DECL(A, B, C, D) HFG_DECL1(A, B, C, D, true) HFG_DECL1(A, B, C, D, false)
HFG_FOR_CLASSES(HFG_DECL)
#undef HFG_DECL
#undef HFG_DECL1

namespace {

template <int A, int B, int C, int D, bool P>
hipError_t jk_class(const JkArgs& a, int nitems, LaunchShape sh, int mixed, hipStream_t s) {
    if constexpr (A + B + C + D >= 7) {
        return hipErrorInvalidValue;
    } else {
        return mixed ? launch_jk_class32<A, B, C, D, P>(a, nitems, sh, s) : launch_jk_class64<A, B, C, D, P>(a, nitems, sh, s);
    }
}

}  // namespace

hipError_t launch_jk(int la, int lb, int lc, int ld, bool pure, int mixed, const JkArgs& a, int nitems,
                     LaunchShape want, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    const LaunchShape sh = scratch_safe_shape(la, lb, lc, ld, false, want, nitems);
#define HFG_CASE(A, B, C, D)                                                                      \
    if (la == A && lb == B && lc == C && ld == D)                                                 \
        return pure ? jk_class<A, B, C, D, true>(a, nitems, sh, mixed, s)                         \
                    : jk_class<A, B, C, D, false>(a, nitems, sh, mixed, s);
    HFG_FOR_CLASSES(HFG_CASE)
#undef HFG_CASE
    return hipErrorInvalidValue;
}

}  // namespace hfg
