// Run-time dispatch from (la lb|lc ld) and precision to the specialised J/K kernel.
#include "jk_impl.hpp"

namespace hfg {

#define HFG_DECL(A, B, C, D)                                                                       \
    template <>                                                                                    \
    hipError_t launch_jk_class64<A, B, C, D>(bool, const JkArgs&, int, hipStream_t);               \
    template <>                                                                                    \
    hipError_t launch_jk_class32<A, B, C, D>(bool, const JkArgs&, int, hipStream_t);
HFG_FOR_CLASSES(HFG_DECL)
#undef HFG_DECL

hipError_t launch_jk(int la, int lb, int lc, int ld, bool pure, int mixed, const JkArgs& a, int nitems,
                     hipStream_t s) {
    if (nitems == 0) return hipSuccess;
#define HFG_CASE(A, B, C, D)                                                                     \
    if (la == A && lb == B && lc == C && ld == D)                                                \
        return mixed ? launch_jk_class32<A, B, C, D>(pure, a, nitems, s)                         \
                     : launch_jk_class64<A, B, C, D>(pure, a, nitems, s);
    HFG_FOR_CLASSES(HFG_CASE)
#undef HFG_CASE
    return hipErrorInvalidValue;
}

}  // namespace hfg
