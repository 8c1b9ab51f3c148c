// Single-precision Boys function values for the mixed-precision builds of the classes whose
// primitive loops carry no VRR work worth the name (L = la+lb+lc+ld <= 2): there the Boys value
// and the prefactor are most of each primitive quartet, and they stay in FP64 in every other
// class and in the FP64 builds.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace hfg {

__host__ __device__ inline float boys32_rsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrtf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

// F_0..F_M(T) (M <= 2) scaled by `scale`, in FP32:
//   F_0(T) = sqrt(pi)/2 erf(sqrt T) / sqrt T (F_0(0) = 1) for every T;
//   T < 1:  F_M from its series sum_k (-T)^k / (k! (2M + 2k + 1)), k <= 9 (the first omitted term
//           is below 1.3e-8 of F_M), and F_1 = (2T F_2 + e^-T) / 3 downward when M = 2;
//   T >= 1: upward, F_{m+1} = ((2m + 1) F_m - e^-T) / 2T, which at T >= 1 loses less than a
//           factor 3 of relative accuracy per step for m < 2.
// Both branches are evaluated and one is selected, so the lanes of a wavefront do not diverge.
template <int M>
__host__ __device__ inline void boys32(float T, float scale, float* F) {
    static_assert(M >= 0 && M <= 2, "FP32 Boys values for orders 0..2 only");
    float f0 = 1.0f;
    float ri = 0.0f;
    if (T > 0.0f) {
        ri = boys32_rsqrt(T);
        f0 = 0.886226925f * erff(T * ri) * ri;  // sqrt(pi)/2
    }
    F[0] = scale * f0;
    if constexpr (M >= 1) {
        const float e = expf(-T);
        const float h = 0.5f * ri * ri;  // 1 / 2T
        float s;
        if constexpr (M == 1) {
            s = -1.0f / 7620480.0f;
            s = 1.0f / 766080.0f + T * s;
            s = -1.0f / 85680.0f + T * s;
            s = 1.0f / 10800.0f + T * s;
            s = -1.0f / 1560.0f + T * s;
            s = 1.0f / 264.0f + T * s;
            s = -1.0f / 54.0f + T * s;
            s = 1.0f / 14.0f + T * s;
            s = -1.0f / 5.0f + T * s;
            s = 1.0f / 3.0f + T * s;
            const float up1 = (f0 - e) * h;
            F[1] = scale * (T < 1.0f ? s : up1);
        } else {
            s = -1.0f / 8346240.0f;
            s = 1.0f / 846720.0f + T * s;
            s = -1.0f / 95760.0f + T * s;
            s = 1.0f / 12240.0f + T * s;
            s = -1.0f / 1800.0f + T * s;
            s = 1.0f / 312.0f + T * s;
            s = -1.0f / 66.0f + T * s;
            s = 1.0f / 18.0f + T * s;
            s = -1.0f / 7.0f + T * s;
            s = 1.0f / 5.0f + T * s;
            const float down1 = (2.0f * T * s + e) * (1.0f / 3.0f);
            const float up1 = (f0 - e) * h;
            const float up2 = (3.0f * up1 - e) * h;
            const bool small = T < 1.0f;
            F[1] = scale * (small ? down1 : up1);
            F[2] = scale * (small ? s : up2);
        }
    }
}

}  // namespace hfg
