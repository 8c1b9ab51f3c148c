// Single-precision Boys function values for the mixed-precision builds of the classes whose
// primitive loops carry little VRR work (L = la+lb+lc+ld <= 3): there the Boys value
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

// Coefficients (-1)^k / (k! (2M + 2k + 1)), k < K, of the series of F_M, at compile time.
template <int M, int K>
struct Boys32Series {
    float c[K];
    constexpr Boys32Series() : c() {
        double fact = 1.0;
        for (int k = 0; k < K; ++k) {
            c[k] = (float)(((k & 1) ? -1.0 : 1.0) / (fact * (2 * M + 2 * k + 1)));
            fact *= (double)(k + 1);
        }
    }
};

// F_0..F_M(T) (M <= 3) scaled by `scale`, in FP32:
//   F_0(T) = sqrt(pi)/2 erf(sqrt T) / sqrt T (F_0(0) = 1); from T = 16 on erf(sqrt T) = 1 in
//           FP32 (erfc(4) < 1.6e-8), and a wavefront whose lanes all lie there skips erff;
//   T < Ts: F_M from its series sum_k (-T)^k / (k! (2M + 2k + 1)) and the lower orders by the
//           stable downward recursion F_m = (2T F_{m+1} + e^-T) / (2m + 1);
//   T >= Ts: upward, F_{m+1} = ((2m + 1) F_m - e^-T) / 2T.
// Ts = 1 for M <= 2 (10 series terms, the first omitted below 1.3e-8 of F_M; the upward steps
// lose less than a factor 3 each), Ts = 2 for M = 3 (16 terms, the first omitted below 2e-9 of
// F_3; at T = 2 the three upward steps together lose a factor of about 5 instead of 21 at T = 1).
// Both branches are evaluated and one is selected, so the lanes of a wavefront do not diverge.
template <int M>
__host__ __device__ inline void boys32(float T, float scale, float* F) {
    static_assert(M >= 0 && M <= 3, "FP32 Boys values for orders 0..3 only");
    float f0 = 1.0f;
    float ri = 0.0f;
    if (T > 0.0f) ri = boys32_rsqrt(T);
    if (T >= 16.0f)
        f0 = 0.886226925f * ri;  // sqrt(pi)/2
    else if (T > 0.0f)
        f0 = 0.886226925f * erff(T * ri) * ri;
    if constexpr (M == 0) {
        F[0] = scale * f0;
    } else {
        const float e = expf(-T);
        const float h = 0.5f * ri * ri;  // 1 / 2T
        constexpr int K = M == 3 ? 16 : 10;
        constexpr Boys32Series<M, K> ser{};
        float s = 0.0f;
#pragma unroll
        for (int k = K - 1; k >= 0; --k) s = ser.c[k] + T * s;
        float down[M + 1], up[M + 1];
        down[M] = s;
#pragma unroll
        for (int m = M - 1; m >= 0; --m) down[m] = (2.0f * T * down[m + 1] + e) * (1.0f / (2 * m + 1));
        up[0] = f0;
#pragma unroll
        for (int m = 0; m < M; ++m) up[m + 1] = ((2 * m + 1) * up[m] - e) * h;
        const bool small = T < (M == 3 ? 2.0f : 1.0f);
        F[0] = scale * f0;
#pragma unroll
        for (int m = 1; m <= M; ++m) F[m] = scale * (small ? down[m] : up[m]);
    }
}

}  // namespace hfg
