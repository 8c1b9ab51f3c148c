// Boys function F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt in FP64 on the device.
//
// T < kBoysTmax: 7-term Taylor expansion of F_M about the nearest grid point T_k = k/32 of a
// table of F_m(T_k) (|T - T_k| <= 1/64, truncation below 1e-16 relative), then the stable
// downward recursion F_m = (2T F_{m+1} + e^-T) / (2m+1).
// T >= kBoysTmax: F_0 = sqrt(pi/T)/2 (erfc(6) < 3e-17) and the upward recursion
// F_{m+1} = ((2m+1) F_m - e^-T) / 2T, which is stable for T > m.
// hfscf evaluates F_nmax through the incomplete gamma function with T clamped to >= 1e-8 and
// recurses down from nmax = 3L (hfscf: src/math/gamma.cpp:18-24, src/integrals/hfscf_oseri.cpp:103-118);
// the clamp changes the result only for an L = 0 quartet, which is mirrored by the caller.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace hfg {

constexpr int kBoysCols = 24;           // m = 0..23 stored per grid point
constexpr int kBoysMaxM = kBoysCols - 7;  // largest order the Taylor step supports
constexpr double kBoysStepInv = 32.0;
constexpr double kBoysTmax = 36.0;
constexpr int kBoysRows = 36 * 32 + 1;

// Host: F_m(k/32) for k < kBoysRows, m < kBoysCols, row-major (long-double series).
std::vector<double> boys_table();

template <int M, typename R>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ tab, double scale, R* F) {
    static_assert(M <= kBoysMaxM, "Boys order beyond table");
    const double eT = exp(-T);
    double f[M + 1];
    if (T < kBoysTmax) {
        const int k = __double2int_rn(T * kBoysStepInv);
        const double dt = (double)k * (1.0 / kBoysStepInv) - T;
        const double* row = tab + k * kBoysCols + M;
        double s = row[6];
        s = row[5] + s * dt * (1.0 / 6.0);
        s = row[4] + s * dt * (1.0 / 5.0);
        s = row[3] + s * dt * (1.0 / 4.0);
        s = row[2] + s * dt * (1.0 / 3.0);
        s = row[1] + s * dt * (1.0 / 2.0);
        s = row[0] + s * dt;
        f[M] = s;
#pragma unroll
        for (int m = M - 1; m >= 0; --m) f[m] = (2.0 * T * f[m + 1] + eT) * (1.0 / (2 * m + 1));
    } else {
        const double rT = 1.0 / T;
        f[0] = 0.88622692545275801365 * sqrt(rT);  // sqrt(pi)/2
#pragma unroll
        for (int m = 0; m < M; ++m) f[m + 1] = ((2 * m + 1) * f[m] - eT) * (0.5 * rT);
    }
#pragma unroll
    for (int m = 0; m <= M; ++m) F[m] = (R)(scale * f[m]);
}

}  // namespace hfg
