// Host generation of the FP64 Boys table F_m(k/32).
#include <cmath>

#include "boys.hpp"

namespace hfg {

// F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)): every term is positive, so the
// long-double sum is accurate to its working precision for the tabulated T <= 36.
std::vector<double> boys_table() {
    std::vector<double> tab((size_t)kBoysRows * kBoysCols);
    for (int k = 0; k < kBoysRows; ++k) {
        const long double T = (long double)k / kBoysStepInv;
        for (int m = 0; m < kBoysCols; ++m) {
            long double term = 1.0L / (2 * m + 1), sum = term;
            for (int j = 1; j < 2000; ++j) {
                term *= 2.0L * T / (2 * m + 2 * j + 1);
                sum += term;
                if (term < sum * 1e-21L) break;
            }
            tab[(size_t)k * kBoysCols + m] = (double)(sum * expl(-T));
        }
    }
    return tab;
}

}  // namespace hfg
