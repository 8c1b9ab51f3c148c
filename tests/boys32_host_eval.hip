// Host evaluation of the FP32 Boys values (boys32.hpp) for tests/test_boys32.py.
#include "../hfscf_amd/csrc/gpu/boys32.hpp"

extern "C" void boys32_eval(int m, int n, const float* T, float* F) {
    for (int i = 0; i < n; ++i) {
        float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (m == 0) hfg::boys32<0>(T[i], 1.0f, f);
        if (m == 1) hfg::boys32<1>(T[i], 1.0f, f);
        if (m == 2) hfg::boys32<2>(T[i], 1.0f, f);
        if (m == 3) hfg::boys32<3>(T[i], 1.0f, f);
        for (int k = 0; k <= m; ++k) F[i * (m + 1) + k] = f[k];
    }
}
