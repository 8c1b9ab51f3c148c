"""The FP32 Boys values of the mixed-precision builds (hfscf_amd/csrc/gpu/boys32.hpp), evaluated
on the CPU against the incomplete gamma function (scipy), m = 0..3: F_m(T) = Gamma(m + 1/2) P(m + 1/2, T) /
(2 T^(m + 1/2)), with the T -> 0 limit 1/(2m + 1). Bar: 2e-6 relative (FP32 arithmetic, the
upward recursion at T >= 1 loses at most a few units of the last place per step)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy.special import gamma, gammainc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("boys32") / "libboys32.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                    os.path.join(HERE, "boys32_host_eval.hip"), "-o", out], check=True, capture_output=True)
    return C.CDLL(out)


def _boys(m, T):
    T = np.asarray(T, dtype=np.float64)
    out = np.empty_like(T)
    small = T < 1e-12
    out[small] = 1.0 / (2 * m + 1)
    t = T[~small]
    out[~small] = gamma(m + 0.5) * gammainc(m + 0.5, t) / (2 * t ** (m + 0.5))
    return out


@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_boys32_against_incomplete_gamma(lib, m):
    T = np.concatenate([[0.0, 1e-8, 1e-4], np.linspace(0.001, 2.4, 600), [0.999999, 1.0, 1.000001, 1.999999, 2.0,
                        2.000001, 15.99999, 16.0, 16.00001], np.geomspace(2.4, 120.0, 400)]).astype(np.float32)
    F = np.zeros((len(T), m + 1), dtype=np.float32)
    lib.boys32_eval(C.c_int(m), C.c_int(len(T)), T.ctypes.data_as(C.POINTER(C.c_float)),
                    F.ctypes.data_as(C.POINTER(C.c_float)))
    for k in range(m + 1):
        ref = _boys(k, T.astype(np.float64))
        rel = np.abs(F[:, k] - ref) / ref
        assert rel.max() < 2e-6, (m, k, float(T[np.argmax(rel)]), float(rel.max()))
