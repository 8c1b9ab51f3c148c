"""hfg_build_jk_noalloc (device pointers, fixed scratch, no allocation) against hfg_build_jk on
the same context, on two calls in a row (the second with a smaller molecule, sharing the GPU's
scratch), and its refusal of a build larger than the fixed scratch. Bar: 1e-14 absolute (the same
kernels in the same order)."""
import ctypes as C
import os

import numpy as np
import pytest

import hfscf_amd as H
from conftest import golden
from oracle import basis_ref

pytestmark = pytest.mark.gpu

HIP = "/opt/rocm/lib/libamdhip64.so"


def _hip():
    lib = C.CDLL(HIP)
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _noalloc_jk(hip, ctx, D):
    n = D.shape[0]
    nb = n * n * 8
    dD, dJ, dK = _dev(hip, nb), _dev(hip, nb), _dev(hip, nb)
    try:
        D = np.ascontiguousarray(D)
        assert hip.hipMemcpy(dD, D.ctypes.data_as(C.c_void_p), nb, 1) == 0
        rc = H.hfgpu().hfg_build_jk_noalloc(ctx._h, dD, dJ, dK)
        ctx._check(rc)
        J, K = np.zeros((n, n)), np.zeros((n, n))
        assert hip.hipMemcpy(J.ctypes.data_as(C.c_void_p), dJ, nb, 2) == 0
        assert hip.hipMemcpy(K.ctypes.data_as(C.c_void_p), dK, nb, 2) == 0
        return J, K
    finally:
        for p in (dD, dJ, dK):
            hip.hipFree(p)


def _mol(nwater, basis):
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    return H.Molecule.from_arrays(Z[:3 * nwater], X[:3 * nwater], basis)


def test_noalloc_matches_build_jk_twice():
    hip = _hip()
    rng = np.random.default_rng(3)
    for nwater, basis in [(2, "cc-pvdz"), (1, "6-31g*")]:  # the second call is the smaller input
        mol = _mol(nwater, basis)
        ctx = H.GpuContext(mol)
        A = rng.standard_normal((mol.nbf, mol.nbf))
        D = 0.05 * (A + A.T)
        J0, K0 = ctx.build_jk(D)
        J1, K1 = _noalloc_jk(hip, ctx, D)
        d = max(np.abs(J1 - J0).max(), np.abs(K1 - K0).max())
        print(f"(H2O){nwater} {basis}: max |noalloc - build_jk| = {d:.2e} (slab {ctx.stats()['slab_bytes']} B)",
              flush=True)
        assert d <= 1e-14


def test_noalloc_refuses_builds_beyond_the_scratch():
    hip = _hip()
    mol = H.Molecule.from_xyz(golden("h2o_16.xyz"), "cc-pvdz")
    ctx = H.GpuContext(mol)
    ctx.build_jk(np.eye(mol.nbf) * 0.1)  # sizes the slab statistics
    if ctx.stats()["slab_bytes"] <= (32 << 20) - 512:
        pytest.skip("(H2O)16 fits the fixed scratch")
    with pytest.raises(RuntimeError, match="fixed maximum"):
        _noalloc_jk(hip, ctx, np.eye(mol.nbf) * 0.1)
