"""The cooperative J/K kernel's programs (hfscf_amd/csrc/gpu/coop.hip) evaluated on the CPU
against the oracle's shell quartets (CPU only).

tests/coop_host_eval.hip runs the steps of jk_coop_kernel (Boys values per order, VRR level by
level, contraction, bra and ket HRR, spherical transform) sequentially from the programs the
library generates, so a wrong index in a program fails here before any GPU run. Bar: 1e-14
absolute against the oracle's chained 1-D Obara-Saika quartets (pinned against an independent
McMurchie-Davidson code in test_oracle.py).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import golden
from oracle import basis_ref

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def evaluator(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("coop") / "libcoopeval.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                    os.path.join(HERE, "coop_host_eval.hip"), "-o", out], check=True, capture_output=True)
    return C.CDLL(out)


def _quartets(L):
    d = [i for i in range(len(L)) if L[i] == 2]
    p = [i for i in range(len(L)) if L[i] == 1]
    qs = []
    for a in d:
        for b in d:
            qs += [(a, b, d[0], d[-1]), (a, b, d[-1], p[0]), (a, p[1], d[0], d[-1]), (a, p[0], b, p[-1]),
                   (a, b, d[0], 0), (a, b, p[2], p[1])]
    return qs + _low_classes(L)


def _low_classes(L):
    # one quartet of every class from L = 0 up to 6 (la >= lb, lc >= ld), on shells of different
    # atoms where the basis has them: the classes the library may also run on these kernels (from
    # L = 4 for J/K, every class for the ERI export)
    by_l = {l: [i for i in range(len(L)) if L[i] == l] for l in range(3)}
    qs = []
    for la in range(3):
        for lb in range(la + 1):
            for lc in range(3):
                for ld in range(lc + 1):
                    if not 0 <= la + lb + lc + ld <= 6:
                        continue
                    pick = [by_l[l][(k * 3 + 1) % len(by_l[l])] for k, l in enumerate((la, lb, lc, ld))]
                    qs.append(tuple(pick))
    return qs


@pytest.mark.parametrize("basis,pure", [("cc-pvdz", None), ("cc-pvdz", False), ("6-31g*", None)])
def test_coop_programs_against_oracle(evaluator, basis, pure):
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    Z, X = Z[:6], X[:6]
    sh, pure, n = basis_ref.shells_for(Z, basis, pure)
    o = oracle.Oracle(sh, X, pure, 1e-30, screen=False)
    L = sh["L"]
    po = np.concatenate([[0], np.cumsum(sh["nprim"])])
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    worst = 0.0
    for q in _quartets(L):
        Ls = [int(L[s]) for s in q]
        cen = np.ascontiguousarray(np.array([X[sh["atom"][s]] for s in q]).reshape(-1))
        np4 = np.array([sh["nprim"][s] for s in q], dtype=np.int32)
        al = np.concatenate([sh["alpha"][po[s]:po[s + 1]] for s in q])
        co = np.concatenate([sh["coef"][po[s]:po[s + 1]] for s in q])
        nf = [(2 * l + 1) if pure else (l + 1) * (l + 2) // 2 for l in Ls]
        out = np.zeros(int(np.prod(nf)))
        lds = evaluator.coop_eval(*Ls, int(pure), cen.ctypes.data_as(dp), np4.ctypes.data_as(ip), al.ctypes.data_as(dp),
                                  co.ctypes.data_as(dp), out.ctypes.data_as(dp))
        assert 0 < lds <= 64 * 1024 - 2048  # FP64 LDS of the launch fits a 64 KB workgroup
        worst = max(worst, float(np.abs(out.reshape(nf) - o.shell_quartet(*q)).max()))
    assert worst < 1e-14, worst
    assert len(_low_classes(L)) == 36 - 3  # every class (la >= lb, lc >= ld) from L = 0 to 6


def test_coop_gradient_programs_against_oracle(evaluator):
    # the cooperative gradient kernel's programs (VRR at L+1, the six shifted classes' HRR and
    # contraction tables, the 2 alpha weights) against the oracle's derivative integrals
    # (pinned by finite differences in test_oracle.py) contracted with a random Cartesian density
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    Z, X = Z[:6], X[:6]
    sh, pure, n = basis_ref.shells_for(Z, "cc-pvdz", False)
    o = oracle.Oracle(sh, X, False, 1e-30, screen=False)
    L = sh["L"]
    po = np.concatenate([[0], np.cumsum(sh["nprim"])])
    d = [i for i in range(len(L)) if L[i] == 2]
    p = [i for i in range(len(L)) if L[i] == 1]
    qs = [(d[0], d[1], d[1], d[0]), (d[0], d[1], d[0], p[0]), (d[1], p[1], d[0], d[1]), (d[0], d[0], d[1], p[2]),
          (d[1], p[0], d[0], p[3]), (d[0], d[1], p[2], p[1]), (d[0], 0, d[1], d[1])] + _low_classes(L)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    rng = np.random.default_rng(0)
    for q in qs:
        Ls = [int(L[s]) for s in q]
        nc = [(l + 1) * (l + 2) // 2 for l in Ls]
        cen = np.ascontiguousarray(np.array([X[sh["atom"][s]] for s in q]).reshape(-1))
        np4 = np.array([sh["nprim"][s] for s in q], dtype=np.int32)
        al = np.concatenate([sh["alpha"][po[s]:po[s + 1]] for s in q])
        co = np.concatenate([sh["coef"][po[s]:po[s + 1]] for s in q])
        G = np.ascontiguousarray(rng.standard_normal(nc))
        out = np.zeros(9)
        lds = evaluator.coop_grad_eval(*Ls, cen.ctypes.data_as(dp), np4.ctypes.data_as(ip), al.ctypes.data_as(dp),
                                       co.ctypes.data_as(dp), G.ctypes.data_as(dp), out.ctypes.data_as(dp))
        assert 0 < lds <= 160 * 1024 - 6144  # fits a gfx950 workgroup next to the static LDS
        der = o.shell_quartet_deriv1(*q)
        ref = np.array([np.sum(G * der[k]) for k in range(9)])
        assert np.abs(out - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max()), (q, out - ref)
