"""ORACLE — test infrastructure only (parity checker and timed CPU baseline).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package.

* ``liboracle.so`` (ref_oracle.cpp): C++ restatement of hfscf's two-electron algorithm
  (chained 1-D Obara-Saika ERIs with incomplete-gamma Boys, mean-Schwarz screening, stored
  ERIs, conventional digestion, on-the-fly derivative-ERI gradient contraction, and the
  second-derivative ERIs of the Hessian contracted with the density).
* ``md_eri.py``: an independent McMurchie-Davidson ERI evaluator in numpy (a textbook
  scheme with no code in common with either the oracle or the GPU kernels).
* ``scf_ref.py``: hfscf's RHF loop semantics (CORE guess, DIIS with max-error eviction,
  energy with the previous extrapolated Fock matrix) restated in numpy/scipy.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def build():
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.argtypes = [C.c_int, _ip, _dp, _ip, _ip, _dp, _dp, C.c_int, C.c_double, C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_nbf.argtypes = [C.c_void_p]
        L.orc_ncart.argtypes = [C.c_void_p]
        L.orc_num_eri.argtypes = [C.c_void_p]
        L.orc_num_eri.restype = C.c_longlong
        L.orc_build_eri.argtypes = [C.c_void_p]
        L.orc_build_eri_rows.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_build_eri_rows.restype = C.c_longlong
        L.orc_get_eri.argtypes = [C.c_void_p, _dp]
        L.orc_get_screen.argtypes = [C.c_void_p, _dp]
        L.orc_eri.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_eri.restype = C.c_double
        L.orc_fock.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_jk.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_shell_quartet.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_shell_quartet_deriv1.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_grad2e.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_grad2e_parts.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp]
        L.orc_shell_quartet_deriv2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_hess2e.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp]
        L.orc_boys.argtypes = [C.c_int, C.c_double]
        L.orc_boys.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    """hfscf's two-electron algorithm on the CPU for one molecule.

    shells: dict with L, atom, nprim, alpha, coef_raw (file coefficients; the oracle
    normalises them itself as hfscf's Shell does), xyz: natom x 3 bohr.
    """

    def __init__(self, shells, xyz, pure, integral_tol=1e-14, screen=True):
        L = np.ascontiguousarray(shells["L"], dtype=np.int32)
        atom = np.ascontiguousarray(shells["atom"], dtype=np.int32)
        nprim = np.ascontiguousarray(shells["nprim"], dtype=np.int32)
        alpha = np.ascontiguousarray(shells["alpha"], dtype=np.float64)
        coef = np.ascontiguousarray(shells["coef_raw"], dtype=np.float64)
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        centres = np.ascontiguousarray(xyz[atom].reshape(-1))
        self.natom = len(xyz)
        self.h = lib().orc_create(len(L), L.ctypes.data_as(_ip), _p(centres), atom.ctypes.data_as(_ip),
                                  nprim.ctypes.data_as(_ip), _p(alpha), _p(coef), int(pure), integral_tol,
                                  int(screen))
        self.nbf = lib().orc_nbf(self.h)
        self.ncart = lib().orc_ncart(self.h)
        self.L = L
        self.pure = pure
        self._built = False

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def build_eri(self):
        lib().orc_build_eri(self.h)
        self._built = True

    def eri_vector(self):
        v = np.zeros(lib().orc_num_eri(self.h))
        lib().orc_get_eri(self.h, _p(v))
        return v

    def screen(self):
        s = np.zeros((self.ncart, self.ncart))
        lib().orc_get_screen(self.h, _p(s))
        return s

    def fock(self, H, D):
        if not self._built:
            self.build_eri()
        F = np.zeros_like(H)
        lib().orc_fock(self.h, _p(np.ascontiguousarray(H)), _p(np.ascontiguousarray(D)), _p(F))
        return F

    def jk(self, D):
        if not self._built:
            self.build_eri()
        n = self.nbf
        J, K = np.zeros((n, n)), np.zeros((n, n))
        lib().orc_jk(self.h, _p(np.ascontiguousarray(D)), _p(J), _p(K))
        return J, K

    def nfun(self, s):
        l = int(self.L[s])
        return 2 * l + 1 if self.pure else (l + 1) * (l + 2) // 2

    def shell_quartet(self, a, b, c, d):
        shape = [self.nfun(s) for s in (a, b, c, d)]
        out = np.zeros(shape)
        lib().orc_shell_quartet(self.h, a, b, c, d, _p(out))
        return out

    def shell_quartet_deriv1(self, a, b, c, d):
        m = [(int(self.L[s]) + 1) * (int(self.L[s]) + 2) // 2 for s in (a, b, c, d)]
        out = np.zeros([9] + m)
        lib().orc_shell_quartet_deriv1(self.h, a, b, c, d, _p(out))
        return out

    def grad2e(self, D):
        g = np.zeros((self.natom, 3))
        lib().orc_grad2e(self.h, _p(np.ascontiguousarray(D)), self.natom, _p(g))
        return g

    def grad2e_parts(self, D):
        """The Coulomb and exchange parts of grad2e (hfscf's gradCoulomb, gradExchange)."""
        gc, gx = np.zeros((self.natom, 3)), np.zeros((self.natom, 3))
        lib().orc_grad2e_parts(self.h, _p(np.ascontiguousarray(D)), self.natom, _p(gc), _p(gx))
        return gc, gx

    def shell_quartet_deriv2(self, a, b, c, d):
        """45 Cartesian blocks, one per coordinate pair X1 <= X2 of (A, B, C) x (x, y, z)."""
        m = [(int(self.L[s]) + 1) * (int(self.L[s]) + 2) // 2 for s in (a, b, c, d)]
        out = np.zeros([45] + m)
        lib().orc_shell_quartet_deriv2(self.h, a, b, c, d, _p(out))
        return out

    def hess2e(self, D):
        """Two-electron Hessian at fixed density: (HC, HX), hfscf's m_hessianC and m_hessianEx."""
        n3 = 3 * self.natom
        hc, hx = np.zeros((n3, n3)), np.zeros((n3, n3))
        lib().orc_hess2e(self.h, _p(np.ascontiguousarray(D)), self.natom, _p(hc), _p(hx))
        return hc, hx


def boys(m, T):
    return lib().orc_boys(int(m), float(T))
