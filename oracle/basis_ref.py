"""ORACLE — test infrastructure only. hfscf's basis/molecule ordering rules in Python.

Restates hfscf: src/basis/hfscf_basis.cpp:19-275 (.gbs reading, SP split into S then P,
shells per atom in file order), src/integrals/hfscf_shell.cpp:44-54 (normalisation),
src/molecule/hfscf_molecule.cpp:92-94 and hfscf_constants.hpp:5 (angstrom -> bohr).
"""
import math
import os

import numpy as np

ANG2BOHR = 1.8897261254535
SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
BASISDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hfscf_amd", "basis")
GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def read_gbs(name):
    path = os.path.join(BASISDIR, name.lower().replace("*", "_st_") + ".gbs")
    lines = open(path).read().splitlines()
    spherical = lines[0].strip() == "spherical"
    lib, z, i = {}, None, 1
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if not ln or ln.startswith("!"):
            continue
        if ln.startswith("*"):
            z = None
            continue
        tok = ln.split()
        if z is None:
            z = SYMBOLS.index(tok[0])
            lib[z] = []
            continue
        kind, nprim = tok[0], int(tok[1])
        rows = [lines[i + k].replace("D", "E").split() for k in range(nprim)]
        i += nprim
        a = [float(r[0]) for r in rows]
        if kind == "SP":
            lib[z].append((0, a, [float(r[1]) for r in rows]))
            lib[z].append((1, a, [float(r[2]) for r in rows]))
        else:
            lib[z].append(("SPDFGH".index(kind), a, [float(r[1]) for r in rows]))
    return spherical, lib


def read_xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0])
    Z, X = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        Z.append(SYMBOLS.index(t[0]))
        X.append([float(v) * ANG2BOHR for v in t[1:4]])
    return np.array(Z), np.array(X)


def read_multi_xyz(path):
    lines = open(path).read().splitlines()
    out, k = [], 0
    while k < len(lines) and lines[k].strip():
        n = int(lines[k])
        Z, X = [], []
        for ln in lines[k + 2:k + 2 + n]:
            t = ln.split()
            Z.append(SYMBOLS.index(t[0]))
            X.append([float(v) * ANG2BOHR for v in t[1:4]])
        out.append((np.array(Z), np.array(X)))
        k += 2 + n
    return out


def dfac(n):
    return 1 if n <= 1 else n * dfac(n - 2)


def normalise(L, a, c):
    m = L + 1.5
    s = sum(c[j] * c[k] * (a[j] * a[k]) ** (m / 2) / (a[j] + a[k]) ** m for j in range(len(a)) for k in range(len(a)))
    norm = 1.0 / math.sqrt(s * math.pi ** 1.5 * dfac(2 * L - 1) / 2 ** L)
    return [c[j] * norm * a[j] ** (m / 2) for j in range(len(a))]


def shells_for(Z, basis, pure=None):
    """Shell table in hfscf order: dict of arrays like hfscf_amd.Molecule.shells()."""
    spherical, lib = read_gbs(basis)
    pure = spherical if pure is None else pure
    L, atom, nprim, co, po, alpha, coef, raw = [], [], [], [], [], [], [], []
    idx = ids = 0
    for at, z in enumerate(Z):
        for l, a, c in lib[int(z)]:
            L.append(l)
            atom.append(at)
            nprim.append(len(a))
            co.append(idx)
            po.append(ids)
            idx += (l + 1) * (l + 2) // 2
            ids += 2 * l + 1
            alpha += a
            raw += c
            coef += normalise(l, a, c)
    sh = dict(L=np.array(L, dtype=np.int32), atom=np.array(atom, dtype=np.int32),
              nprim=np.array(nprim, dtype=np.int32), cart_off=np.array(co, dtype=np.int32),
              pure_off=np.array(po, dtype=np.int32), alpha=np.array(alpha), coef=np.array(coef),
              coef_raw=np.array(raw))
    return sh, pure, (ids if pure else idx)
