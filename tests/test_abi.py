"""The C-ABI libraries load and export every function the headers declare (CPU, no GPU calls
beyond device discovery)."""
import os
import re

import numpy as np
import pytest

import hfscf_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:hfg|hfscf)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header,loader", [("hfgpu.h", H.hfgpu), ("hfscf.h", H.hfscf)])
def test_exports(header, loader):
    lib = loader()
    names = _declared(header)
    assert len(names) > 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_fails_cleanly():
    import ctypes as C
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = H.hfgpu()
    h = C.c_void_p()
    assert lib.hfg_create(C.byref(h), 0, None) == -4  # HFG_ERR_UNSUPPORTED
    with pytest.raises(RuntimeError):
        H.GpuContext(H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", "h2o.xyz"), "sto-3g"))


def test_shell_table_matches_independent_reader():
    from oracle import basis_ref
    for xyz, basis in [("h2o_8.xyz", "cc-pvdz"), ("c12h26.xyz", "6-31g**"), ("h2o.xyz", "6-31g*")]:
        mol = H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", xyz), basis)
        Z, X = basis_ref.read_xyz(os.path.join(ROOT, "tests", "golden", xyz))
        ref, pure, nbf = basis_ref.shells_for(Z, basis)
        sh = mol.shells()
        assert mol.pure == pure and mol.nbf == nbf
        for k in ["L", "atom", "nprim", "cart_off", "pure_off"]:
            assert (sh[k] == ref[k]).all(), k
        assert abs(sh["coef"] - ref["coef"]).max() < 1e-13
        assert abs(mol.geometry()[1] - X).max() < 1e-12


def test_config_sizes():
    # basis-function counts of the workloads (cc-pVDZ pure: 192 / 384; 6-31G** Cartesian: 310)
    sizes = {("h2o_8.xyz", "cc-pvdz"): 192, ("h2o_16.xyz", "cc-pvdz"): 384, ("c12h26.xyz", "6-31g**"): 310,
             ("h2o.xyz", "sto-3g"): 7}
    for (xyz, basis), n in sizes.items():
        assert H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", xyz), basis).nbf == n


def test_struct_layouts_match_headers(tmp_path):
    # ctypes mirrors of the option/result/stats structs against the C compiler's layout
    import ctypes as C
    import subprocess
    structs = {"hfg_options": H.HfgOptions, "hfg_stats": H.HfgStats, "hfscf_scf_options": H.ScfOptions,
               "hfscf_scf_result": H.ScfResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hfgpu.h"', '#include "hfscf.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, fname, v = line.split()
        got[(cname, fname)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_batch_geometry_frames():
    # the batching workload: 512 seeded (H2O)n clusters, n in 1..10 (seed 5), angstrom frames
    frames = H.read_xyz_frames(os.path.join(ROOT, "tests", "golden", "batch_h2o.xyz"))
    assert len(frames) == 512
    ns = [len(Z) // 3 for Z, _ in frames]
    assert min(ns) == 1 and max(ns) == 10
    for Z, X in frames:
        assert len(Z) % 3 == 0 and X.shape == (len(Z), 3)
        assert (Z[0::3] == 8).all() and (Z[1::3] == 1).all() and (Z[2::3] == 1).all()
        # O-H bonds of 0.9572 angstrom in bohr
        d = np.linalg.norm(X[1::3] - X[0::3], axis=1) / H.ANGSTROM_TO_BOHR
        assert np.abs(d - 0.9572).max() < 1e-9
    mol = H.Molecule.from_arrays(*frames[0], "6-31g*")
    assert mol.nbf == 19 * (len(frames[0][0]) // 3)


def _private_segments():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "hfscf_amd", "csrc"))
    import private_segments
    return private_segments


def _limit():
    text = open(os.path.join(ROOT, "hfscf_amd", "csrc", "gpu", "private_segments.hpp")).read()
    return int(re.search(r"kMaxPrivateBytes\s*=\s*(\d+)", text).group(1))


def test_private_segments_within_dispatch_limit():
    # every class kernel in the built library needs at most the per-lane scratch that the library
    # dispatches (the largest that has run), and none uses a dynamic stack
    P = _private_segments()
    ks = P.library_kernels(os.path.join(ROOT, "hfscf_amd", "lib", "libhfgpu.so"))
    cls = [k for k in ks if any(t in k[0] for t in ("jk_kernel", "jk_coop_kernel", "grad_kernel", "grad_coop_kernel"))]
    assert sum("jk_coop_kernel" in k[0] for k in cls) == 2  # FP64 and mixed
    assert sum("grad_coop_kernel" in k[0] for k in cls) == 2
    assert len(cls) >= 150, len(cls)
    limit = _limit()
    over = [(name, b) for name, b, dyn, _ in cls if b > limit or dyn]
    assert not over, over
    print(f"{len(cls)} class kernels, largest private segment {max(k[1] for k in cls)} bytes (limit {limit})")


def test_private_segment_table_matches_library():
    # the table the library checks before a dispatch lists the same per-class maxima
    text = open(os.path.join(ROOT, "hfscf_amd", "_build", "private_segments.cpp")).read()
    rows = re.findall(r"\{(\d), (\d), (\d), (\d), (\d), (\d), (\d), (\d+), (\d)\}", text)
    assert len(rows) >= 150
    assert max(int(r[7]) for r in rows) <= _limit()
    assert all(r[8] == "0" for r in rows)
