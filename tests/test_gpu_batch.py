"""Batching mode: many molecules in one context (one persistent launch per class over the work
items of every molecule), HIP-graph replay, the launch-shape autotuner and the lockstep
batched SCF, each against the same molecules handled one at a time."""
import os

import numpy as np
import pytest

import hfscf_amd as H

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _clusters(count, basis="6-31g*"):
    frames = H.read_xyz_frames(os.path.join(GOLD, "batch_h2o.xyz"))[:count]
    return [H.Molecule.from_arrays(Z, X, basis) for Z, X in frames]


def _rand_density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return 0.05 * (A + A.T)


def test_batch_jk_matches_single_contexts():
    mols = _clusters(8)
    ctx = H.GpuContext(mols)
    Ds = [_rand_density(m.nbf, 40 + k) for k, m in enumerate(mols)]
    J, K = ctx.build_jk(np.concatenate([d.ravel() for d in Ds]))
    for k, (m, Jb, Kb) in enumerate(zip(mols, ctx.split(J), ctx.split(K))):
        Js, Ks = H.GpuContext(m).build_jk(Ds[k])
        err = max(np.abs(Jb - Js).max(), np.abs(Kb - Ks).max())
        print(f"molecule {k} ({m.nbf} bf): batch vs single J/K {err:.2e}", flush=True)
        assert err < 1e-12


def test_batch_gradient_matches_single_contexts():
    mols = _clusters(6)
    ctx = H.GpuContext(mols)
    Ds = [_rand_density(m.nbf, 70 + k) for k, m in enumerate(mols)]
    g = ctx.energy_backward(np.concatenate([d.ravel() for d in Ds]))
    a0 = 0
    for k, m in enumerate(mols):
        gs = H.GpuContext(m).energy_backward(Ds[k])
        err = np.abs(g[a0:a0 + m.natom] - gs).max()
        print(f"molecule {k}: batch vs single gradient {err:.2e}", flush=True)
        assert err < 1e-11
        a0 += m.natom


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replays_match_direct_launches(batched):
    mols = _clusters(4) if batched else _clusters(1)[0]
    direct = H.GpuContext(mols)
    graph = H.GpuContext(mols, graph=True)
    other = H.GpuContext(_clusters(2)[1])
    n = sum(m.nbf ** 2 for m in (mols if batched else [mols]))
    rng = np.random.default_rng(9)
    for rep in range(3):
        D = 0.05 * rng.standard_normal(n)
        if not batched:
            D = (D.reshape(mols.nbf, mols.nbf) + D.reshape(mols.nbf, mols.nbf).T).ravel()
        else:
            D = np.concatenate([(d + d.T).ravel() for d in direct.split(D)])
        Jd, Kd = direct.build_jk(D)
        Jg, Kg = graph.build_jk(D)
        other.build_jk(_rand_density(other.nbf, rep))  # eager work between replays
        err = max(np.abs(Jg - Jd).max(), np.abs(Kg - Kd).max())
        print(f"replay {rep}: graph vs direct {err:.2e}", flush=True)
        assert err < 1e-13


def test_autotune_keeps_results():
    mol = _clusters(10)[9]
    ctx = H.GpuContext(mol)
    D = _rand_density(mol.nbf, 3)
    J0, K0 = ctx.build_jk(D)
    g0 = ctx.energy_backward(D)
    ctx.autotune(3)
    shapes = ctx.launch_shapes()
    # tuned register kernels report their block, classes on the cooperative kernels -1, the
    # (ss|ss) J/K launch on the shell-group kernel -2
    assert shapes and all(s["jk"][1] in (-2, -1, 128, 256) and s["grad"][1] in (-1, 128, 256) for s in shapes)
    assert all(s["jk"][1] == -1 for s in shapes if sum(int(x) for x in s["cls"]) >= 7)
    print("autotuned shapes:", shapes, flush=True)
    J1, K1 = ctx.build_jk(D)
    g1 = ctx.energy_backward(D)
    assert max(np.abs(J1 - J0).max(), np.abs(K1 - K0).max()) < 1e-13
    assert np.abs(g1 - g0).max() < 1e-12


@pytest.mark.parametrize("precision", [H.FP64, H.MIXED])
def test_batch_rhf_matches_solo(precision):
    mols = _clusters(12)
    # no FP64 finish in the mixed solo runs: batches stay in one precision, and each molecule
    # follows the solo trajectory of the same options
    opts = dict(diis_range=6, energy_tol=1e-10, rms_tol=1e-9, incremental=1, precision=precision, polish_rms=0.0)
    res, st = H.batch_rhf(mols, gradient=True, graph=1, autotune=1, **opts)
    print(f"batch: {st}", flush=True)
    for k, m in enumerate(mols):
        solo = H.rhf(m, **opts)
        gs = H.rhf_gradient(m, solo["D"], solo["F"], precision=precision)
        de = res[k]["energy_final"] - solo["energy_final"]
        dg = np.abs(res[k]["gradient"] - gs).max()
        print(f"molecule {k} ({m.nbf} bf): dE {de:.2e} dg {dg:.2e} iterations {res[k]['iterations']} "
              f"vs {solo['iterations']}", flush=True)
        assert res[k]["converged"] and solo["converged"]
        assert abs(de) <= 1e-9
        # FP64: the same numbers. Mixed: the FP32 gradient kernels evaluated at two densities that
        # agree to the SCF tolerance differ by their FP32 rounding, about 1e-7 Eh/bohr
        assert dg <= (1e-8 if precision == H.FP64 else 2e-6)


def test_batch_seeded_subset_all_converge():
    # 64 clusters drawn with a fixed seed from the 512 of the batching workload, spanning
    # n = 1..10 waters; FP64 density-difference builds with density screening and a full build
    # every 8 iterations: every molecule converges, and a sample matches its solo run to 1e-9 Eh
    frames = H.read_xyz_frames(os.path.join(GOLD, "batch_h2o.xyz"))
    rng = np.random.default_rng(64)
    idx = sorted(rng.choice(len(frames), size=64, replace=False).tolist())
    mols = [H.Molecule.from_arrays(*frames[k], "6-31g*") for k in idx]
    sizes = sorted({m.natom // 3 for m in mols})
    assert sizes == list(range(1, 11)), sizes
    opts = dict(diis_range=6, energy_tol=1e-10, rms_tol=1e-9, incremental=1)
    res, st = H.batch_rhf(mols, gradient=True, graph=1, autotune=1, **opts)
    conv = [r["converged"] for r in res]
    print(f"batch of 64: {sum(conv)} converged, {st['iterations']} iterations, "
          f"max molecule iterations {max(r['iterations'] for r in res)}", flush=True)
    assert all(conv), [idx[k] for k, c in enumerate(conv) if not c]
    for k in range(0, 64, 11):
        solo = H.rhf(mols[k], **opts)
        assert solo["converged"]
        assert abs(res[k]["energy_final"] - solo["energy_final"]) <= 1e-9
