"""Benchmark of the RHF two-electron hot path on one MI355X.

A step is one direct Fock build F = H + 2J - K (every surviving shell quartet evaluated,
digested and summed on the GPU) for the converged density of the workload. Default workload:
(H2O)8 / cc-pVDZ (pure d, 192 basis functions), Schwarz threshold 1e-12, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload h2o8|c12h26|h2o16]
                    [--no-cpu] [--no-extra]

Prints one JSON line (value = FP64 Fock builds/s on the GPU) with a roofline object for the
J/K kernels (model FLOPs of the surviving quartets / HIP-event kernel time, against the FP64
vector peak) and the CPU baseline (the oracle, timed on a bounded sample on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import hfscf_amd as H  # noqa: E402

METRIC = "Fock builds/s + gradients/s, mixed vs FP64 at |dE|<=1e-6 Eh; % of VALU peak"
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
WORKLOADS = {
    "h2o8": ("h2o_8.xyz", "cc-pvdz", "(H2O)8 / cc-pVDZ (pure d), 192 bf, direct RHF Fock build, Schwarz 1e-12"),
    "c12h26": ("c12h26.xyz", "6-31g**", "C12H26 / 6-31G** (Cartesian d), 310 bf, direct RHF Fock build"),
    "h2o16": ("h2o_16.xyz", "cc-pvdz", "(H2O)16 / cc-pVDZ (pure d), 384 bf, direct RHF Fock build"),
}


def time_builds(ctx, Hc, D, steps, warmup):
    for _ in range(warmup):
        ctx.fock(Hc, D)
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.fock(Hc, D)
        kern += ctx.stats()["kernel_seconds"]
    el = time.perf_counter() - t0
    return el, kern


def cpu_baseline(mol, budget_rows=32):
    """Oracle (CPU restatement of hfscf's stored-ERI path) on this host: ERI rows sampled with a
    fixed stride (an unbiased estimate of the full ERI build) plus one full digestion."""
    import oracle

    lib = oracle.lib()
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=True)
    ns = mol.nshell
    stride = max(1, ns // budget_rows)
    t_rows = 0.0
    q = 0
    for r in range(0, ns, stride):
        t = time.perf_counter()
        q += lib.orc_build_eri_rows(orc.h, r, r + 1)
        t_rows += time.perf_counter() - t
    t_eri = t_rows * stride
    orc._built = True
    n = mol.nbf
    D = np.eye(n) * 0.5
    Hc = np.zeros((n, n))
    t = time.perf_counter()
    orc.fock(Hc, D)
    t_dig = time.perf_counter() - t
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / (t_eri + t_dig),
        "unit": "Fock builds/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle ERI build on shell rows 0..{ns - 1} step {stride} ({q} shell quartets, "
                  f"{t_rows:.2f} s, scaled x{stride} to {t_eri:.1f} s) + one full stored-ERI digestion "
                  f"({t_dig:.2f} s), {threads} OpenMP threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="h2o8", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    xyz, basis, desc = WORKLOADS[args.workload]
    mol = H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", xyz), basis)
    S, T, V = mol.oneint()
    Hc = T + V
    # converged density from a GPU SCF (untimed)
    t0 = time.perf_counter()
    scf = H.rhf(mol, diis_range=8)
    t_scf = time.perf_counter() - t0
    D = scf["D"]
    print(f"# SCF {args.workload}: E = {scf['energy']:.10f} Eh, {scf['iterations']} iterations, "
          f"{t_scf:.2f} s", file=sys.stderr, flush=True)
    ctx = H.GpuContext(mol)
    el, kern = time_builds(ctx, Hc, D, args.steps, args.warmup)
    st = ctx.stats()
    builds_s = args.steps / el
    kern_per = kern / args.steps
    achieved = st["model_flops"] / kern_per / 1e12
    out = {
        "metric": METRIC,
        "value": builds_s,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "nbf": mol.nbf, "shell_quartets": st["quartets"], "precision": "fp64",
                   "scf_energy": scf["energy"], "scf_iterations": scf["iterations"], "scf_seconds": t_scf},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": None,
                     "model_gflop_per_build": st["model_flops"] / 1e9, "kernel_ms_per_build": 1e3 * kern_per},
    }
    ref_path = os.path.join(ROOT, "tests", "golden", "cfg2_h2o8_ccpvdz.npz")
    if args.workload == "h2o8" and os.path.exists(ref_path):
        out["config"]["dE_vs_oracle"] = scf["energy"] - float(np.load(ref_path)["energy"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
