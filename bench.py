"""Benchmark of the RHF two-electron hot path on one MI355X.

A step is one direct Fock build F = H + 2J - K (every surviving shell quartet evaluated,
digested and summed on the GPU) for the converged density of the workload. Default workload:
(H2O)8 / cc-pVDZ (pure d, 192 basis functions), Schwarz threshold 1e-12, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload h2o8|c12h26|h2o16]
                    [--no-cpu] [--no-extra] [--no-mixed] [--no-scf] [--batch-molecules N]

Prints one JSON line (value = FP64 Fock builds/s on the GPU) with a roofline object for the
J/K kernels (model FLOPs of the surviving quartets / HIP-event kernel time, against the FP64
vector peak) and the CPU baseline (the oracle, timed on a bounded sample on this host).
"extra" holds the other figures of the metric: two-electron gradients/s (adjoint kernel),
the mixed-precision Fock build and gradient rates at the same density with their errors
against FP64, whole-SCF timings of the FP64 and mixed (incremental) solvers with |dE|, and the
batching mode (512 seeded water clusters in one lockstep SCF) against one-at-a-time solves.
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
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector (packed; 78.6 without packed FP32)
WORKLOADS = {
    "h2o8": ("h2o_8.xyz", "cc-pvdz", "(H2O)8 / cc-pVDZ (pure d), 192 bf, direct RHF Fock build, Schwarz 1e-12"),
    "c12h26": ("c12h26.xyz", "6-31g**", "C12H26 / 6-31G** (Cartesian d), 310 bf, direct RHF Fock build"),
    "h2o16": ("h2o_16.xyz", "cc-pvdz", "(H2O)16 / cc-pVDZ (pure d), 384 bf, direct RHF Fock build"),
}
SCF_OPTS = dict(diis_range=8, energy_tol=1e-10, rms_tol=1e-10)


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


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


def time_grads(ctx, D, steps, warmup):
    g = None
    for _ in range(warmup):
        g = ctx.energy_backward(D)
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        g = ctx.energy_backward(D)
        kern += ctx.stats()["kernel_seconds"]
    return time.perf_counter() - t0, kern, g


def copy_calibration(mib, reps=3):
    """A plain device-to-device copy of a known size, so that a profiler's HBM byte counters can
    be scaled against it."""
    import torch

    n = int(mib * 2 ** 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} copies of {n * 4} bytes (reads {n * 4} B, writes {n * 4} B each)")


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


def dominant_roofline(classes):
    """Roofline of the dominant J/K kernel (the class launch with the longest duration in the last
    build): model FLOPs of its surviving quartets / its HIP-event duration, against FP64 vector peak.
    Traffic is None: the profiler's HBM byte counters could not be collected for these kernels."""
    c = max(classes, key=lambda c: c["kernel_ms"])
    ach = c["model_flops"] / (c["kernel_ms"] * 1e-3) / 1e12
    return {"bound": "valu", "kernel": f"jk_kernel ({c['cls'][0]}{c['cls'][1]}|{c['cls'][2]}{c['cls'][3]}) fp64",
            "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach / PEAK_FP64, "traffic": None,
            "model_gflop_per_launch": c["model_flops"] / 1e9, "kernel_ms": c["kernel_ms"]}


def gradient_figures(mol, D, steps, warmup, autotune=True):
    """FP64 two-electron gradients/s (adjoint kernel) at the workload's density."""
    ctx = H.GpuContext(mol)
    if autotune:
        ctx.autotune(2)
    el, kern, g = time_grads(ctx, D, steps, warmup)
    st = ctx.stats()
    ctx.close()
    log(f"FP64 gradient: {steps / el:.2f}/s, kernel {1e3 * kern / steps:.2f} ms")
    return {"main_fp64_gradients_per_s": steps / el, "main_fp64_gradient_kernel_ms": 1e3 * kern / steps,
            "main_fp64_gradient_model_tflops": st["model_flops"] / (kern / steps) / 1e12}


def mixed_figures(workload, steps, warmup):
    """Mixed precision against FP64 on one workload ((H2O)16 / cc-pVDZ by default): Fock builds/s
    and gradients/s at the converged FP64 density (full builds, no density difference), and whole
    SCF runs, FP64 with full builds as hfscf does and mixed with FP32 density-difference builds,
    density screening and a final FP64 build, with |dE| and the gradient error."""
    xyz, basis, desc = WORKLOADS[workload]
    mol = H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", xyz), basis)
    S, T, V = mol.oneint()
    Hc = T + V
    out = {"mixed_workload": desc}
    t0 = time.perf_counter()
    scf64 = H.rhf(mol, **SCF_OPTS)
    out["scf_fp64_seconds"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    # FP32 density-difference builds leave about 1e-8 of error in the accumulated Fock matrix, so
    # the mixed SCF stops at rms 1e-7 (its energy is the final FP64 build's)
    scfm = H.rhf(mol, precision=H.MIXED, incremental=1, **dict(SCF_OPTS, energy_tol=1e-9, rms_tol=1e-7))
    out["scf_mixed_seconds"] = time.perf_counter() - t0
    out["scf_fp64_fock_seconds"] = scf64["fock_seconds"]
    out["scf_mixed_fock_seconds"] = scfm["fock_seconds"]
    out["scf_fp64_iterations"] = scf64["iterations"]
    out["scf_mixed_iterations"] = scfm["iterations"]
    out["scf_fp64_converged"] = bool(scf64["converged"])
    out["scf_mixed_converged"] = bool(scfm["converged"])
    out["scf_mixed_dE_vs_fp64"] = scfm["energy_final"] - scf64["energy_final"]
    out["scf_fock_speedup_mixed_vs_fp64"] = scf64["fock_seconds"] / max(scfm["fock_seconds"], 1e-12)
    log(f"{workload} SCF FP64 {out['scf_fp64_seconds']:.2f} s (Fock {scf64['fock_seconds']:.2f} s), mixed "
        f"{out['scf_mixed_seconds']:.2f} s (Fock {scfm['fock_seconds']:.2f} s), dE {out['scf_mixed_dE_vs_fp64']:.2e}")
    D = scf64["D"]
    res = {}
    for tag, prec in [("fp64", H.FP64), ("mixed", H.MIXED)]:
        ctx = H.GpuContext(mol, precision=prec)
        el, kern = time_builds(ctx, Hc, D, steps, warmup)
        F = ctx.fock(Hc, D)
        elg, kerng, g = time_grads(ctx, D, steps, warmup)
        ctx.close()
        res[tag] = (F, g)
        out[f"{tag}_fock_builds_per_s"] = steps / el
        out[f"{tag}_fock_kernel_ms"] = 1e3 * kern / steps
        out[f"{tag}_gradients_per_s"] = steps / elg
        out[f"{tag}_gradient_kernel_ms"] = 1e3 * kerng / steps
    out["mixed_fock_max_abs_err_vs_fp64"] = float(np.abs(res["mixed"][0] - res["fp64"][0]).max())
    out["mixed_energy_err_at_fixed_D"] = float(np.sum(D * (res["mixed"][0] - res["fp64"][0])))
    out["mixed_gradient_max_abs_err_vs_fp64"] = float(np.abs(res["mixed"][1] - res["fp64"][1]).max())
    out["fock_speedup_mixed_vs_fp64"] = out["mixed_fock_builds_per_s"] / out["fp64_fock_builds_per_s"]
    out["gradient_speedup_mixed_vs_fp64"] = out["mixed_gradients_per_s"] / out["fp64_gradients_per_s"]
    log(f"{workload} builds/s FP64 {out['fp64_fock_builds_per_s']:.2f} mixed {out['mixed_fock_builds_per_s']:.2f}; "
        f"gradients/s FP64 {out['fp64_gradients_per_s']:.2f} mixed {out['mixed_gradients_per_s']:.2f}")
    return out


def batch_figures(count, sample, seed=11):
    """Batching mode on the seeded (H2O)n clusters (n in 1..10, 6-31G*): all molecules in one
    lockstep SCF with one batched build per iteration (HIP graphs, autotuned launch shapes), then
    gradients in one batched launch; against one-at-a-time GPU solves of a random sample."""
    frames = H.read_xyz_frames(os.path.join(ROOT, "tests", "golden", "batch_h2o.xyz"))[:count]
    mols = [H.Molecule.from_arrays(Z, X, "6-31g*") for Z, X in frames]
    opts = dict(diis_range=6, energy_tol=1e-10, rms_tol=1e-9, incremental=1)
    t0 = time.perf_counter()
    res, st = H.batch_rhf(mols, gradient=True, graph=1, autotune=1, **opts)
    tb = time.perf_counter() - t0
    log(f"batch of {len(mols)}: {tb:.2f} s ({len(mols) / tb:.1f} molecules/s), stats {st}")
    rng = np.random.default_rng(seed)
    idx = sorted(rng.choice(len(mols), size=min(sample, len(mols)), replace=False).tolist())
    t_solo, de, dg = 0.0, 0.0, 0.0
    for k in idx:
        t0 = time.perf_counter()
        r = H.rhf(mols[k], **opts)
        g = H.rhf_gradient(mols[k], r["D"], r["F"])
        t_solo += time.perf_counter() - t0
        de = max(de, abs(r["energy_final"] - res[k]["energy_final"]))
        dg = max(dg, float(np.abs(g - res[k]["gradient"]).max()))
    solo_rate = len(idx) / t_solo
    log(f"solo sample of {len(idx)}: {solo_rate:.2f} molecules/s, max |dE| {de:.2e}, max |dg| {dg:.2e}")
    nconv = sum(r["converged"] for r in res)
    return {"batch_molecules": len(mols), "batch_converged": nconv, "batch_seconds": tb,
            "batch_molecules_per_s": len(mols) / tb, "batch_iterations": st["iterations"],
            "batch_ms_per_iteration": 1e3 * st["scf_seconds"] / max(st["iterations"], 1),
            "batch_setup_seconds": st["setup_seconds"], "batch_autotune_seconds": st["autotune_seconds"],
            "batch_gradient_seconds": st["gradient_seconds"], "batch_quartets_per_build": st["quartets"],
            "solo_sample": len(idx), "solo_molecules_per_s": solo_rate,
            "batch_speedup_vs_solo": (len(mols) / tb) / solo_rate,
            "batch_vs_solo_max_dE": de, "batch_vs_solo_max_dgrad": dg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="h2o8", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes and FLOP-model stream balance")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-scf", action="store_true",
                    help="time the builds on a seeded random density instead of the converged one "
                         "(FP64 build cost does not depend on the density values; for profiler runs)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="MIB",
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--mixed-workload", default="h2o16", choices=sorted(WORKLOADS),
                    help="workload of the mixed-precision against FP64 figures")
    ap.add_argument("--batch-molecules", type=int, default=512,
                    help="batching-mode figures over this many seeded water clusters (0 = skip)")
    ap.add_argument("--solo-sample", type=int, default=24,
                    help="molecules solved one at a time for the batching comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    xyz, basis, desc = WORKLOADS[args.workload]
    mol = H.Molecule.from_xyz(os.path.join(ROOT, "tests", "golden", xyz), basis)
    S, T, V = mol.oneint()
    Hc = T + V
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    if args.no_scf:
        rng = np.random.default_rng(7)
        A = rng.standard_normal((mol.nbf, mol.nbf))
        D = 0.05 * (A + A.T)
        scf = dict(energy=float("nan"), energy_final=float("nan"), iterations=0, seconds=0.0, fock_seconds=0.0)
    else:
        # converged density from a GPU SCF (untimed)
        t0 = time.perf_counter()
        scf = H.rhf(mol, **SCF_OPTS)
        scf["seconds"] = time.perf_counter() - t0
        D = scf["D"]
        log(f"SCF {args.workload}: E = {scf['energy']:.10f} Eh (final build {scf['energy_final']:.10f}), "
            f"{scf['iterations']} iterations, {scf['seconds']:.2f} s")
    ctx = H.GpuContext(mol)
    if not args.no_autotune:
        ctx.autotune(1)  # launch shapes and stream balance from measured class times (untimed set-up)
    el, kern = time_builds(ctx, Hc, D, args.steps, args.warmup)
    st = ctx.stats()
    classes = ctx.class_stats()  # the last build's class launches, bracketed by HIP events
    ctx.close()
    builds_s = args.steps / el
    kern_per = kern / args.steps
    achieved = st["model_flops"] / kern_per / 1e12
    log(f"FP64 Fock build: {builds_s:.2f}/s, kernel {1e3 * kern_per:.2f} ms, {achieved:.2f} TFLOP/s model")
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
                   "scf_energy": scf["energy"], "scf_energy_final": scf["energy_final"],
                   "scf_iterations": scf["iterations"], "scf_seconds": scf["seconds"],
                   "quartets_per_s": st["quartets"] * builds_s,
                   "launch_shapes": "default" if args.no_autotune else "autotuned"},
        "roofline": dominant_roofline(classes),
        "build_roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP64, "model_gflop_per_build": st["model_flops"] / 1e9,
                           "kernel_ms_per_build": 1e3 * kern_per,
                           "note": "all class launches of one build (four streams) over the event span"},
        "classes": [{"cls": c["cls"], "quartets": c["quartets"], "model_gflop": c["model_flops"] / 1e9,
                     "kernel_ms": c["kernel_ms"],
                     "tflops": c["model_flops"] / c["kernel_ms"] / 1e9 if c["kernel_ms"] > 0 else None}
                    for c in sorted(classes, key=lambda c: -c["kernel_ms"])],
    }
    ref_path = os.path.join(ROOT, "tests", "golden", "cfg2_h2o8_ccpvdz.npz")
    if args.workload == "h2o8" and os.path.exists(ref_path):
        f = np.load(ref_path)
        out["config"]["dE_vs_oracle"] = scf["energy"] - float(f["energy"])
        if "energy_final" in f:
            out["config"]["dE_final_vs_oracle"] = scf["energy_final"] - float(f["energy_final"])
    if not args.no_extra:
        out["extra"] = gradient_figures(mol, D, max(2, args.steps // 2), 1, not args.no_autotune)
        if not args.no_mixed:
            out["extra"].update(mixed_figures(args.mixed_workload, max(2, args.steps // 2), 1))
        if args.batch_molecules > 0:
            out["extra"].update(batch_figures(args.batch_molecules, args.solo_sample))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
