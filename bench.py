"""Benchmark of the REPEAT fit path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mof1000]

A step is one whole charge fit from a device-resident cube (vdW filter, basis
streaming with normal-equation accumulation, constrained solve), the hot path
the reference runs as ``_filter_grid`` + ``fit_charges``.  The chosen form is
the split precision path (int8 fixed-point slices on MFMA with exact int32
accumulation, fp64 across tiles and in the solve); the plain fp64 GPU path
runs on the same inputs as the denominator of the speed ratio.  Inputs are the
seeded synthetic workload (exact Ewald ESP plus noise) named in ``--workload``.

Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT8_TOPS = 5000.0     # MI355X dense int8 MFMA
PEAK_FP64_TFLOPS = 78.6     # MI355X fp64 (vector and matrix)


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _workload(name):
    from pyrepeat_amd import synth
    if name.startswith("stream"):
        return synth.stream_request(int(name[6:]))
    return getattr(synth, name)()


def _constraints(system, tmpdir):
    from pyrepeat_amd import params
    symm = resp = None
    if system.symm_groups:
        p = os.path.join(tmpdir, "symm.txt")
        with open(p, "w") as f:
            f.write(system.symm_text())
        symm = params.read_symmetry(p)
    if system.resp_text:
        p = os.path.join(tmpdir, "resp.txt")
        with open(p, "w") as f:
            f.write(system.resp_text)
        resp = params.read_resp_params(p)
    return params.constraint_arrays(system.n_atoms, system.atomic_numbers, system.q_tot, symm, resp, None)


def run_gpu(ctx, system, con, precision, steps, warmup):
    from pyrepeat_amd import _abi
    prec = _abi.RP_PREC_FP64 if precision == "fp64" else _abi.RP_PREC_OZAKI
    fit = _abi.Fit(ctx, system.n_grid, system.axis, system.atomic_numbers, system.atom_pos, system.esp_cube,
                   _abi.RP_ORDER_CUBE, 1, prec, 1.0, 1000.0, 20.0)
    prepared = fit._constraints(con)
    for _ in range(warmup):
        q = fit.refit(con, prepared)
    ctx.profile(True)
    lat = []
    t_start = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        q = fit.refit(con, prepared)
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    stats = ctx.kernel_stats()
    ctx.profile(False)
    info = fit.info
    out = {"q": q, "lat": np.array(lat), "total": total, "kernels": stats, "M": int(info.n_keep),
           "N": int(fit.n_atoms), "K": int(info.n_k), "timings": fit.timings()}
    fit.close()
    return out


def cpu_baseline(system, con, target_seconds=15.0):
    """The oracle (C restatement of the reference CPU algorithm, OpenMP on the host cores) on a
    bounded sample of kept points, extrapolated to the whole fit."""
    from oracle import repeat_oracle as O
    threads = len(os.sched_getaffinity(0))
    box, vol = O.box_and_volume(system.n_grid, system.axis)
    ew = O.ewald_setup(box, vol, 20.0, 1)
    eff, rmin, rmax, _ = O.effective_atoms(system.atom_pos, system.atomic_numbers, system.axis, box, 1.0, 1000.0)
    t0 = time.perf_counter()
    keep = O.filter_mask(system.n_grid, system.axis, eff, rmin, rmax, threads)
    t_filter = time.perf_counter() - t0
    pos, v, _ = O.compact(keep, system.esp_flat(), system.n_grid, system.axis)
    M = len(v)
    # calibrate the per-row cost on a small sample, then time a sample sized to ~target_seconds
    rng = np.random.default_rng(0)
    probe = np.sort(rng.choice(M, size=min(M, 256), replace=False))
    t0 = time.perf_counter()
    O.phi_matrix(pos[probe], system.atom_pos, ew, nthreads=threads)
    per_row = (time.perf_counter() - t0) / len(probe)
    n_sample = int(min(M, max(256, target_seconds / max(per_row, 1e-9))))
    sample = np.sort(rng.choice(M, size=n_sample, replace=False))
    t0 = time.perf_counter()
    phi = O.phi_matrix(pos[sample], system.atom_pos, ew, nthreads=threads)
    t_phi = time.perf_counter() - t0
    t0 = time.perf_counter()
    _ = phi.T @ phi
    _ = phi.T @ v[sample]
    t_gram = time.perf_counter() - t0
    t_fit = t_filter + (t_phi + t_gram) * (M / n_sample)
    return {"value": round(1.0 / t_fit, 6), "unit": "fits/s", "cores": threads, "kind": "port",
            "sample": (f"oracle Phi (direct cos over K={len(ew.kcoefs)} k, erfc over {len(ew.shifts)} shifts) "
                       f"+ NumPy Gram on {n_sample} of M={M} kept points ({t_phi + t_gram:.1f} s), "
                       f"filter on the full cube ({t_filter:.1f} s), scaled to M"),
            "seconds_per_fit": round(t_fit, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mof1000")
    ap.add_argument("--baseline-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the fit path is measured on one MI355X; "
                                   "multi-GPU sharding is described, not run"}))
        return 2
    import tempfile
    from pyrepeat_amd import _abi
    from pyrepeat_amd.build import HERE  # noqa: F401  (package import check)

    t0 = time.perf_counter()
    system = _workload(args.workload)
    _log(f"workload {args.workload}: N={system.n_atoms} grid={tuple(int(v) for v in system.n_grid)} "
         f"generated in {time.perf_counter() - t0:.1f} s")
    with tempfile.TemporaryDirectory() as td:
        con = _constraints(system, td)
    ctx = _abi.Context(0, 0)
    _log("chosen (split) path ...")
    split = run_gpu(ctx, system, con, "split", args.steps, args.warmup)
    _log(f"  {args.steps / split['total']:.2f} fits/s, p50 {1e3 * np.median(split['lat']):.2f} ms")
    bsteps = args.baseline_steps or args.steps
    _log("plain fp64 path ...")
    f64 = run_gpu(ctx, system, con, "fp64", bsteps, min(args.warmup, 2))
    _log(f"  {bsteps / f64['total']:.2f} fits/s, p50 {1e3 * np.median(f64['lat']):.2f} ms")
    ctx.close()

    dq_vs_fp64 = float(np.max(np.abs(split["q"] - f64["q"])))
    golden = os.path.join(ROOT, "tests", "golden", f"{args.workload}.json")
    err_oracle = {}
    if os.path.exists(golden):
        g = json.load(open(golden))
        ref = np.array(g["charges"])
        err_oracle = {"split_max_dq_e": float(np.max(np.abs(split["q"] - ref))),
                      "fp64_max_dq_e": float(np.max(np.abs(f64["q"] - ref)))}

    fits_s = args.steps / split["total"]
    fits_s64 = bsteps / f64["total"]
    M, N, K = split["M"], split["N"], split["K"]
    # dominant kernel of the chosen path and its roofline
    kern = split["kernels"]
    dom = max(("kspace", "gram"), key=lambda k: kern[k][0])
    ms, launches = kern[dom]
    per_fit_launches = launches / args.steps
    if dom == "kspace":
        alg = 4.0 * M * N * K                       # F_k = 4 M N K per fit
    else:
        alg = M * N * (N + 1) + 3.0 * M * N         # F_G per fit
    achieved = alg / per_fit_launches / (ms / launches * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "k_oz_kspace" if dom == "kspace" else "k_oz_gram",
            "achieved": round(achieved, 3), "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_INT8_TOPS, 5), "traffic": None,
            "note": "algorithmic fp flops per launch / mean launch time (HIP events); peak = dense int8 "
                    "MFMA rate, the unit the kernel runs on (16 int8 slice products per fp product)"}
    cpu = None
    if not args.no_cpu:
        _log("cpu baseline (oracle sample) ...")
        cpu = cpu_baseline(system, con, args.cpu_seconds)
    kms = {k: {"total_ms": round(v[0], 3), "launches": v[1]} for k, v in kern.items() if v[1]}
    kms64 = {k: {"total_ms": round(v[0], 3), "launches": v[1]} for k, v in f64["kernels"].items() if v[1]}
    line = {
        "metric": "charge fits/s + p50/p99 latency, chosen vs plain FP64 GPU at <=1e-5 e",
        "value": round(fits_s, 4), "unit": "fits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * split["total"] / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "i8 slices -> i32 exact, f64 accumulate/solve",
        "data": "synthetic (seeded exact Ewald ESP + 1e-5 noise)",
        "config": {"workload": args.workload, "n_atoms": N, "grid": [int(v) for v in system.n_grid],
                   "kept_points": M, "k_vectors": K, "precision": "split (int8 Ozaki)"},
        "p50_ms": round(1e3 * float(np.median(split["lat"])), 3),
        "p99_ms": round(1e3 * float(np.percentile(split["lat"], 99)), 3),
        "fp64_fits_per_s": round(fits_s64, 4),
        "fp64_p50_ms": round(1e3 * float(np.median(f64["lat"])), 3),
        "speedup_vs_fp64_gpu": round(fits_s / fits_s64, 3),
        "max_dq_vs_fp64_e": dq_vs_fp64, "within_1e-5": dq_vs_fp64 <= 1e-5,
        "error_vs_oracle": err_oracle,
        "kernels_chosen": kms, "kernels_fp64": kms64,
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(fits_s / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
