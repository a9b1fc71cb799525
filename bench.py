"""Benchmark of the REPEAT fit path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mof1000]

A step is one whole charge fit from a device-resident cube (vdW filter, basis
streaming with normal-equation accumulation, constrained solve), the hot path
the reference runs as ``_filter_grid`` + ``fit_charges``.  The chosen form is
the refined split form (int8 fixed-point slices on MFMA with exact int32
accumulation, fp64 across tiles and in the solve, fp64 iterative refinement
against the kept basis); the plain fp64 GPU path (the reference's cos/sin
GEMM formulation) runs on the same inputs as the denominator of the speed
ratio, and the fp64 form with the factorized basis beside it, so the gain of
the precision form itself shows.  Inputs are the seeded synthetic workload
(exact Ewald ESP plus noise) named in ``--workload``.

Prints one JSON line (rank 0, N = 1).  Its "batch" object runs the same fit as
24 whole requests from host arrays queued at once on the serving loop with 4
slots (PCIe included), and its "serving" object is the serving stream below at
400 requests/s with 8 slots (``--no-serve`` leaves both out).

    python bench.py --serve [--requests 512] [--slots 8] [--rate 20] [--budget-gb 200]

runs the serving stream instead: seeded requests of 20-800 atoms arrive as a
Poisson process (20 requests/s unless ``--rate``) and are fitted concurrently
by the serving loop; the line reports open-loop fits/s with p50/p99 latency
(completion - arrival), the saturated capacity (every request queued at once),
and the charges of a sample against the same requests run alone and in plain
fp64.
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


def calibrate_copy(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime, a kernel of known traffic
    (reads and writes mib MiB) whose FETCH_SIZE / WRITE_SIZE counters scale those of the fit kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib) << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (src, dst):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) != 0:      # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def flops_from_profile(kernel, workload):
    """Executed fp64 flops per launch of `kernel` from the committed counter summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_fp64_flops.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("executed_flops_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


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
    prec = {"fp64": _abi.RP_PREC_FP64, "split": _abi.RP_PREC_OZAKI, "refined": _abi.RP_PREC_OZAKI_REFINED,
            "fp64-factorized": _abi.RP_PREC_FP64_FACT}[precision]
    fit = _abi.Fit(ctx, system.n_grid, system.axis, system.atomic_numbers, system.atom_pos, system.esp_cube,
                   _abi.RP_ORDER_CUBE, 1, prec, 1.0, 1000.0, 20.0)
    prepared = fit._constraints(con)
    for _ in range(warmup):
        q = fit.refit(con, prepared)
    # inside the timed region only the roofline kernel (the basis, the most time per fit) is timed:
    # every timed launch adds two event records to the stream, each a few microseconds of idle GPU
    ctx.profile(True, kinds=("kspace",))
    lat = []
    t_start = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        q = fit.refit(con, prepared)
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    dominant = ctx.kernel_stats()["kspace"]
    # per-kernel breakdown: every kind timed, on further fits after the timed region
    breakdown_fits = max(1, min(steps, 5))
    ctx.profile(True)
    for _ in range(breakdown_fits):
        fit.refit(con, prepared)
    stats = ctx.kernel_stats()
    ctx.profile(False)
    report = fit.solve_report()
    info = fit.info
    M = int(info.n_keep)
    N = int(fit.n_atoms)
    plan = _abi.chunk_plan(M, N, int(info.n_k), prec)
    rows = plan["rows"]
    # grid lines holding kept points, per chunk (a line cut by a chunk boundary counts in both chunks,
    # as the basis kernel forms its reciprocal table in both)
    _, _, idx = fit.grid()
    key = idx // int(system.n_grid[0])
    lines = [int(np.unique(key[c * rows:(c + 1) * rows]).size) for c in range(plan["n_chunks"])]
    # end to end from host arrays: the cube upload, the host Ewald setup and the filter (rp_fit_reset), the
    # accumulate and the solve, i.e. a whole fit as a caller holding the ESP in host memory sees it
    e2e = []
    for _ in range(3):
        t0 = time.perf_counter()
        fit.reset(system.n_grid, system.axis, system.atomic_numbers, system.atom_pos, system.esp_cube,
                  _abi.RP_ORDER_CUBE, 1, prec, 1.0, 1000.0, 20.0)
        fit.accumulate()
        fit.solve(con)
        e2e.append(time.perf_counter() - t0)
    out = {"q": q, "lat": np.array(lat), "total": total, "kernels": stats, "breakdown_fits": breakdown_fits,
           "dominant": dominant, "M": M, "rows": rows, "plan": plan, "lines_per_chunk": lines,
           "slices": plan["slices"], "N": N, "K": int(info.n_k), "volume": abs(float(info.volume)),
           "timings": fit.timings(), "report": report, "e2e": np.array(e2e),
           "refine": fit.refine_stats() if prec == _abi.RP_PREC_OZAKI_REFINED else None}
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
    # three runs on independent seeded samples, each ~1/3 of the time budget; the median per-row cost
    runs = 3
    n_sample = int(min(M, max(256, target_seconds / runs / max(per_row, 1e-9))))
    per_point = []
    for _ in range(runs):
        sample = np.sort(rng.choice(M, size=n_sample, replace=False))
        t0 = time.perf_counter()
        phi = O.phi_matrix(pos[sample], system.atom_pos, ew, nthreads=threads)
        _ = phi.T @ phi
        _ = phi.T @ v[sample]
        per_point.append((time.perf_counter() - t0) / n_sample)
    t_fit = t_filter + float(np.median(per_point)) * M
    return {"value": round(1.0 / t_fit, 6), "unit": "fits/s", "cores": threads, "kind": "port",
            "sample": (f"oracle Phi (direct cos over K={len(ew.kcoefs)} k, erfc over {len(ew.shifts)} shifts) "
                       f"+ NumPy Gram, {runs} runs on seeded samples of {n_sample} of M={M} kept points "
                       f"(median {1e6 * float(np.median(per_point)):.1f} us per point, spread "
                       f"{1e6 * min(per_point):.1f}-{1e6 * max(per_point):.1f}), filter on the full cube "
                       f"({t_filter:.1f} s), scaled to M; strict IEEE libm cos/erfc, where the reference's "
                       f"numba path uses fastmath"),
            "seconds_per_fit": round(t_fit, 2)}


def _serve_args(system):
    return (system.n_grid, system.axis, system.atomic_numbers, system.atom_pos, system.esp_cube)


def _run_stream(srv, reqs, opts, arrivals):
    """Submit reqs[r] at arrival time arrivals[r] (seconds after the start); returns results in order."""
    import threading
    t0 = srv.now() + 0.05
    tickets = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]

    def submitter():
        for r, (system, con) in enumerate(reqs):
            wait = t0 + arrivals[r] - srv.now()
            if wait > 0:
                time.sleep(wait)
            tickets[r] = srv.submit(_serve_args(system), con, opts)
            ready[r].set()

    th = threading.Thread(target=submitter)
    th.start()
    out = []
    for r in range(len(reqs)):
        ready[r].wait()
        q, res = srv.poll(tickets[r], -1)
        if res["status"] != 0:
            raise RuntimeError(f"request {r} failed: {res['error']}")
        res["arrival"] = t0 + arrivals[r]
        out.append((q, res))
    th.join()
    return out, t0


def serve_run(R, slots, rate, budget_gb, chosen):
    """The serving stream: R requests, open loop at ``rate`` req/s, then all queued at once (capacity),
    then a sample rerun alone and in fp64.  Returns the result line as a dict."""
    from pyrepeat_amd import _abi, synth
    arrivals_all, _ = synth.stream_schedule()
    arrivals = arrivals_all[:R] * (20.0 / rate)
    t0 = time.perf_counter()
    systems = synth.stream_requests(R)
    _log(f"{R} stream requests generated in {time.perf_counter() - t0:.1f} s")
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        reqs = [(s, _constraints(s, td)) for s in systems]
    chosen_prec = _abi.RP_PREC_OZAKI_REFINED if chosen == "refined" else _abi.RP_PREC_OZAKI
    split = _abi.make_options(1, chosen_prec, 1.0, 1000.0, 20.0)
    f64 = _abi.make_options(1, _abi.RP_PREC_FP64, 1.0, 1000.0, 20.0)
    srv = _abi.Server(0, slots, int(budget_gb * 1e9))
    # warm-up: every slot sees the largest request once, so the timed stream allocates nothing
    big = max(range(R), key=lambda r: int(np.prod(systems[r].n_grid)) * systems[r].n_atoms)
    tk = [srv.submit(_serve_args(reqs[big][0]), reqs[big][1], split) for _ in range(slots)]
    for t in tk:
        srv.poll(t, -1)
    _log("open-loop stream ...")
    res, t_start = _run_stream(srv, reqs, split, arrivals)
    lat = np.array([r["t_done"] - r["arrival"] for _, r in res])
    span = max(r["t_done"] for _, r in res) - t_start
    st_open = srv.stats()
    _log(f"  {R / span:.2f} fits/s, p50 {1e3 * np.median(lat):.1f} ms, p99 {1e3 * np.percentile(lat, 99):.1f} ms")
    _log("saturated (all queued at once) ...")
    sat, t_sat = _run_stream(srv, reqs, split, np.zeros(R))
    span_sat = max(r["t_done"] for _, r in sat) - t_sat
    st_sat = srv.stats()
    # a sample run alone (one request on the server at a time), in split and in fp64
    sample = sorted(set(np.linspace(0, R - 1, min(16, R)).astype(int).tolist()))
    dq_alone, dq_f64 = 0.0, 0.0
    for r in sample:
        system, con = reqs[r]
        q_alone, _ = srv.poll(srv.submit(_serve_args(system), con, split), -1)
        q64, _ = srv.poll(srv.submit(_serve_args(system), con, f64), -1)
        dq_alone = max(dq_alone, float(np.max(np.abs(q_alone - res[r][0]))))
        dq_f64 = max(dq_f64, float(np.max(np.abs(q_alone - q64))))
    srv.close()
    service = np.array([r["t_done"] - r["t_admit"] for _, r in res])
    queue_wait = np.array([r["t_admit"] - r["arrival"] for _, r in res])
    n_atoms = np.array([s.n_atoms for s in systems])
    # service time of the saturated run by request size: the requests sorted by atom count into eighths
    # (with one slot the requests run one after another, so these are the sizes' device times)
    svc_sat = np.array([r["t_done"] - r["t_admit"] for _, r in sat])
    order = np.argsort(n_atoms, kind="stable")
    eighths = [order[i * R // 8:(i + 1) * R // 8] for i in range(8)] if R >= 8 else [order]
    by_size = [{"atoms": [int(n_atoms[e].min()), int(n_atoms[e].max())],
                "service_ms": round(1e3 * float(svc_sat[e].mean()), 3),
                "share": round(float(svc_sat[e].sum() / svc_sat.sum()), 3)} for e in eighths]
    line = {
        "metric": f"charge fits/s + p50/p99 latency, serving stream (Poisson {rate:g} req/s, 20-800 atoms)",
        "value": round(R / span, 4), "unit": "fits/s", "n_gpus": 1, "steps": R, "warmup": slots,
        "ms_per_step": round(1e3 * span / R, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 basis/solve/refinement, i8 slices -> i32 exact Gram" if chosen == "refined" else
                 "f64 basis/solve, i8 slices -> i32 exact Gram",
        "data": "synthetic stream (seeded Ewald ESP + 1e-5 noise, fp32 cubes)",
        "config": {"workload": f"stream{R}", "requests": R, "rate_per_s": rate, "slots": slots,
                   "hbm_budget_gb": budget_gb, "atoms_min": int(n_atoms.min()), "atoms_max": int(n_atoms.max()),
                   "grid_points_total": int(sum(int(np.prod(s.n_grid)) for s in systems)),
                   "precision": "split (int8 Ozaki, 4 slices)" if chosen == "split" else
                   "refined (int8 Ozaki, 3 slices, fp64 residual refinement)"},
        "p50_ms": round(1e3 * float(np.median(lat)), 3), "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "service_p50_ms": round(1e3 * float(np.median(service)), 3),
        "queue_wait_p99_ms": round(1e3 * float(np.percentile(queue_wait, 99)), 3),
        "capacity_fits_per_s": round(R / span_sat, 4),
        "max_running": int(st_sat["max_running"]), "peak_reserved_gb": round(st_sat["peak_reserved"] / 1e9, 3),
        "open_loop_max_running": int(st_open["max_running"]),
        "mix_vs_alone_max_dq_e": dq_alone, "chosen_vs_fp64_max_dq_e": dq_f64, "sample_requests": sample,
        "saturated_service_by_size": by_size,
    }
    return line


def batch_run(system, con, chosen, slots, requests, budget_gb):
    """``requests`` copies of one fit queued at once on a server with ``slots`` slots, each a whole
    request from host arrays (cube upload, host Ewald setup, filter, accumulate, solve)."""
    from pyrepeat_amd import _abi
    prec = _abi.RP_PREC_OZAKI_REFINED if chosen == "refined" else _abi.RP_PREC_OZAKI
    opt = _abi.make_options(1, prec, 1.0, 1000.0, 20.0)
    sargs = _serve_args(system)
    srv = _abi.Server(0, slots, int(budget_gb * 1e9))
    for t in [srv.submit(sargs, con, opt) for _ in range(slots)]:   # every slot sized once
        srv.poll(t, -1)
    t0 = time.perf_counter()
    tickets = [srv.submit(sargs, con, opt) for _ in range(requests)]
    qs = [srv.poll(t, -1)[0] for t in tickets]
    dt = time.perf_counter() - t0
    srv.close()
    return {"fits_per_s": round(requests / dt, 3), "ms_per_fit": round(1e3 * dt / requests, 3), "slots": slots,
            "requests": requests, "max_dq_between_requests_e": max(float(np.max(np.abs(q - qs[0]))) for q in qs),
            "note": "the workload's fit as whole requests from host arrays (cube upload over PCIe, host Ewald "
                    "setup, filter, accumulate, solve), all queued at once on the serving loop"}


def serve_main(args):
    print(json.dumps(serve_run(args.requests, args.slots, args.rate, args.budget_gb, args.chosen)))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mof1000")
    ap.add_argument("--chosen", choices=["split", "refined"], default="refined",
                    help="the chosen path: split (4 slices) or refined (3 slices + fp64 refinement, the default: "
                         "fp64-level charges with an error bound per fit)")
    ap.add_argument("--baseline-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving stream (512 requests at 400 req/s, 8 slots) of the default line")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve", action="store_true", help="run the serving stream instead of the single-fit bench")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--rate", type=float, default=20.0,
                    help="serving: arrival rate of the open-loop stream (requests/s); the seeded exponential gaps "
                         "of the 20 req/s schedule are scaled to it")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the fit path is measured on one MI355X; "
                                   "multi-GPU sharding is described, not run"}))
        return 2
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
    if args.serve:
        return serve_main(args)
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
    _log(f"chosen ({args.chosen}) path ...")
    split = run_gpu(ctx, system, con, args.chosen, args.steps, args.warmup)
    _log(f"  {args.steps / split['total']:.2f} fits/s, p50 {1e3 * np.median(split['lat']):.2f} ms")
    bsteps = args.baseline_steps or args.steps
    _log("plain fp64 path ...")
    f64 = run_gpu(ctx, system, con, "fp64", bsteps, min(args.warmup, 2))
    _log(f"  {bsteps / f64['total']:.2f} fits/s, p50 {1e3 * np.median(f64['lat']):.2f} ms")
    _log("fp64 path with the factorized basis ...")
    f64f = run_gpu(ctx, system, con, "fp64-factorized", bsteps, min(args.warmup, 2))
    _log(f"  {bsteps / f64f['total']:.2f} fits/s, p50 {1e3 * np.median(f64f['lat']):.2f} ms")
    ctx.close()

    dq_vs_fp64 = float(np.max(np.abs(split["q"] - f64["q"])))
    golden = os.path.join(ROOT, "tests", "golden", f"{args.workload}.json")
    err_oracle = {}
    if os.path.exists(golden):
        g = json.load(open(golden))
        if "charges" in g:
            ref = np.array(g["charges"])
            err_oracle = {"chosen_max_dq_e": float(np.max(np.abs(split["q"] - ref))),
                          "fp64_max_dq_e": float(np.max(np.abs(f64["q"] - ref))),
                          "fp64_factorized_max_dq_e": float(np.max(np.abs(f64f["q"] - ref)))}
        if "n_keep" in g:
            # the golden of a configuration too large for the oracle's solve holds its filter result only
            err_oracle["kept_points_equal_oracle"] = int(g["n_keep"]) == int(split["M"])

    fits_s = args.steps / split["total"]
    fits_s64 = bsteps / f64["total"]
    fits_s64f = bsteps / f64f["total"]
    M, N, K = split["M"], split["N"], split["K"]
    # dominant kernel of the chosen path: the basis (k_sep_phi, the most time per fit), bound by its fp64
    # vector arithmetic.  Its algorithmic fp64 flops per launch, from the factorized sum's units:
    #   per (grid line, atom)   358: T2 for h = 0..6, folded by conjugate symmetry into T1[0] plus 6 pairs
    #                                (P_k Re Y + Q_k Im Y): 24 multiply-adds (48 flops) per h, 336; the line's
    #                                part of the nearest-image test and of the displacement (axes 1 and 2), 22
    #   per (kept point, atom)   62: Re sum_h T2 X^h by Horner (5 complex steps + the last, 44) and the
    #                                point's part of the test and displacement (axis 0) with r^2, 18
    #   per real-space term      47: erfc(sqrt(a) r)/r = 1/r - sqrt(a) P(u), P of degree 21 by Horner
    #                                (u, 21 FMAs, the last FMA) plus the reciprocal square root
    # with the terms inside the cutoff counted as the sphere's share of the cell, (4/3) pi Rc^3 / V.
    ms, launches = split["dominant"]
    launches_per_fit = launches / args.steps
    t_launch = ms / launches * 1e-3
    rows, lines = split["rows"], split["lines_per_chunk"]
    frac_in = min(1.0, 4.0 / 3.0 * np.pi * 20.0 ** 3 / split["volume"])
    alg_fit = 358.0 * sum(lines) * N + 62.0 * M * N + 47.0 * frac_in * M * N
    alg_launch = alg_fit / launches_per_fit
    achieved = alg_launch / t_launch / 1e12
    executed = flops_from_profile("k_sep_phi", args.workload)
    traffic = traffic_from_profile("k_sep_phi", args.workload)
    roof = {"bound": "mfma", "compute_unit": "fp64 vector ALU (v_fma_f64); the MI355X fp64 vector and matrix "
                                              "peaks are both 78.6 TFLOP/s",
            "kernel": "k_sep_phi", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
            "launch_ms": round(ms / launches, 4), "rows_per_launch": rows,
            "algorithmic_flops_per_launch": round(alg_launch), "lines_per_fit": int(sum(lines)),
            "executed_flops_per_launch": executed,
            "executed_tflops": round(executed / t_launch / 1e12, 3) if executed else None,
            "executed_frac": round(executed / t_launch / 1e12 / PEAK_FP64_TFLOPS, 4) if executed else None,
            "hbm_gbs": round(traffic / t_launch / 1e9, 1) if traffic else None,
            "hbm_frac": round(traffic / t_launch / 8e12, 4) if traffic else None,
            "fk_equivalent_tflops": round(4.0 * M * N * K / launches_per_fit / t_launch / 1e12, 1),
            "erfc_terms_per_launch": round(frac_in * M * N / launches_per_fit),
            "erfc_terms_per_s": round(frac_in * M * N / launches_per_fit / t_launch, -6),
            "note": "achieved = the factorized sum's algorithmic fp64 flops of one launch (358 per line and atom, "
                    "62 per kept point and atom, 47 per real-space term) / mean launch time (HIP events on the "
                    "fit's stream); executed = rocprof SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 x 64 lanes (idle "
                    "lanes included); traffic = calibrated FETCH_SIZE + WRITE_SIZE; fk_equivalent counts "
                    "4 M N K, the reference's cos/sin GEMM formulation, which this kernel does not execute; "
                    "erfc_terms = the real-space terms inside the cutoff (the sphere's share of the cell), each an "
                    "fp64 polynomial erfc, per launch and per second of the launch"}
    # the int8 Gram (MFMA), from the per-kernel breakdown fits after the timed region
    gms, glaunch = split["kernels"]["gram"]
    g_launch = gms / glaunch * 1e-3
    alg_per_launch = (M * N * (N + 1) + 3.0 * M * N) / (glaunch / split["breakdown_fits"])
    np_ = split["plan"]["np"]
    tiles = split["plan"]["n_tiles"]
    P = split["slices"]
    executed = 2.0 * rows * tiles * 64 * 128 * P * P / g_launch / 1e12    # int8 ops issued, P^2 slice products
    g_traffic = traffic_from_profile("k_oz_gram", args.workload)
    # slices once, the row-split partial tiles read and written once, and the slots of the items that the
    # library halves to complete an XCD's last round (its rule, for the 32 CUs of an MI355X XCD)
    n_splits = split["plan"]["splits"]
    halve = 0
    if n_splits % 8 == 0:
        rem = (tiles * (n_splits // 8)) % 32
        if 0 < rem <= 16 and rem <= tiles and (rows // n_splits) // 2 >= (128 if P == 3 else 64):
            halve = rem
    g_alg_bytes = P * np_ * rows + 2 * 8.0 * (n_splits * tiles + 8 * halve) * 64 * 128
    roof_gram = {"bound": "mfma", "kernel": "k_oz_gram", "achieved": round(alg_per_launch / g_launch / 1e12, 3),
                 "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                 "frac": round(alg_per_launch / g_launch / 1e12 / PEAK_INT8_TOPS, 5),
                 "traffic": g_traffic, "algorithmic_bytes": round(g_alg_bytes),
                 "traffic_over_algorithmic": round(g_traffic / g_alg_bytes, 3) if g_traffic else None,
                 "executed_int8_tops": round(executed, 1), "executed_frac": round(executed / PEAK_INT8_TOPS, 4),
                 "launch_ms": round(gms / glaunch, 4), "slices": P,
                 "note": "achieved = fp64-equivalent Gram flops F_G = M N (N + 1) + 3 M N of the launch's rows / "
                         f"mean launch time (events, breakdown fits); peak = dense int8 MFMA; each fp product costs "
                         f"P^2 = {P * P} int8 slice products, which executed_int8_tops counts"}
    # the vdW filter phase (mask, scan, compaction with the ESP gather, mean, centring and bounds), bound by
    # HBM in principle: its algorithmic bytes B_f = n_total (s_V + 1) + M (4 + s_V) (read the ESP, write the
    # mask byte, write the kept index and the gathered value) over the phase's device time (HIP events
    # around its kernels, last breakdown fit)
    n_total = int(np.prod(system.n_grid))
    s_v = system.esp_cube.dtype.itemsize
    b_f = n_total * (s_v + 1) + M * (4 + s_v)
    t_f = split["timings"]["filter_ms"] * 1e-3
    f_traffic = traffic_from_profile("filter", args.workload)
    fk = split["kernels"].get("filter")   # k_filter_mask alone: (ms, launches) over the breakdown fits
    roof_filter = {"bound": "hbm", "kernel": "filter phase (k_filter_mask, k_scan_counts, k_compact, k_mean, "
                                             "k_center_and_bounds, k_filter_publish)",
                   "achieved": round(b_f / t_f / 1e9, 1) if t_f > 0 else None, "peak": 8000.0, "unit": "GB/s",
                   "frac": round(b_f / t_f / 8e12, 4) if t_f > 0 else None, "traffic": f_traffic,
                   "algorithmic_bytes": int(b_f), "phase_ms": round(1e3 * t_f, 4),
                   "mask_kernel_ms": round(fk[0] / fk[1], 4) if fk and fk[1] else None,
                   "note": "compute-bound on the per-point cell-list scan (k_filter_mask); the phase reads the "
                           "cube once and writes the mask and the kept points"}
    cpu = None
    if not args.no_cpu:
        _log("cpu baseline (oracle sample) ...")
        cpu = cpu_baseline(system, con, args.cpu_seconds)
    batch = None
    if not args.no_serve:
        # the same fit as whole requests through the serving loop, several in flight: the latency-bound
        # phases of one fit (solve, refinement substitutions, host set-up) overlap the others' kernels
        _log("batch of whole requests through the server (4 slots) ...")
        batch = batch_run(system, con, args.chosen, slots=4, requests=24, budget_gb=args.budget_gb)
        _log(f"  {batch['fits_per_s']:.1f} fits/s")
    serving = None
    if not args.no_serve:
        _log("serving stream (512 requests, 400 req/s, 8 slots) ...")
        sv = serve_run(512, 8, 400.0, args.budget_gb, args.chosen)
        serving = {k: sv[k] for k in ("metric", "value", "unit", "p50_ms", "p99_ms", "service_p50_ms",
                                      "queue_wait_p99_ms", "capacity_fits_per_s", "max_running",
                                      "peak_reserved_gb", "mix_vs_alone_max_dq_e", "chosen_vs_fp64_max_dq_e")}
        serving["config"] = sv["config"]
    def per_fit(run):
        nf = run["breakdown_fits"]
        return {k: {"ms_per_fit": round(v[0] / nf, 4), "launches_per_fit": v[1] / nf}
                for k, v in run["kernels"].items() if v[1]}
    kms, kms64 = per_fit(split), per_fit(f64)
    line = {
        "metric": "charge fits/s + p50/p99 latency, chosen vs plain FP64 GPU at <=1e-5 e",
        "value": round(fits_s, 4), "unit": "fits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * split["total"] / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 basis/solve/refinement, i8 slices -> i32 exact Gram" if args.chosen == "refined" else
                 "f64 basis/solve, i8 slices -> i32 exact Gram",
        "data": "synthetic (seeded exact Ewald ESP + 1e-5 noise)",
        "config": {"workload": args.workload, "n_atoms": N, "grid": [int(v) for v in system.n_grid],
                   "kept_points": M, "k_vectors": K,
                   "precision": "split (int8 Ozaki, 4 slices)" if args.chosen == "split" else
                   "refined (int8 Ozaki, 3 slices, fp64 residual refinement)"},
        "p50_ms": round(1e3 * float(np.median(split["lat"])), 3),
        "p99_ms": round(1e3 * float(np.percentile(split["lat"], 99)), 3),
        "fp64_fits_per_s": round(fits_s64, 4),
        "fp64_p50_ms": round(1e3 * float(np.median(f64["lat"])), 3),
        "speedup_vs_fp64_gpu": round(fits_s / fits_s64, 3),
        "fp64_factorized_fits_per_s": round(fits_s64f, 4),
        "fp64_factorized_p50_ms": round(1e3 * float(np.median(f64f["lat"])), 3),
        "speedup_vs_fp64_factorized": round(fits_s / fits_s64f, 3),
        "max_dq_vs_fp64_e": dq_vs_fp64, "within_1e-5": dq_vs_fp64 <= 1e-5,
        "max_dq_fp64_factorized_vs_fp64_e": float(np.max(np.abs(f64f["q"] - f64["q"]))),
        "error_vs_oracle": err_oracle,
        "end_to_end_p50_ms": {"chosen": round(1e3 * float(np.median(split["e2e"])), 2),
                              "fp64": round(1e3 * float(np.median(f64["e2e"])), 2),
                              "fp64_factorized": round(1e3 * float(np.median(f64f["e2e"])), 2),
                              "note": "from host arrays: cube upload, host Ewald setup, filter, accumulate and "
                                      "solve (rp_fit_reset + rp_fit_accumulate + rp_fit_solve); cube text "
                                      "parsing excluded"},
        "refine_steps_last_correction_e": split["refine"],
        "solve_report": split["report"],
        "kernels_chosen": kms, "kernels_fp64": kms64, "kernels_fp64_factorized": per_fit(f64f),
        "roofline": roof,
        "roofline_gram": roof_gram,
        "roofline_filter": roof_filter,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(fits_s / cpu["value"], 1) if cpu else None,
        "batch": batch,
        "serving": serving,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
