"""Writes oracle fixtures for the seeded workloads (run from the repo root, CPU only).

    python tests/golden/make_golden.py water crystal200 mof1000 large5000

Each ``<name>.json`` holds the fit settings, the kept-point count, a checksum of
the keep mask, the oracle's charges (full precision) and its RMS errors.  The
inputs are regenerated from their seeds by ``pyrepeat_amd.synth``; the oracle
is ``oracle/repeat_oracle.py`` (CPU restatement of the reference).

The 5000-atom workload (943k kept points x 5000 atoms x 754 k-vectors) is beyond
the oracle's fit in reasonable time, so its fixture holds the filter only: the
kept count, the mask checksum, the effective-atom count and the mean ESP over
the kept points.  The 1000-atom workload keeps its basis (2 GB) to give the RMS.
"""
import hashlib
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import repeat_oracle as O  # noqa: E402
from pyrepeat_amd import synth  # noqa: E402


def mask_digest(mask):
    return hashlib.sha256(np.packbits(np.asarray(mask, dtype=np.uint8)).tobytes()).hexdigest()


FILTER_ONLY = {"large5000"}
STATS_ALWAYS = {"mof1000"}


def run_filter_only(name, s):
    t0 = time.time()
    box, _ = O.box_and_volume(s.n_grid, s.axis)
    eff, rmin, rmax, _ = O.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, 1.0, 1000.0)
    keep = O.filter_mask(s.n_grid, s.axis, eff, rmin, rmax)
    _, V, flat = O.compact(keep, s.esp_flat(), s.n_grid, s.axis)
    v_raw = np.asarray(s.esp_flat())[flat].astype(np.float64)
    out = {
        "workload": name, "n_grid": [int(v) for v in s.n_grid], "n_atoms": int(s.n_atoms), "q_tot": s.q_tot,
        "filter_only": True, "n_eff": int(len(eff)), "n_keep": int(len(flat)), "mask_sha256": mask_digest(keep),
        "v_mean": float(np.mean(v_raw)), "oracle_seconds": round(time.time() - t0, 1),
    }
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, "(filter only) done in", out["oracle_seconds"], "s; M =", out["n_keep"], flush=True)


def run(name):
    s = getattr(synth, name)()
    if name in FILTER_ONLY:
        return run_filter_only(name, s)
    symm = resp = None
    with tempfile.TemporaryDirectory() as td:
        if s.symm_groups:
            p = os.path.join(td, "symm.txt")
            open(p, "w").write(s.symm_text())
            symm = O.read_symmetry(p)
        if s.resp_text:
            p = os.path.join(td, "resp.txt")
            open(p, "w").write(s.resp_text)
            resp = O.read_resp_params(p)
    stats = name in STATS_ALWAYS or s.n_atoms * 1.0 * np.prod(s.n_grid) < 5e8
    t0 = time.time()
    r = O.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), q_tot=s.q_tot, symm=symm, resp=resp,
              stats=stats)
    out = {
        "workload": name, "n_grid": [int(v) for v in s.n_grid], "n_atoms": int(s.n_atoms),
        "q_tot": s.q_tot, "symmetry_groups": len(s.symm_groups), "resp": s.resp_text,
        "n_keep": int(len(r.V)), "mask_sha256": mask_digest(r.mask),
        "n_k": int(len(r.ewald.kvecs)), "n_shift": int(len(r.ewald.shifts)),
        "n_shift_pruned": int(len(r.ewald.shifts_pruned)),
        "charges": [float(q) for q in r.charges], "used_fallback": bool(r.used_fallback),
        "rms": r.rms, "rms_corrected": r.rms_corrected, "sum_V": float(r.sum_V),
        "oracle_seconds": round(time.time() - t0, 1),
    }
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, "done in", out["oracle_seconds"], "s; M =", out["n_keep"], flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:]:
        run(n)
