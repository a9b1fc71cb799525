"""GPU parity at the two largest configurations, through the C ABI.

mof1000 (1000 atoms, 160^3 grid, RESP on H, the headline configuration): the keep mask and kept count equal
the oracle's (digest of the packed mask), every precision's charges match the oracle's golden charges
(1e-6 e for the fp64 and refined forms, 1e-5 e for the split form), the total charge holds to 1e-9 and the
RMS of the fit matches the oracle's to 1e-8 a.u.

large5000 (5000 atoms, 232^3 grid, 943k kept points): beyond the oracle's fit in test time, so its golden
holds the filter (kept count, mask digest, mean ESP over the kept points); the refined charges match the
device's plain fp64 path within 1e-5 e.
"""
import hashlib

import numpy as np
import pytest

from pyrepeat_amd import _abi, params, synth

from conftest import golden
from test_gpu_parity import _calc

pytestmark = pytest.mark.gpu


def _digest(mask):
    return hashlib.sha256(np.packbits(np.asarray(mask, dtype=np.uint8)).tobytes()).hexdigest()


def test_mof1000_mask_charges_and_rms(gpu_ctx, tmp_path):
    g = golden("mof1000")
    s = synth.mof1000()
    ref = np.array(g["charges"])
    assert g["rms"] is not None
    for precision, tol in (("refined", 1e-6), ("fp64", 1e-6), ("split", 1e-5)):
        print(f"mof1000 {precision}", flush=True)
        calc = _calc(gpu_ctx, s, precision, tmp_path)
        if precision == "refined":
            _, _, idx, mask = calc._fit.grid(mask=True)
            assert calc._fit.info.n_keep == g["n_keep"] == len(idx)
            assert _digest(mask) == g["mask_sha256"]
            assert calc._fit.info.n_k == g["n_k"] and calc._fit.info.n_shift_pruned == g["n_shift_pruned"]
        q = calc.fit_charges()
        dq = np.max(np.abs(q - ref))
        st = calc.compute_stats(q, need_v_coul=False)
        print(f"  max|dq| vs oracle {dq:.2e}, rms {st['rms_error']:.12e} (oracle {g['rms']:.12e}), "
              f"report {calc.solve_report}", flush=True)
        assert dq <= tol
        assert abs(q.sum() - s.q_tot) <= 1e-9
        assert abs(st["rms_error"] - g["rms"]) <= 1e-8
        calc.close()


def test_large5000_mask_and_refined_vs_fp64(gpu_ctx):
    g = golden("large5000")
    print("large5000: generating the 232^3 cube", flush=True)
    s = synth.large5000()
    con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None)
    out = {}
    for prec in (_abi.RP_PREC_OZAKI_REFINED, _abi.RP_PREC_FP64):
        print(f"large5000 precision {prec}", flush=True)
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1, prec)
        if prec == _abi.RP_PREC_OZAKI_REFINED:
            _, _, idx, mask = f.grid(mask=True)
            assert f.info.n_keep == g["n_keep"] and f.info.n_eff == g["n_eff"]
            assert _digest(mask) == g["mask_sha256"]
            assert abs(f.info.v_mean - g["v_mean"]) <= max(1e-12 * abs(g["v_mean"]), 1e-14)
        f.accumulate()
        q, _ = f.solve(con)
        out[prec] = (q, f.solve_report())
        f.close()
    dq = np.max(np.abs(out[_abi.RP_PREC_OZAKI_REFINED][0] - out[_abi.RP_PREC_FP64][0]))
    print(f"  refined - fp64 {dq:.2e}; reports {out[_abi.RP_PREC_OZAKI_REFINED][1]} "
          f"{out[_abi.RP_PREC_FP64][1]}", flush=True)
    assert dq <= 1e-5
    for q, _ in out.values():
        assert abs(q.sum()) <= 1e-9
