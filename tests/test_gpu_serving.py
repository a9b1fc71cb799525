"""Serving loop and fit reuse on the GPU, through the C ABI.

Bars: a fit reloaded with rp_fit_reset, and a request served in a mix of
concurrent requests, give bitwise the same charges as the same inputs fitted
alone on a fresh fit (every kernel is deterministic: no floating-point
atomics, fixed reduction orders); split-precision charges are within 1e-5 e
of fp64; device memory stays within the admission bound of each request.
"""
import os
import tempfile

import numpy as np
import pytest

from pyrepeat_amd import _abi, params, synth

pytestmark = pytest.mark.gpu

SPLIT = _abi.make_options(1, _abi.RP_PREC_OZAKI, 1.0, 1000.0, 20.0)
F64 = _abi.make_options(1, _abi.RP_PREC_FP64, 1.0, 1000.0, 20.0)
REFINED = _abi.make_options(1, _abi.RP_PREC_OZAKI_REFINED, 1.0, 1000.0, 20.0)


def _con(s):
    with tempfile.TemporaryDirectory() as td:
        symm = None
        if s.symm_groups:
            p = os.path.join(td, "symm.txt")
            with open(p, "w") as f:
                f.write(s.symm_text())
            symm = params.read_symmetry(p)
        return params.constraint_arrays(s.n_atoms, s.atomic_numbers, s.q_tot, symm, None, None)


def _args(s):
    return (s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube)


def _alone(s, con, precision):
    """Fresh context and fit: create (filter), accumulate, solve."""
    ctx = _abi.Context(0, 0)
    f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1, precision)
    f.accumulate()
    q, _ = f.solve(con)
    ne = f.normal_equations()
    f.close()
    ctx.close()
    return q, ne


@pytest.fixture(scope="module")
def stream12():
    reqs = synth.stream_requests(12)
    return [(s, _con(s)) for s in reqs]


@pytest.mark.parametrize("precision", [_abi.RP_PREC_OZAKI, _abi.RP_PREC_FP64, _abi.RP_PREC_OZAKI_REFINED])
def test_reset_matches_fresh_fit_without_growth(stream12, precision):
    big = max(stream12, key=lambda sc: int(np.prod(sc[0].n_grid)) * sc[0].n_atoms)
    small = min(stream12, key=lambda sc: int(np.prod(sc[0].n_grid)) * sc[0].n_atoms)
    print(f"reset case precision={precision} big={big[0].name} small={small[0].name}", flush=True)
    ctx = _abi.Context(0, 0)
    f = _abi.Fit(ctx, *_args(big[0]), _abi.RP_ORDER_CUBE, 1, precision)
    f.accumulate()
    f.solve(big[1])
    held = ctx.hbm_in_use()
    f.reset(*_args(small[0]), _abi.RP_ORDER_CUBE, 1, precision)
    f.accumulate()
    q, _ = f.solve(small[1])
    ne = f.normal_equations()
    # the smaller system reuses every chunk buffer; only the real-space shift list may grow
    # (a smaller cell has more images within the cutoff), by 24 bytes per shift
    assert 0 <= ctx.hbm_in_use() - held <= 24 * 125
    f.close()
    ctx.close()
    q0, ne0 = _alone(small[0], small[1], precision)
    np.testing.assert_array_equal(q, q0)
    for a, b in zip(ne, ne0):
        np.testing.assert_array_equal(a, b)


def test_request_bound_covers_device_memory(stream12):
    cases = [(stream12[0][0], stream12[0][1]), (synth.water(), None), (synth.crystal200(), None)]
    for s, con in cases:
        con = con or _con(s)
        for opts, prec in ((SPLIT, _abi.RP_PREC_OZAKI), (F64, _abi.RP_PREC_FP64),
                           (REFINED, _abi.RP_PREC_OZAKI_REFINED)):
            print(f"bound case {s.name} precision={prec}", flush=True)
            bound = _abi.Server.request_bytes(_args(s), con, opts)
            # a context capped at the bound: the fit must run within it (the refined form streams its
            # refinement when keeping the basis would not fit), --stats path included
            ctx = _abi.Context(0, bound)
            f = _abi.Fit(ctx, *_args(s), _abi.RP_ORDER_CUBE, 1, prec)
            f.accumulate()
            q, _ = f.solve(con)
            f.phi_matvec(q)
            used = ctx.hbm_in_use()
            f.close()
            ctx.close()
            assert used <= bound, (s.name, prec, used, bound)
            # and not absurdly loose (the bound takes the kept count as the whole grid)
            assert bound < 4 * used + (64 << 20), (s.name, prec, used, bound)
            if prec == _abi.RP_PREC_OZAKI_REFINED:
                # uncapped, the refined form keeps its basis on top of the bound's buffers
                ctx = _abi.Context(0, 0)
                f = _abi.Fit(ctx, *_args(s), _abi.RP_ORDER_CUBE, 1, prec)
                f.accumulate()
                q2, _ = f.solve(con)
                plan = _abi.chunk_plan(f.info.n_keep, s.n_atoms, f.info.n_k, prec)
                kept = 8 * plan["np"] * plan["rows"] * plan["n_chunks"]
                assert ctx.hbm_in_use() <= bound + kept
                f.close()
                ctx.close()
                np.testing.assert_array_equal(q2, q)


def test_server_mix_equals_alone(stream12):
    """Requests served concurrently (4 slots, all queued at once) equal the same requests run alone."""
    print("server mix case: 12 stream requests, 4 slots", flush=True)
    srv1 = _abi.Server(0, 1, 200 * 10**9)
    alone = []
    for s, con in stream12:
        q, res = srv1.poll(srv1.submit(_args(s), con, SPLIT), -1)
        assert res["status"] == 0, res["error"]
        alone.append(q)
    srv1.close()
    srv = _abi.Server(0, 4, 200 * 10**9)
    tickets = [srv.submit(_args(s), con, SPLIT) for s, con in stream12]
    mix = [srv.poll(t, -1) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(stream12) and st["failed"] == 0
    assert st["max_running"] >= 2, st          # the fits overlapped on the device
    assert st["peak_reserved"] <= st["budget"]
    for (s, con), qa, (qm, res) in zip(stream12, alone, mix):
        assert res["status"] == 0, res["error"]
        np.testing.assert_array_equal(qm, qa)
        assert abs(qm.sum() - s.q_tot) < 1e-9
    # the served result is the plain fit's, and split stays within 1e-5 e of fp64
    for i in (0, 5):
        s, con = stream12[i]
        q_plain, _ = _alone(s, con, _abi.RP_PREC_OZAKI)
        np.testing.assert_array_equal(mix[i][0], q_plain)
        q64, _ = _alone(s, con, _abi.RP_PREC_FP64)
        assert np.max(np.abs(mix[i][0] - q64)) <= 1e-5


def test_server_result_matches_oracle(oracle, stream12):
    s, con = min(stream12, key=lambda sc: int(np.prod(sc[0].n_grid)) * sc[0].n_atoms)
    print(f"server oracle case {s.name} N={s.n_atoms} grid={tuple(s.n_grid)}", flush=True)
    srv = _abi.Server(0, 2, 200 * 10**9)
    q, res = srv.poll(srv.submit(_args(s), con, F64), -1)
    srv.close()
    assert res["status"] == 0, res["error"]
    symm = None
    if s.symm_groups:
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "symm.txt")
            with open(p, "w") as f:
                f.write(s.symm_text())
            symm = oracle.read_symmetry(p)
    ref = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat().astype(np.float64),
                     q_tot=s.q_tot, symm=symm)
    assert np.max(np.abs(q - ref.charges)) <= 1e-6


def test_server_budget_admission_and_rejection(stream12):
    reqs = stream12[:6]
    bounds = [_abi.Server.request_bytes(_args(s), con, SPLIT) for s, con in reqs]
    budget = int(1.2 * max(bounds))
    print(f"budget case budget={budget}", flush=True)
    srv = _abi.Server(0, 4, budget)
    tickets = [srv.submit(_args(s), con, SPLIT) for s, con in reqs]
    out = [srv.poll(t, -1) for t in tickets]
    st = srv.stats()
    assert st["peak_reserved"] <= budget
    for (q, res), (s, con) in zip(out, reqs):
        assert res["status"] == 0, res["error"]
        assert res["bytes_bound"] <= budget
    # a request whose bound exceeds the whole budget is refused at submission
    small = _abi.Server(0, 1, max(bounds) - 1)
    big = reqs[int(np.argmax(bounds))]
    with pytest.raises(_abi.RepeatHipError) as e:
        small.submit(_args(big[0]), big[1], SPLIT)
    assert e.value.code == _abi.RP_EBUDGET
    small.close()
    srv.close()


def test_server_failed_request_does_not_stop_the_stream(stream12):
    """A request whose vdW filter keeps no point fails alone; the next request on that slot is correct."""
    s, con = stream12[0]
    empty = _abi.make_options(1, _abi.RP_PREC_OZAKI, 50.0, 1000.0, 20.0)   # r_min covers the whole cell
    print("server error-path case", flush=True)
    srv = _abi.Server(0, 1, 200 * 10**9)
    t_bad = srv.submit(_args(s), con, empty)
    t_good = srv.submit(_args(s), con, SPLIT)
    q_bad, r_bad = srv.poll(t_bad, -1)
    q_good, r_good = srv.poll(t_good, -1)
    st = srv.stats()
    srv.close()
    assert r_bad["status"] == _abi.RP_EEMPTY and "kept no grid point" in r_bad["error"]
    assert r_good["status"] == 0
    q0, _ = _alone(s, con, _abi.RP_PREC_OZAKI)
    np.testing.assert_array_equal(q_good, q0)
    assert st["failed"] == 1 and st["completed"] == 2
