"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: keep mask bit-exact; kept points and ESP identical; charges within
1e-6 e (fp64 path) or 1e-5 e (split path) of the oracle; RMS within 1e-8 a.u.
"""
import json
import os

import numpy as np
import pytest

from pyrepeat_amd import _abi, synth
from pyrepeat_amd.calculator import ChargeCalculator
from pyrepeat_amd.cube import cube_from_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fit(ctx, s, precision="fp64", vdw_factor=1.0, vdw_max=1000.0, fit_type=1):
    prec = _abi.RP_PREC_FP64 if precision == "fp64" else _abi.RP_PREC_OZAKI
    return _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, fit_type,
                    prec, vdw_factor, vdw_max)


def _oracle_mask(O, s, vdw_factor=1.0, vdw_max=1000.0):
    box, _ = O.box_and_volume(s.n_grid, s.axis)
    eff, rmin, rmax, _ = O.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, vdw_factor, vdw_max)
    return O.filter_mask(s.n_grid, s.axis, eff, rmin, rmax), len(eff)


FILTER_CASES = [("water", None, 1.0, 1000.0), ("crystal200", None, 1.0, 1000.0),
                ("tiny", 0, 1.0, 1000.0), ("tiny", 1, 1.0, 1000.0), ("tiny", 2, 1.0, 1.6), ("tiny", 3, 0.8, 1.3),
                ("tiny", 5, 1.4, 2.5)]


@pytest.mark.parametrize("name,seed,vf,vm", FILTER_CASES)
def test_filter_mask_bit_exact(oracle, gpu_ctx, name, seed, vf, vm):
    s = synth.tiny(seed=seed, n_atoms=5, n=14, L=10.0, triclinic=seed % 2 == 1) if name == "tiny" \
        else getattr(synth, name)()
    print(f"filter case {name} seed={seed} vdw {vf}/{vm}", flush=True)
    ref, n_eff = _oracle_mask(oracle, s, vf, vm)
    f = _fit(gpu_ctx, s, vdw_factor=vf, vdw_max=vm)
    xyz, v, idx, mask = f.grid(mask=True)
    assert f.info.n_eff == n_eff
    np.testing.assert_array_equal(mask, ref)
    pos, vref, flat = oracle.compact(ref, s.esp_flat(), s.n_grid, s.axis)
    np.testing.assert_array_equal(idx, flat)
    np.testing.assert_array_equal(xyz, pos)
    np.testing.assert_allclose(v, vref, rtol=0, atol=1e-15)
    f.close()


def test_filter_tie_exact_radius(gpu_ctx):
    """A point at distance exactly r_min is dropped; the device matches the strict '>' of the reference."""
    L, n = 8.0, 8
    axis = np.eye(3) * (L / n)
    R = 2.72687  # H
    # vdw factor with fl(f * R) == 5.0 exactly (a 3-4-5 point is at distance 5)
    f = 5.0 / R
    for _ in range(64):
        if f * R == 5.0:
            break
        f = np.nextafter(f, 10.0 if f * R < 5.0 else 0.0)
    assert f * R == 5.0
    esp = np.arange(n ** 3, dtype=np.float64)
    fit = _abi.Fit(gpu_ctx, [n, n, n], axis, [1], np.array([[0.0, 0.0, 0.0]]), esp, _abi.RP_ORDER_FLAT, 1,
                   _abi.RP_PREC_FP64, f, 1000.0)
    _, _, idx, mask = fit.grid(mask=True)
    m = mask.reshape(n, n, n)  # [k][j][i]
    assert m[0, 4, 3] == 0 and m[0, 3, 4] == 0
    assert m[0, 4, 4] == 1
    fit.close()


def test_ewald_setup_matches_oracle(oracle, gpu_ctx):
    for s in (synth.water(), synth.crystal200()):
        f = _fit(gpu_ctx, s)
        kv, kc, sh, sp = f.ewald()
        box, vol = oracle.box_and_volume(s.n_grid, s.axis)
        ew = oracle.ewald_setup(box, vol, 20.0, 1)
        np.testing.assert_allclose(kv, ew.kvecs, rtol=1e-15, atol=1e-16)
        np.testing.assert_allclose(kc, ew.kcoefs, rtol=1e-13)
        np.testing.assert_array_equal(sh, ew.shifts)
        xyz, _, _ = f.grid()
        np.testing.assert_array_equal(sp, oracle.prune_shifts(ew.shifts, xyz, box, s.atom_pos, 20.0))
        assert abs(f.info.volume - vol) <= 1e-12 * abs(vol)
        f.close()


def test_phi_at_points_matches_oracle(oracle, gpu_ctx):
    s = synth.crystal200()
    f = _fit(gpu_ctx, s)
    rng = np.random.default_rng(0)
    g = rng.random((64, 3)) @ s.box
    ref_ew = oracle.ewald_setup(*oracle.box_and_volume(s.n_grid, s.axis), 20.0, 1)
    ref = oracle.phi_matrix(g, s.atom_pos, ref_ew)
    got = f.phi(g)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13)
    f.close()


@pytest.mark.parametrize("name", ["water", "tiny3"])
def test_normal_equations_match_oracle(oracle, gpu_ctx, name):
    s = synth.water() if name == "water" else synth.tiny(seed=3, n_atoms=6, n=16, L=10.0, triclinic=True)
    f = _fit(gpu_ctx, s)
    f.accumulate()
    ata, atb, sp, sv = f.normal_equations()
    box, vol = oracle.box_and_volume(s.n_grid, s.axis)
    ew = oracle.ewald_setup(box, vol, 20.0, 1)
    xyz, v, _ = f.grid()
    ATA, ATb, sum_phi, sum_V, _ = oracle.accumulate(xyz, v, s.atom_pos, ew)
    np.testing.assert_allclose(ata, ATA, rtol=1e-11)
    np.testing.assert_allclose(sp, sum_phi, rtol=1e-11)
    np.testing.assert_allclose(atb, ATb, rtol=1e-9, atol=1e-12)
    assert abs(sv - sum_V) < 1e-9
    f.close()


def _golden(name):
    with open(os.path.join(GOLDEN, f"{name}.json")) as fh:
        return json.load(fh)


def _calc(ctx, s, precision, tmp_path):
    symm = resp = None
    if s.symm_groups:
        symm = tmp_path / "symm.txt"
        symm.write_text(s.symm_text())
    if s.resp_text:
        resp = tmp_path / "resp.txt"
        resp.write_text(s.resp_text)
    return ChargeCalculator(cube=cube_from_system(s), q_tot=s.q_tot, symm_file=str(symm) if symm else None,
                            resp_file=str(resp) if resp else None, precision=precision, context=ctx)


@pytest.mark.parametrize("name", ["water", "crystal200"])
def test_fit_fp64_matches_golden(gpu_ctx, tmp_path, name):
    g = _golden(name)
    s = getattr(synth, name)()
    print(f"fit fp64 {name}", flush=True)
    calc = _calc(gpu_ctx, s, "fp64", tmp_path)
    assert calc.n_grid_points == g["n_keep"]
    q = calc.fit_charges()
    dq = np.max(np.abs(q - np.array(g["charges"])))
    print(f"  max|dq| = {dq:.3e}", flush=True)
    assert dq <= 1e-6
    assert abs(q.sum() - s.q_tot) < 1e-9
    if g.get("rms") is not None:
        st = calc.compute_stats(q, need_v_coul=True)
        assert abs(st["rms_error"] - g["rms"]) <= 1e-8
        assert abs(calc.rms_from_normal_equations(q) - g["rms"]) <= 1e-8
    calc.close()


def test_stats_match_oracle_small(oracle, gpu_ctx):
    s = synth.tiny(seed=6, n_atoms=4, n=14, L=10.0, triclinic=True)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), stats=True)
    calc = ChargeCalculator(cube=cube_from_system(s), precision="fp64", context=gpu_ctx)
    q = calc.fit_charges()
    np.testing.assert_allclose(q, r.charges, atol=1e-9)
    st = calc.compute_stats(q)
    np.testing.assert_allclose(st["V_coul"], r.v_coul, rtol=1e-10, atol=1e-12)
    assert abs(st["rms_error"] - r.rms) < 1e-10
    calc.close()


def test_symmetry_fit_small_matches_oracle(oracle, gpu_ctx, tmp_path):
    s = synth.tiny(seed=12, n_atoms=8, n=16, L=11.0, triclinic=True)
    p = tmp_path / "symm.txt"
    p.write_text("1,3\n2-2\n4,6,8\n")
    symm = oracle.read_symmetry(str(p))
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), q_tot=1.0, symm=symm, stats=True)
    calc = ChargeCalculator(cube=cube_from_system(s), q_tot=1.0, symm_file=str(p), precision="fp64", context=gpu_ctx)
    q = calc.fit_charges()
    np.testing.assert_allclose(q, r.charges, atol=1e-9)
    assert q[0] == q[2] and q[3] == q[5] == q[7]
    st = calc.compute_stats(q, need_v_coul=False)
    assert abs(st["rms_error"] - r.rms) < 1e-10          # reference RMS (sum_phi quirk under symmetry)
    assert abs(calc.rms_from_normal_equations(q, corrected=True) - r.rms_corrected) < 1e-10
    calc.close()


def test_fit_type0_matches_oracle(oracle, gpu_ctx):
    s = synth.tiny(seed=13, n_atoms=4, n=14, L=10.0)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), fit_flag=0)
    calc = ChargeCalculator(cube=cube_from_system(s), fit_flag=0, precision="fp64", context=gpu_ctx)
    np.testing.assert_allclose(calc.fit_charges(), r.charges, atol=1e-9)
    calc.close()


@pytest.mark.parametrize("name", ["water", "crystal200"])
def test_fp64_factorized_matches_fp64(gpu_ctx, tmp_path, name):
    """The fp64 form with the factorized basis: its normal equations equal the plain fp64 path's to fp64
    rounding (the same basis by another summation), its charges the golden charges within 1e-6 e."""
    g = _golden(name)
    s = getattr(synth, name)()
    print(f"fp64-factorized {name}", flush=True)
    out = {}
    for prec in (_abi.RP_PREC_FP64, _abi.RP_PREC_FP64_FACT):
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1, prec)
        f.accumulate()
        out[prec] = f.normal_equations()
        f.close()
    (a64, b64, s64, v64), (af, bf, sf, vf) = out[_abi.RP_PREC_FP64], out[_abi.RP_PREC_FP64_FACT]
    np.testing.assert_allclose(af, a64, rtol=1e-11, atol=1e-13 * np.abs(a64).max())
    np.testing.assert_allclose(bf, b64, rtol=1e-11, atol=1e-13 * np.abs(b64).max())
    np.testing.assert_allclose(sf, s64, rtol=1e-11, atol=1e-13 * np.abs(s64).max())
    calc = _calc(gpu_ctx, s, "fp64-factorized", tmp_path)
    q = calc.fit_charges()
    calc.close()
    dq = np.max(np.abs(q - np.array(g["charges"])))
    print(f"  max|dq| vs golden {dq:.2e}", flush=True)
    assert dq <= 1e-6


def _duplicate_atom_case(oracle):
    """Three atoms, the second at the first's place: identical basis columns, an exactly singular KKT
    matrix.  Its minimum-norm solution (the reference's lstsq branch) is the fit without the second atom,
    the first atom's charge split equally between the two."""
    s = synth.tiny(seed=14, n_atoms=3, n=12, L=9.0)
    pos = s.atom_pos.copy()
    pos[1] = pos[0]
    keep = [0, 2]
    merged = oracle.fit(s.n_grid, s.axis, s.atomic_numbers[keep], pos[keep], s.esp_flat())
    expect = np.array([merged.charges[0] / 2, merged.charges[0] / 2, merged.charges[1]])
    cube = {"n_atoms": 3, "n_grid": s.n_grid, "atom_index": s.atomic_numbers, "atom_pos": pos, "axis_vector": s.axis,
            "esp_cube": s.esp_cube}
    return cube, expect


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("split", 1e-5)])
def test_singular_kkt_least_squares(oracle, gpu_ctx, capfd, precision, tol):
    """An exactly singular KKT matrix (two atoms at one place) is solved by the minimum-norm least
    squares, as the reference's lstsq fallback (repeat.py:2306-2310), with its warning."""
    cube, expect = _duplicate_atom_case(oracle)
    print(f"singular KKT, {precision}: lstsq fallback", flush=True)
    calc = ChargeCalculator(cube=cube, precision=precision, context=gpu_ctx)
    q = calc.fit_charges()
    rep = calc._fit.solve_report()
    calc.close()
    print(f"  q {q}, expected {expect}, report {rep}", flush=True)
    assert rep["solver"] == _abi.RP_SOLVER_LSTSQ
    assert "least squares fallback" in capfd.readouterr().err
    np.testing.assert_allclose(q, expect, rtol=0, atol=tol)
    assert abs(q[0] - q[1]) <= 1e-9
    assert abs(q.sum()) < 1e-9


def test_backsub_wait_limit_falls_back_to_lu(gpu_ctx, tmp_path):
    """The one-launch backward substitution waits on flags with a bounded poll count.  With the
    limit forced to zero from the host (RP_BACKSUB_SPIN=0, a fresh process), every wait gives up
    without polling; the solve must notice and answer through the LU, with the Cholesky's charges."""
    import subprocess
    import sys
    script = tmp_path / "spin.py"
    script.write_text(
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {str(ROOT)!r})\n"
        "from pyrepeat_amd import _abi, params, synth\n"
        "s = synth.crystal200()\n"
        "ctx = _abi.Context(0, 0)\n"
        "f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,\n"
        "             _abi.RP_PREC_FP64)\n"
        "f.accumulate()\n"
        "q, _ = f.solve(params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None))\n"
        "np.save(sys.argv[1], q)\n"
        "print('lu', int(f.last_solve_used_lu))\n")
    out = {}
    for spin in ("0", None):
        env = dict(os.environ)
        env.pop("RP_BACKSUB_SPIN", None)
        if spin is not None:
            env["RP_BACKSUB_SPIN"] = spin
        path = tmp_path / f"q_{spin}.npy"
        print(f"backsub wait limit {spin}", flush=True)
        r = subprocess.run([sys.executable, str(script), str(path)], env=env, capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        out[spin] = (np.load(path), r.stdout.split()[-1])
    assert out["0"][1] == "1" and out[None][1] == "0"
    assert np.max(np.abs(out["0"][0] - out[None][0])) <= 1e-8


def test_backsub_wait_limit_refined_refines_the_lu(gpu_ctx, tmp_path):
    """Refined form with every triangular-solve wait giving up (RP_BACKSUB_SPIN=0, a fresh process): the
    solve falls back to LU, and the LU solution of the 3-slice system is refined against the kept fp64
    basis with LU corrections, so the charges still meet the fp64 path's (instead of the 3-slice
    system's 1e-4 e)."""
    import subprocess
    import sys
    script = tmp_path / "spin_refined.py"
    script.write_text(
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {str(ROOT)!r})\n"
        "from pyrepeat_amd import _abi, params, synth\n"
        "s = synth.crystal200()\n"
        "ctx = _abi.Context(0, 0)\n"
        "con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None)\n"
        "out = []\n"
        "for prec in (_abi.RP_PREC_OZAKI_REFINED, _abi.RP_PREC_FP64):\n"
        "    f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1, prec)\n"
        "    f.accumulate()\n"
        "    q, _ = f.solve(con)\n"
        "    out.append(q)\n"
        "    if prec == _abi.RP_PREC_OZAKI_REFINED:\n"
        "        rep = f.solve_report()\n"
        "    f.close()\n"
        "np.save(sys.argv[1], np.array(out))\n"
        "print('report', rep)\n"
        "print('solver', rep['solver'], 'steps', rep['refine_steps'], 'within', int(rep['within_target']))\n")
    env = dict(os.environ)
    env["RP_BACKSUB_SPIN"] = "0"
    path = tmp_path / "q_refined.npy"
    print("refined, backsub wait limit 0", flush=True)
    r = subprocess.run([sys.executable, str(script), str(path)], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    print(r.stdout, flush=True)
    tail = r.stdout.split()
    assert tail[-5] == str(_abi.RP_SOLVER_LU) and int(tail[-3]) >= 1 and tail[-1] == "1"
    q_ref, q64 = np.load(path)
    assert np.max(np.abs(q_ref - q64)) <= 1e-8


def test_solve_large_n_matches_dense_solve(oracle, gpu_ctx):
    """N' = 1600 takes the Cholesky path for large systems: while the trailing triangle has more
    64 x 64 tiles than the GPU has CUs, the panel is its own launch and the update reads it; the last
    steps fuse them; the backward substitution runs 25 chained blocks.  The device's normal equations,
    centred and solved densely by the oracle (LU), must give the device solve's charges."""
    from pyrepeat_amd import params
    s = synth.tiny(seed=21, n_atoms=1600, n=76, L=90.0)
    print("large-N solve: 1600 atoms, 76x77x78 grid", flush=True)
    f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                 _abi.RP_PREC_FP64)
    try:
        f.accumulate()
        ata, atb, sp, sv = f.normal_equations()
        M = int(f.info.n_keep)
        assert M > 4 * s.n_atoms
        A, b = oracle.centre(ata, atb, sp, sv, M)
        sol, _ = oracle.solve_kkt(A, b, np.ones(s.n_atoms), 0.0)
        q, _ = f.solve(params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None))
        assert not f.last_solve_used_lu
        assert np.max(np.abs(q - sol[:s.n_atoms])) <= 1e-6
    finally:
        f.close()


def test_deterministic_bitwise(gpu_ctx):
    s = synth.crystal200()
    qs = []
    for _ in range(2):
        calc = ChargeCalculator(cube=cube_from_system(s), precision="fp64", context=gpu_ctx)
        qs.append(calc.fit_charges())
        calc.close()
    np.testing.assert_array_equal(qs[0], qs[1])


# ---------------------------------------------------------------- split precision path
# (factorized fp64 reciprocal sum, fp64 real space, int8-sliced Gram with exact int32 groups)

def test_fit_type0_split_matches_oracle(oracle, gpu_ctx):
    """Bare 1/r basis (no reciprocal part) through the split path."""
    s = synth.tiny(seed=13, n_atoms=4, n=14, L=10.0)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), fit_flag=0)
    calc = ChargeCalculator(cube=cube_from_system(s), fit_flag=0, precision="split", context=gpu_ctx)
    assert np.max(np.abs(calc.fit_charges() - r.charges)) <= 1e-5
    calc.close()


@pytest.mark.parametrize("name", ["water", "crystal200"])
def test_fit_split_matches_golden(gpu_ctx, tmp_path, name):
    g = _golden(name)
    s = getattr(synth, name)()
    print(f"fit split {name}", flush=True)
    calc = _calc(gpu_ctx, s, "split", tmp_path)
    q = calc.fit_charges()
    dq = np.max(np.abs(q - np.array(g["charges"])))
    print(f"  max|dq| = {dq:.3e}", flush=True)
    assert dq <= 1e-5
    assert abs(q.sum() - s.q_tot) < 1e-9
    calc.close()


def test_split_normal_equations_close_to_fp64(gpu_ctx):
    s = synth.tiny(seed=3, n_atoms=6, n=16, L=10.0, triclinic=True)
    out = []
    for prec in ("fp64", "split"):
        f = _fit(gpu_ctx, s, precision=prec)
        f.accumulate()
        out.append(f.normal_equations())
        f.close()
    (a64, b64, s64, v64), (a8, b8, s8, v8) = out
    np.testing.assert_allclose(a8, a64, rtol=2e-6, atol=1e-9 * np.abs(a64).max())
    np.testing.assert_allclose(s8, s64, rtol=2e-6)
    # sum(V) of the sliced V column: 4 slices hold V to 2^-29 of max|V| per point
    xyz, v, _ = _fit(gpu_ctx, s).grid()
    assert abs(v8 - v64) <= len(v) * np.abs(v).max() * 2.0 ** -28


def test_split_deterministic_and_refit(gpu_ctx, tmp_path):
    s = synth.crystal200()
    calc = _calc(gpu_ctx, s, "split", tmp_path)
    q1 = calc.fit_charges()
    q2 = calc._fit.refit(calc.constraints())
    np.testing.assert_array_equal(q1, q2)
    calc.close()


def test_refit_fp64_equals_fit_charges(gpu_ctx, tmp_path):
    s = synth.water()
    calc = _calc(gpu_ctx, s, "fp64", tmp_path)
    q1 = calc.fit_charges()
    q2 = calc._fit.refit(calc.constraints())
    np.testing.assert_array_equal(q1, q2)
    calc.close()
