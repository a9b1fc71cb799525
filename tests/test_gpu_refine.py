"""Split form with iterative refinement (RP_PREC_OZAKI_REFINED) on the GPU, through the C ABI.

The Gram comes from 3 int8 slices per basis element; each refinement step forms the residual of
the fp64 normal equations from the fp64 basis kept in HBM and corrects the solve with the split
system's Cholesky factor.  The charges converge to the fp64 system's solution, so the bars are
the fp64 path's: within 1e-6 e of the oracle's golden charges, within 1e-9 e of the oracle on
small systems (symmetry with merge multiplicities, RESP, non-zero total charge), and within
rounding (1e-8 e at condition numbers up to ~2e6) of the device's own fp64 path.
"""
import numpy as np
import pytest

from pyrepeat_amd import _abi, params, synth
from pyrepeat_amd.calculator import ChargeCalculator
from pyrepeat_amd.cube import cube_from_system

from test_gpu_parity import _calc, _golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["water", "crystal200"])
def test_refined_matches_golden_and_fp64(gpu_ctx, tmp_path, name):
    g = _golden(name)
    s = getattr(synth, name)()
    print(f"fit refined {name}", flush=True)
    calc = _calc(gpu_ctx, s, "refined", tmp_path)
    q = calc.fit_charges()
    steps, last = calc._fit.refine_stats()
    dq = np.max(np.abs(q - np.array(g["charges"])))
    print(f"  max|dq| vs golden = {dq:.3e}, steps {steps}, last correction {last:.2e}", flush=True)
    assert not calc._fit.last_solve_used_lu
    assert 1 <= steps <= 2
    assert dq <= 1e-6
    assert abs(q.sum() - s.q_tot) < 1e-9
    calc.close()
    c64 = _calc(gpu_ctx, s, "fp64", tmp_path)
    q64 = c64.fit_charges()
    c64.close()
    assert np.max(np.abs(q - q64)) <= 1e-8


@pytest.mark.parametrize("symm_text,q_tot,resp", [
    ("1,3\n2-2\n4,6,8\n", 1.0, None),
    ("1,3\n1,3\n5,7\n", 0.0, "H 0.1 0.0\nC 0.05 -0.1\n"),   # atom 3 twice in set 1: multiplicity 2
])
def test_refined_symmetry_resp_matches_oracle(oracle, gpu_ctx, tmp_path, symm_text, q_tot, resp):
    s = synth.tiny(seed=12, n_atoms=8, n=16, L=11.0, triclinic=True)
    p = tmp_path / "symm.txt"
    p.write_text(symm_text)
    rp = None
    if resp:
        rp = tmp_path / "resp.txt"
        rp.write_text(resp)
    print(f"refined symmetry case {symm_text!r} q_tot={q_tot} resp={resp is not None}", flush=True)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), q_tot=q_tot,
                   symm=oracle.read_symmetry(str(p)), resp=oracle.read_resp_params(str(rp)) if rp else None)
    calc = ChargeCalculator(cube=cube_from_system(s), q_tot=q_tot, symm_file=str(p),
                            resp_file=str(rp) if rp else None, precision="refined", context=gpu_ctx)
    q = calc.fit_charges()
    assert calc._fit.refine_stats()[0] >= 1
    np.testing.assert_allclose(q, r.charges, atol=1e-9)
    calc.close()


@pytest.mark.parametrize("n_atoms,r", [(20, 0), (24, 0), (50, 1), (64, 0)])
def test_refined_small_serving_cells(oracle, gpu_ctx, tmp_path, n_atoms, r):
    """Serving-stream cells of 20-64 atoms keep 10-56k points; their 3-slice plans once cut a chunk into
    row splits of 1088 rows, which the Gram's 128-row stages do not divide, and the fit failed.  The
    refined charges match the oracle (total charge and symmetry groups of the request included) and
    the device's fp64 path."""
    s = synth.stream_request(r, n_atoms=n_atoms)
    symm = None
    if s.symm_groups:
        symm = tmp_path / "symm.txt"
        symm.write_text(s.symm_text())
    print(f"refined serving cell: {n_atoms} atoms, request {r}, grid {list(s.n_grid)}, "
          f"q_tot {s.q_tot}, {len(s.symm_groups)} groups", flush=True)
    ref = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), q_tot=s.q_tot,
                     symm=oracle.read_symmetry(str(symm)) if symm else None)
    out = {}
    for precision in ("refined", "fp64"):
        calc = _calc(gpu_ctx, s, precision, tmp_path)
        out[precision] = calc.fit_charges()
        plan = _abi.chunk_plan(calc._fit.info.n_keep, s.n_atoms, calc._fit.info.n_k,
                               _abi.RP_PREC_OZAKI_REFINED)
        calc.close()
    print(f"  plan {plan}", flush=True)
    dq = np.max(np.abs(out["refined"] - ref.charges))
    print(f"  max|dq| refined vs oracle {dq:.2e}, vs fp64 {np.max(np.abs(out['refined'] - out['fp64'])):.2e}",
          flush=True)
    assert dq <= 1e-6
    assert np.max(np.abs(out["refined"] - out["fp64"])) <= 1e-8
    assert abs(out["refined"].sum() - s.q_tot) < 1e-9


def test_refined_flags_ill_conditioned_cell(gpu_ctx, capfd):
    """A 400-atom cluster in a 36 Bohr box: points only outside it, cond(A) ~ 2e12, unrestrained charges of
    hundreds of e.  Any two fp64 solutions of this system differ far above 1e-5 e (the refined and the
    plain fp64 path by ~1e-2 e), so the refined charges either meet the fp64 path's within 1e-5 e or the
    solve report flags the fit: its sensitivity (charge change under an 8-ulp perturbation of the system)
    must cover the gap, and the calculator warns."""
    s = synth.cluster(400)
    con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None)
    print("ill-conditioned cluster: 400 atoms, 64^3 grid", flush=True)
    out = {}
    for prec in (_abi.RP_PREC_FP64, _abi.RP_PREC_OZAKI_REFINED):
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1, prec)
        f.accumulate()
        q, _ = f.solve(con)
        out[prec] = (q, f.solve_report())
        if prec == _abi.RP_PREC_FP64:
            ata, _, sp, _ = f.normal_equations()
            ev = np.linalg.eigvalsh(ata - np.outer(sp, sp) / f.info.n_keep)
        f.close()
    cond = ev[-1] / ev[0]
    dq = np.max(np.abs(out[_abi.RP_PREC_FP64][0] - out[_abi.RP_PREC_OZAKI_REFINED][0]))
    rep = out[_abi.RP_PREC_OZAKI_REFINED][1]
    print(f"  cond(A) {cond:.2e}, refined - fp64 {dq:.2e}, report {rep}", flush=True)
    assert cond >= 1e11
    assert dq <= 1e-5 or not rep["well_conditioned"] or not rep["within_target"]
    assert rep["sensitivity"] >= dq
    calc = ChargeCalculator(cube=cube_from_system(s), precision="refined", context=gpu_ctx)
    calc.fit_charges()
    calc.close()
    assert "ill-conditioned fit" in capfd.readouterr().err
    # a well-conditioned cell is not flagged
    w = synth.crystal200()
    f = _abi.Fit(gpu_ctx, w.n_grid, w.axis, w.atomic_numbers, w.atom_pos, w.esp_cube, _abi.RP_ORDER_CUBE, 1,
                 _abi.RP_PREC_OZAKI_REFINED)
    f.accumulate()
    f.solve(params.constraint_arrays(w.n_atoms, w.atomic_numbers, 0.0, None, None, None))
    rep = f.solve_report()
    f.close()
    assert rep["well_conditioned"] and rep["within_target"] and 0.0 <= rep["sensitivity"] <= 1e-8


def test_refined_streamed_equals_kept(tmp_path):
    """The refined form without the kept basis (each step recomputes the basis chunk by chunk): forced by
    the option flag, and chosen by the fit itself under a budget that cannot hold the kept basis.  Four
    chunks of 65 536 points (mof1000): the charges equal the kept form's bitwise, within less memory."""
    s = synth.mof1000()
    symm = resp = None
    if s.resp_text:
        resp = tmp_path / "resp.txt"
        resp.write_text(s.resp_text)
    con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, s.q_tot, None,
                                   params.read_resp_params(str(resp)) if resp else None, None)
    out = {}
    for mode in ("kept", "forced", "budget"):
        print(f"refined mof1000, basis {mode}", flush=True)
        budget = 0 if mode != "budget" else out["forced"][2] + (256 << 20)
        ctx = _abi.Context(0, budget)
        f = _abi.Fit(ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                     _abi.RP_PREC_OZAKI_REFINED, flags=_abi.RP_OPT_STREAM_REFINE if mode == "forced" else 0)
        f.accumulate()
        q, _ = f.solve(con)
        rep = f.solve_report()
        out[mode] = (q, rep, ctx.hbm_in_use())
        print(f"  device bytes {out[mode][2] / 1e9:.2f} GB, report {rep}", flush=True)
        f.close()
        ctx.close()
    plan = _abi.chunk_plan(252456, s.n_atoms, 754, _abi.RP_PREC_OZAKI_REFINED)
    kept_bytes = 8 * plan["np"] * plan["rows"] * plan["n_chunks"]
    assert plan["n_chunks"] == 4
    for mode in ("forced", "budget"):
        np.testing.assert_array_equal(out[mode][0], out["kept"][0])
        assert out[mode][1]["refine_steps"] == out["kept"][1]["refine_steps"] >= 1
        assert out[mode][2] <= out["kept"][2] - kept_bytes // 2
    assert out["budget"][2] <= out["forced"][2] + (256 << 20)


def test_refined_deterministic_and_basis_kept(gpu_ctx, tmp_path):
    """Two fits are bitwise equal; the --stats matvec (its own chunk buffer) between two solves leaves
    the kept basis intact, so the second solve repeats the first bitwise; refit does the same."""
    s = synth.crystal200()
    print("refined determinism: crystal200", flush=True)
    calc = _calc(gpu_ctx, s, "refined", tmp_path)
    q1 = calc.fit_charges()
    calc._fit.phi_matvec(q1)
    q2, _ = calc._fit.solve(calc.constraints())
    q3 = calc._fit.refit(calc.constraints())
    calc.close()
    calc = _calc(gpu_ctx, s, "refined", tmp_path)
    q4 = calc.fit_charges()
    calc.close()
    for q in (q2, q3, q4):
        np.testing.assert_array_equal(q, q1)


def test_refined_step_limit(gpu_ctx):
    """refine_steps bounds the steps; a step runs only while the error bound the previous ones left
    (rho / (1 - rho) |dq|, rho = 1e-2 after the first step) exceeds 1e-6 e.  A 400-atom cluster with
    cond(A) ~ 2e12 needs more than the two steps queued with the solve, so the host queues the rest."""
    s = synth.water()
    con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None)
    out = {}
    for steps in (1, 3):
        print(f"refined step limit {steps}", flush=True)
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                     _abi.RP_PREC_OZAKI_REFINED, refine_steps=steps)
        f.accumulate()
        q, _ = f.solve(con)
        out[steps] = (q, f.solve_report())
        f.close()
    assert out[1][1]["refine_steps"] == 1
    r3 = out[3][1]
    assert 1 <= r3["refine_steps"] <= 3
    assert r3["refine_steps"] == 3 or r3["error_bound"] <= 1e-6
    assert np.max(np.abs(out[1][0] - out[3][0])) <= 1e-4
    c = synth.cluster(400)
    con = params.constraint_arrays(c.n_atoms, c.atomic_numbers, 0.0, None, None, None)
    reps = {}
    for steps in (2, 6):
        print(f"refined step limit {steps}: 400-atom cluster", flush=True)
        f = _abi.Fit(gpu_ctx, c.n_grid, c.axis, c.atomic_numbers, c.atom_pos, c.esp_cube, _abi.RP_ORDER_CUBE, 1,
                     _abi.RP_PREC_OZAKI_REFINED, refine_steps=steps)
        f.accumulate()
        f.solve(con)
        reps[steps] = f.solve_report()
        f.close()
        print(f"  {reps[steps]}", flush=True)
    assert reps[2]["refine_steps"] == 2 and not reps[2]["within_target"]
    assert 2 < reps[6]["refine_steps"] <= 6 and reps[6]["within_target"]


def test_refined_large_n_matches_fp64(gpu_ctx):
    """N = 1600: 25 diagonal blocks in the ticketed forward and backward substitutions of each step,
    the separate panel launches of the large-N Cholesky.  The refined charges equal the device's fp64
    path (same fp64 normal equations) to rounding."""
    s = synth.tiny(seed=21, n_atoms=1600, n=76, L=90.0)
    con = params.constraint_arrays(s.n_atoms, s.atomic_numbers, 0.0, None, None, None)
    print("refined large-N: 1600 atoms, 76x77x78 grid", flush=True)
    qs = {}
    for prec in (_abi.RP_PREC_OZAKI_REFINED, _abi.RP_PREC_FP64):
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                     prec)
        f.accumulate()
        qs[prec], _ = f.solve(con)
        assert not f.last_solve_used_lu
        if prec == _abi.RP_PREC_OZAKI_REFINED:
            print(f"  steps, last correction = {f.refine_stats()}", flush=True)
        f.close()
    assert np.max(np.abs(qs[_abi.RP_PREC_OZAKI_REFINED] - qs[_abi.RP_PREC_FP64])) <= 1e-7


def test_refined_fit_type0_matches_oracle(oracle, gpu_ctx):
    """Bare 1/r basis (no reciprocal part): the kept basis and the refinement are the same code."""
    s = synth.tiny(seed=13, n_atoms=4, n=14, L=10.0)
    print("refined fit-type 0", flush=True)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), fit_flag=0)
    calc = ChargeCalculator(cube=cube_from_system(s), fit_flag=0, precision="refined", context=gpu_ctx)
    np.testing.assert_allclose(calc.fit_charges(), r.charges, atol=1e-9)
    calc.close()


def test_refined_singular_kkt_least_squares(oracle, gpu_ctx):
    """Two atoms at one place give identical basis columns, identical slices and an exactly singular
    split Gram: the Cholesky and the LU stop, the minimum-norm least squares solves the 3-slice system,
    and the refinement corrects it against the kept fp64 basis with the same pseudo-inverse, so the
    charges are the fp64 system's minimum-norm solution (the merged atom's charge split equally)."""
    from test_gpu_parity import _duplicate_atom_case
    cube, expect = _duplicate_atom_case(oracle)
    print("refined singular KKT: lstsq fallback, refined", flush=True)
    calc = ChargeCalculator(cube=cube, precision="refined", context=gpu_ctx)
    q = calc.fit_charges()
    rep = calc._fit.solve_report()
    calc.close()
    print(f"  q {q}, expected {expect}, report {rep}", flush=True)
    assert rep["solver"] == _abi.RP_SOLVER_LSTSQ and rep["refine_steps"] >= 1 and rep["within_target"]
    np.testing.assert_allclose(q, expect, rtol=0, atol=1e-6)
    assert abs(q[0] - q[1]) <= 1e-9


def test_refined_stats_from_kept_basis(oracle, gpu_ctx, tmp_path):
    """compute_stats of a refined fit streams the kept basis instead of recomputing it: V_coul and the
    RMS match the oracle on a small cell, and V_coul matches the fp64 recomputation on crystal200."""
    s = synth.tiny(seed=6, n_atoms=4, n=14, L=10.0, triclinic=True)
    print("refined stats: tiny6", flush=True)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), stats=True)
    calc = ChargeCalculator(cube=cube_from_system(s), precision="refined", context=gpu_ctx)
    q = calc.fit_charges()
    np.testing.assert_allclose(q, r.charges, atol=1e-9)
    st = calc.compute_stats(q, need_v_coul=True)
    np.testing.assert_allclose(st["V_coul"], r.v_coul, rtol=1e-10, atol=1e-12)
    assert abs(st["rms_error"] - r.rms) < 1e-10
    calc.close()
    s = synth.crystal200()
    print("refined stats: crystal200 against the fp64 recomputation", flush=True)
    calc = _calc(gpu_ctx, s, "refined", tmp_path)
    q = calc.fit_charges()
    v_kept = calc._fit.phi_matvec(q)
    calc.close()
    c64 = _calc(gpu_ctx, s, "fp64", tmp_path)   # not accumulated: phi_matvec recomputes the basis
    v64 = c64._fit.phi_matvec(q)
    c64.close()
    np.testing.assert_allclose(v_kept, v64, rtol=1e-11, atol=1e-13 * np.abs(v64).max())


def test_refined_gram_close_to_fp64(gpu_ctx):
    """The refinement converges for any Gram close enough to the fp64 one, so the charge bars alone would
    not notice a faulty 3-slice Gram: its normal equations are checked directly against the fp64 path's
    (3 slices hold each element to 2^-22 of its column maximum: relative ~2^-21 per Gram entry)."""
    s = synth.tiny(seed=3, n_atoms=6, n=16, L=10.0, triclinic=True)
    print("refined Gram vs fp64: tiny3", flush=True)
    out = []
    for prec in (_abi.RP_PREC_FP64, _abi.RP_PREC_OZAKI_REFINED):
        f = _abi.Fit(gpu_ctx, s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube, _abi.RP_ORDER_CUBE, 1,
                     prec)
        f.accumulate()
        out.append(f.normal_equations())
        f.close()
    (a64, b64, s64, v64), (a3, b3, s3, v3) = out
    # the 4-slice bars of test_split_normal_equations_close_to_fp64 (rtol 2e-6, atol 1e-9 max) times 2^7
    np.testing.assert_allclose(a3, a64, rtol=2.6e-4, atol=1.3e-7 * np.abs(a64).max())
    np.testing.assert_allclose(b3, b64, rtol=2.6e-4, atol=1.3e-7 * np.abs(b64).max())
    np.testing.assert_allclose(s3, s64, rtol=2.6e-4)
    # and it differs from fp64 at all (the slices, not a silent fp64 copy)
    assert np.max(np.abs(a3 - a64)) > 0.0


@pytest.mark.parametrize("triclinic", [False, True])
def test_multi_image_basis_matches_fp64_on_mixed_cells(gpu_ctx, tmp_path, triclinic):
    """A 30 Bohr cell against Rc = 20 Bohr: about a third of the (point, atom) pairs have only their
    nearest image within the cutoff (the basis kernel's term) and the rest several (k_sep_images adds
    their sum over the cutoff sphere's lattice points).  The refined form's kept basis times a charge
    vector equals the fp64 path's, whose GEMM epilogue sums the images over the shift list's box, to
    fp64 rounding; a pair counted by both kernels or by neither would show at once."""
    s = synth.tiny(seed=31, n_atoms=12, n=48, L=30.0, triclinic=triclinic)
    # both kinds of pairs are present: the nearest-image bound over the kept points and atoms
    c = ChargeCalculator(cube=cube_from_system(s), precision="refined", context=gpu_ctx)
    xyz, _, _ = c._fit.grid()
    box = s.axis * np.asarray(s.n_grid)[:, None]
    inv = np.linalg.inv(box)
    vol = abs(np.linalg.det(box))
    h = np.array([vol / np.linalg.norm(np.cross(box[(i + 1) % 3], box[(i + 2) % 3])) for i in range(3)])
    d = (xyz[::7, None, :] - s.atom_pos[None, :, :]) @ inv
    w = d - np.rint(d)
    near_only = np.min((1.0 - np.abs(w)) * h, axis=2) > 20.0
    print(f"mixed cell (triclinic={triclinic}): {near_only.mean():.2f} of the pairs near-only", flush=True)
    assert 0.05 < near_only.mean() < 0.95
    q = c.fit_charges()
    v_ref = c._fit.phi_matvec(q)
    c.close()
    c64 = ChargeCalculator(cube=cube_from_system(s), precision="fp64", context=gpu_ctx)
    v64 = c64._fit.phi_matvec(q)
    q64 = c64.fit_charges()
    c64.close()
    np.testing.assert_allclose(v_ref, v64, rtol=1e-11, atol=1e-13 * np.abs(v64).max())
    assert np.max(np.abs(q - q64)) < 1e-8
