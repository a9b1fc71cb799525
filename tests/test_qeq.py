"""QEq charge model (ChargeCalculator.compute_qeq_charges, repeat.py:2538-2634).

CPU: the oracle's restatement pinned by a two-atom closed form (fit-type 0), by its reciprocal and
real parts against the oracle's basis kernel (which the Madelung and alpha-independence pins hold),
and by the defining property of the solution (equal chemical potentials chi_i + (J q)_i).
GPU: the device Coulomb matrix and charges against the oracle, the invariants, the CLI output.
The reference's fit-type-1 pair term calls scipy.spatial.erfc, which does not exist, so its QEq
path raises there; both the oracle and the device use scipy.special.erfc's function.
"""
import dataclasses

import numpy as np
import pytest
from scipy.special import erfc

from pyrepeat_amd import synth
from pyrepeat_amd.cube import cube_from_system


def _ewald(O, s, fit_flag=1, rc=20.0):
    box, vol = O.box_and_volume(s.n_grid, s.axis)
    return O.ewald_setup(box, vol, rc, fit_flag)


def test_oracle_two_atoms_fit_type0_closed_form(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [1.1, 2.3, -0.7]])
    d = float(np.linalg.norm(pos[1] - pos[0]))
    chi = np.array([0.32, 0.17])
    jii = np.array([0.63, 0.51])
    for q_tot in (0.0, 1.0, -0.5):
        J = oracle.qeq_coulomb_matrix(pos, None, jii, 0)
        q = oracle.qeq_solve(J, chi, q_tot)
        # q0 + q1 = Q and (J00 - 1/d) q0 + (1/d - J11) q1 = chi1 - chi0
        q0 = (chi[1] - chi[0] + (jii[1] - 1 / d) * q_tot) / (jii[0] + jii[1] - 2 / d)
        np.testing.assert_allclose(q, [q0, q_tot - q0], rtol=0, atol=1e-14)


def test_oracle_pair_term_is_the_basis_kernel(oracle):
    """J_ij (i != j) equals the basis function of atom j at atom i with the cutoff lifted: the reciprocal
    part alone (no shifts) and the full sum (every shift, r^2 > 1e-20 vs r > 1e-10: same for distinct atoms)."""
    s = synth.tiny(seed=21, n_atoms=4, n=8, L=8.5, triclinic=True)
    ew = _ewald(oracle, s)
    ew_all = dataclasses.replace(ew, r_cutoff=1e9)
    R = s.atom_pos
    for i, j in ((0, 1), (2, 3), (3, 0)):
        pair = oracle.ewald_pairwise(R[i], R[j], ew)
        full = oracle.phi_matrix(R[i:i + 1], R[j:j + 1], ew_all)[0, 0]
        recip = oracle.phi_matrix(R[i:i + 1], R[j:j + 1], ew, shifts=np.zeros((0, 3)))[0, 0]
        assert abs(pair - full) < 1e-13 * max(1.0, abs(full))
        real = sum(erfc(ew.sqrt_alpha * r) / r for r in np.linalg.norm(R[i] - R[j] + ew.shifts, axis=1))
        assert abs((pair - real) - recip) < 1e-13


def test_oracle_qeq_equalizes_chemical_potential(oracle):
    s = synth.tiny(seed=22, n_atoms=6, n=8, L=9.0, triclinic=True)
    ew = _ewald(oracle, s)
    chi, jii, method = oracle.qeq_parameters(s.atomic_numbers)
    assert method == "QEq"
    J = oracle.qeq_coulomb_matrix(s.atom_pos, ew, jii, 1)
    assert np.array_equal(J, J.T)
    q = oracle.qeq_solve(J, chi, 1.0)
    mu = chi + J @ q
    assert abs(q.sum() - 1.0) < 1e-12
    assert np.ptp(mu) < 1e-10 * max(1.0, np.max(np.abs(mu)))


def test_qeq_parameter_resolution(oracle):
    from pyrepeat_amd.tables import DEFAULT_QEQ_PARAMS
    Z = np.array([1, 8, 6, 118, 0])          # H, O, C, Og (no parameters), X
    chi, jii, method = oracle.qeq_parameters(Z)
    assert method == "QEq"
    assert chi[1] == DEFAULT_QEQ_PARAMS["O"][0] and jii[1] == 2 * DEFAULT_QEQ_PARAMS["O"][1]
    assert chi[3] == 0.15 and jii[3] == 0.50 and chi[4] == 0.15
    chi2, jii2, method2 = oracle.qeq_parameters(Z, qeq_data={"O": {"electroneg": 0.3, "hardness": 0.2}},
                                                param_set=DEFAULT_QEQ_PARAMS)
    assert method2 == "EQeq" and chi2[1] == 0.3 and jii2[1] == 0.4
    ev = 1 / 27.211386245988
    assert chi2[0] == 0.5 * (13.598 - 2.0) * ev and jii2[0] == (13.598 + 2.0) * ev


# ------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("fit_flag", [1, 0])
def test_qeq_matches_oracle(oracle, gpu_ctx, fit_flag):
    from pyrepeat_amd.calculator import ChargeCalculator
    s = synth.tiny(seed=23, n_atoms=10, n=12, L=15.0, triclinic=True)
    print(f"qeq fit_flag={fit_flag} n_atoms=10", flush=True)
    calc = ChargeCalculator(cube=cube_from_system(s), fit_flag=fit_flag, q_tot=-1.0, context=gpu_ctx)
    q, J = calc.compute_qeq_charges(return_matrix=True)
    chi, jii, _ = oracle.qeq_parameters(s.atomic_numbers)
    Jo = oracle.qeq_coulomb_matrix(s.atom_pos, _ewald(oracle, s, fit_flag), jii, fit_flag)
    np.testing.assert_allclose(J, Jo, rtol=0, atol=1e-12)
    qo = oracle.qeq_solve(Jo, chi, -1.0)
    np.testing.assert_allclose(q, qo, rtol=0, atol=1e-10)
    assert abs(q.sum() + 1.0) < 1e-11
    calc.close()


@pytest.mark.gpu
def test_qeq_crystal200_invariants_and_oracle(oracle, gpu_ctx):
    from pyrepeat_amd.calculator import ChargeCalculator
    s = synth.crystal200()
    print("qeq crystal200", flush=True)
    calc = ChargeCalculator(cube=cube_from_system(s), q_tot=0.0, context=gpu_ctx)
    q, J = calc.compute_qeq_charges(return_matrix=True)
    chi, jii, _ = oracle.qeq_parameters(s.atomic_numbers)
    mu = chi + J @ q
    assert abs(q.sum()) < 1e-11
    assert np.ptp(mu) < 1e-9 * np.max(np.abs(mu))
    assert np.max(np.abs(J - J.T)) < 1e-13
    # oracle on the upper triangle of a 40-atom subset of rows (the full oracle matrix is slow in Python)
    ew = _ewald(oracle, s)
    R = s.atom_pos
    for i in range(0, 200, 5):
        for j in range(i + 1, 200, 7):
            assert abs(J[i, j] - oracle.ewald_pairwise(R[i], R[j], ew)) < 1e-12
    # a common translation of all atoms changes no pair difference
    calc2 = ChargeCalculator(cube=dict(cube_from_system(s), atom_pos=R + np.array([0.37, -1.2, 2.05])),
                             q_tot=0.0, context=gpu_ctx)
    np.testing.assert_allclose(calc2.compute_qeq_charges(), q, rtol=0, atol=1e-9)
    calc.close()
    calc2.close()


@pytest.mark.gpu
def test_cli_qeq_and_repeat_output(tmp_path, capsys):
    from pyrepeat_amd import repeat as cli
    s = synth.tiny(seed=24, n_atoms=5, n=12, L=9.0)
    p = tmp_path / "t.cube"
    s.write_cube(str(p))
    out = tmp_path / "q.txt"
    print("cli qeq", flush=True)
    assert cli.main([str(p), "--charge", "qeq", "--output", str(out)]) == 0
    text = capsys.readouterr().out
    assert "QEq Charges:" in text and "Total charge:" in text
    rows = out.read_text().splitlines()
    assert len(rows) == 5
    f = rows[0].split()
    assert f[0] == "1" and len(f) == 6
    print("cli repeat --stats --output", flush=True)
    out2 = tmp_path / "v.txt"
    assert cli.main([str(p), "--fp64", "--stats", "--output", str(out2)]) == 0
    text = capsys.readouterr().out
    assert "REPEAT Charges:" in text and "RMS error:" in text
    line = out2.read_text().splitlines()[0].split()
    assert len(line) == 5
