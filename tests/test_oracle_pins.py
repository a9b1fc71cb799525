"""Pins the CPU oracle against what the mathematics and the reference's own data fix.

The reference ships no tests, fixtures or golden vectors, so each oracle
function is pinned here by a closed form, a textbook constant, a library
routine or a brute-force computation, chosen so that a dropped term, a wrong
sign or index or a transposed operand fails.
"""
import itertools
import math
import os
import re

import numpy as np
import pytest

from pyrepeat_amd import synth, tables

REF = "/root/reference/repeat.py"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ tables

@pytest.mark.skipif(not os.path.exists(REF), reason="reference source not mounted")
def test_vdw_table_matches_reference_text():
    src = open(REF).read()
    m = re.search(r"VDW_RADII = np.array\(\[(.*?)\]", src, re.S)
    vals = [float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
    assert len(vals) == 107
    np.testing.assert_array_equal(tables.VDW_RADII_BOHR, np.array(vals))
    m2 = re.search(r"ATOM_SYMBOLS = \[(.*?)\]", src, re.S)
    syms = [s.strip().strip('"') for s in m2.group(1).replace("\n", " ").split(",") if s.strip()]
    assert tuple(syms) == tables.ATOM_SYMBOLS


def _cpp_vdw_table():
    src = open(os.path.join(ROOT, "pyrepeat_amd", "csrc", "rp_setup.cpp")).read()
    m = re.search(r"const double kVdw\[(\d+)\] = \{(.*?)\};", src, re.S)
    vals = [float(x) for x in m.group(2).replace("\n", " ").split(",") if x.strip()]
    return int(m.group(1)), np.array(vals)


def test_device_vdw_table_matches_host_table():
    """The library's own radius table (the filter's thresholds come from it) holds the same 107 values
    as the host table, which the next test ties to the reference text."""
    n, vals = _cpp_vdw_table()
    assert n == 107 and len(vals) == 107
    np.testing.assert_array_equal(vals, tables.VDW_RADII_BOHR)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference source not mounted")
def test_device_vdw_table_matches_reference_text():
    src = open(REF).read()
    m = re.search(r"VDW_RADII = np.array\(\[(.*?)\]", src, re.S)
    ref = np.array([float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()])
    np.testing.assert_array_equal(_cpp_vdw_table()[1], ref)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference source not mounted")
def test_qeq_default_table_matches_reference_text():
    """DEFAULT_QEQ_PARAMS (repeat.py:611-650), parsed from the reference's text with dict-literal
    semantics (a repeated key keeps its last value), equals the table the QEq paths use."""
    src = open(REF).read()
    m = re.search(r"DEFAULT_QEQ_PARAMS = \{(.*?)\n\}", src, re.S)
    ref = {}
    for sym, chi, hard in re.findall(r'"([A-Z][a-z]?)":\s*\(\s*([-0-9.eE+]+)\s*,\s*([-0-9.eE+]+)\s*\)', m.group(1)):
        ref[sym] = (float(chi), float(hard))
    assert len(ref) == 65
    assert ref == dict(tables.DEFAULT_QEQ_PARAMS)


@pytest.mark.parametrize("trees,n,expected", [
    # one tree: linked atom 2 is dropped from the independent list, then copies its base
    ([{"base": 0, "linked": [2]}], 5, [10, 11, 10, 12, 13]),
    # two trees: independents 0, 1, 2, 5 take g0..g3; linked atoms copy their bases
    ([{"base": 1, "linked": [4, 6]}, {"base": 0, "linked": [3]}], 7, [10, 11, 12, 10, 11, 13, 11]),
    # a linked atom that is the base of a later tree: the copies run in tree order
    ([{"base": 0, "linked": [2]}, {"base": 2, "linked": [4]}], 5, [10, 11, 10, 12, 10]),
    # the same trees in the other order: atom 4 copies atom 2 before atom 2 is set, so stays 0
    ([{"base": 2, "linked": [4]}, {"base": 0, "linked": [2]}], 5, [10, 11, 10, 12, 0]),
])
def test_expand_sum_phi_known_answers(oracle, trees, n, expected):
    """_expand_sum_phi (repeat.py:2491-2509) as fit_charges calls it (repeat.py:2317-2320): with the
    unreduced sum_phi g = (10, 11, 12, ...), scattered as if it were reduced.  Expected vectors are
    worked by hand from the reference's loops; the oracle and the product's stats path must both give
    them."""
    from pyrepeat_amd.calculator import _expand_sum_phi_quirk
    g = np.arange(10.0, 10.0 + n)
    np.testing.assert_array_equal(oracle.expand_sum_phi(g, trees, n), np.array(expected, dtype=float))
    np.testing.assert_array_equal(_expand_sum_phi_quirk(g, {"trees": trees}, n), np.array(expected, dtype=float))


def test_vdw_lookup_index_rules():
    assert tables.vdw_radius_lookup([1])[0] == 2.72687          # H -> index 0
    assert tables.vdw_radius_lookup([8])[0] == 3.30702          # O
    assert tables.vdw_radius_lookup([0])[0] == tables.VDW_RADII_BOHR[-1]  # Z=0 -> last entry
    with pytest.raises(IndexError):
        tables.vdw_radius_lookup([108])


# ------------------------------------------------------------------ Ewald setup

def _cells():
    yield np.eye(3) * 12.0
    yield synth.triclinic_box(24.0, 22.0, 26.0, 85.0, 95.0, 100.0)
    yield np.eye(3) * 40.0


@pytest.mark.parametrize("box", list(_cells()))
def test_kset_count_order_and_weights(oracle, box):
    vol = float(np.dot(box[0], np.cross(box[1], box[2])))
    ew = oracle.ewald_setup(box, vol, 20.0, 1)
    assert len(ew.kvecs) == 754
    assert int(np.sum(ew.kidx[:, 0] == 0)) == 144
    # loop order kx, ky, kz ascending
    assert [tuple(t) for t in ew.kidx] == sorted(tuple(t) for t in ew.kidx)
    # reciprocal vectors: b_i . a_j = 2 pi delta_ij
    np.testing.assert_allclose(ew.recip @ box.T, 2 * np.pi * np.eye(3), atol=1e-12)
    # closed form of the weights from the integer indices
    k = ew.kidx @ ew.recip
    k2 = np.sum(k * k, axis=1)
    alpha = (np.pi / 20.0) ** 2
    np.testing.assert_allclose(ew.kcoefs, 4 * np.pi / vol * np.exp(-k2 / (4 * alpha)) / k2, rtol=1e-13)
    np.testing.assert_allclose(ew.sqrt_alpha, np.pi / 20.0, rtol=1e-15)


def test_shift_lattice(oracle):
    for box, n in ((np.eye(3) * 12.0, 125), (np.eye(3) * 40.0, 27), (np.eye(3) * 70.0, 27)):
        ew = oracle.ewald_setup(box, float(np.linalg.det(box)), 20.0, 1)
        assert len(ew.shifts) == n
        idx = [(ix, iy, iz) for ix in range(-ew.nmax[0], ew.nmax[0] + 1)
               for iy in range(-ew.nmax[1], ew.nmax[1] + 1) for iz in range(-ew.nmax[2], ew.nmax[2] + 1)]
        np.testing.assert_array_equal(ew.shifts, np.array(idx, dtype=float) @ box)


def _full_sphere(box, alpha, kmax=7, ksqmax=49):
    """All k with 0 < |h|^2 < ksqmax, both half spaces: the textbook reciprocal sum."""
    vol = abs(np.linalg.det(box))
    recip = 2 * np.pi * np.linalg.inv(box).T
    hs = np.array([h for h in itertools.product(range(-kmax, kmax + 1), repeat=3)
                   if 0 < h[0] ** 2 + h[1] ** 2 + h[2] ** 2 < ksqmax], dtype=float)
    k = hs @ recip
    k2 = np.sum(k * k, axis=1)
    return k, 4 * np.pi / vol * np.exp(-k2 / (4 * alpha)) / k2, hs


def test_half_space_quirk(oracle):
    """Reference reciprocal part = 1/2 full sphere + 1/2 of the kx = 0 plane (no factor 2 on kx > 0)."""
    box = synth.triclinic_box(11.0, 12.0, 10.0, 88.0, 93.0, 97.0)
    vol = float(np.dot(box[0], np.cross(box[1], box[2])))
    ew = oracle.ewald_setup(box, vol, 20.0, 1)
    rng = np.random.default_rng(0)
    g = rng.random((40, 3)) @ box
    a = rng.random((2, 3)) @ box
    ref_only = oracle.Ewald(box=box, volume=vol, r_cutoff=20.0, fit_flag=1, sqrt_alpha=ew.sqrt_alpha,
                            kvecs=ew.kvecs, kcoefs=ew.kcoefs, shifts=np.zeros((0, 3)))
    rec_ref = oracle.phi_matrix(g, a, ref_only, shifts=np.zeros((0, 3)))
    k, w, hs = _full_sphere(box, ew.alpha)
    d = g[:, None, :] - a[None, :, :]
    full = np.einsum("mak,k->ma", np.cos(d @ k.T), w)
    plane = np.einsum("mak,k->ma", np.cos(d @ k[hs[:, 0] == 0].T), w[hs[:, 0] == 0])
    np.testing.assert_allclose(rec_ref, 0.5 * full + 0.5 * plane, rtol=1e-11, atol=1e-15)


def test_madelung_nacl(oracle):
    """Ewald potential at a Na site of rock salt = -M / r0 with the Madelung constant 1.747564594633.

    Uses the oracle's basis kernel with a converged splitting and the full k-sphere,
    plus the standard self term -2 sqrt(alpha)/sqrt(pi).
    """
    a = 10.0
    box = np.eye(3) * a
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    pos = np.vstack([fcc, fcc + [a / 2, 0, 0]])
    q = np.array([1.0] * 4 + [-1.0] * 4)
    sqrt_alpha = 0.35
    alpha = sqrt_alpha ** 2
    k, w, _ = _full_sphere(box, alpha, kmax=9, ksqmax=82)
    nmax = 3
    shifts = np.array([(i, j, l) for i in range(-nmax, nmax + 1) for j in range(-nmax, nmax + 1)
                       for l in range(-nmax, nmax + 1)], dtype=float) * a
    ew = oracle.Ewald(box=box, volume=a ** 3, r_cutoff=20.0, fit_flag=1, sqrt_alpha=sqrt_alpha, kvecs=k, kcoefs=w,
                      shifts=shifts)
    phi = oracle.phi_matrix(pos[:1], pos, ew, shifts=shifts)[0]
    pot = float(phi @ q) - 2 * sqrt_alpha / math.sqrt(math.pi) * q[0]
    assert pot == pytest.approx(-1.747564594633 / (a / 2), rel=1e-9)


def test_real_space_alpha_independence(oracle):
    """Total Ewald potential of a neutral set does not depend on the splitting parameter."""
    box = synth.triclinic_box(9.0, 9.5, 8.5, 80.0, 97.0, 104.0)
    vol = abs(np.linalg.det(box))
    rng = np.random.default_rng(3)
    atoms = rng.random((4, 3)) @ box
    q = np.array([0.7, -0.2, -0.9, 0.4])
    g = rng.random((25, 3)) @ box
    shifts = np.array(list(itertools.product(range(-5, 6), repeat=3)), dtype=float) @ box
    vals = []
    for sa in (0.30, 0.38):
        k, w, _ = _full_sphere(box, sa * sa, kmax=11, ksqmax=122)
        ew = oracle.Ewald(box=box, volume=vol, r_cutoff=30.0, fit_flag=1, sqrt_alpha=sa, kvecs=k, kcoefs=w,
                          shifts=shifts)
        vals.append(oracle.phi_matrix(g, atoms, ew, shifts=shifts) @ q)
    np.testing.assert_allclose(vals[0], vals[1], rtol=0, atol=1e-9)


def test_fit_type0_is_inverse_distance(oracle):
    box = np.eye(3) * 10.0
    ew = oracle.ewald_setup(box, 1000.0, 20.0, 0)
    g = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 0.0]])
    a = np.array([[0.0, 0.0, 0.0]])
    np.testing.assert_allclose(oracle.phi_matrix(g, a, ew)[:, 0], [1.0 / 3.0, 0.0])


def test_pruned_shifts_keep_every_contribution(oracle):
    s = synth.water()
    box, vol = oracle.box_and_volume(s.n_grid, s.axis)
    ew = oracle.ewald_setup(box, vol, 20.0, 1)
    rng = np.random.default_rng(1)
    g = rng.random((200, 3)) @ box
    pr = oracle.prune_shifts(ew.shifts, g, box, s.atom_pos, 20.0)
    assert 0 < len(pr) < len(ew.shifts)
    # pruned list is an ordered subsequence
    it = iter(map(tuple, ew.shifts))
    assert all(any(t == u for u in it) for t in map(tuple, pr))
    real_only = oracle.Ewald(box=box, volume=vol, r_cutoff=20.0, fit_flag=1, sqrt_alpha=ew.sqrt_alpha,
                             kvecs=np.zeros((0, 3)), kcoefs=np.zeros(0), shifts=ew.shifts)
    np.testing.assert_array_equal(oracle.phi_matrix(g, s.atom_pos, real_only, shifts=ew.shifts),
                                  oracle.phi_matrix(g, s.atom_pos, real_only, shifts=pr))


# ------------------------------------------------------------------ filter

def test_distance_is_cdist_bitwise(oracle):
    import ctypes
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(5)
    x = rng.normal(size=(300, 3)) * 7
    y = rng.normal(size=(200, 3)) * 7
    out = np.empty((300, 200))
    dp = ctypes.POINTER(ctypes.c_double)
    oracle.lib().oracle_cdist(300, x.ctypes.data_as(dp), 200, y.ctypes.data_as(dp), out.ctypes.data_as(dp))
    np.testing.assert_array_equal(out, cdist(x, y))


@pytest.mark.parametrize("seed,tric,vmax", [(0, False, 1000.0), (1, True, 1000.0), (2, True, 1.6), (3, False, 1.3)])
def test_filter_matches_cdist_predicate(oracle, seed, tric, vmax):
    s = synth.tiny(seed=seed, n_atoms=4, n=10, L=8.0, triclinic=tric)
    box, _ = oracle.box_and_volume(s.n_grid, s.axis)
    eff, rmin, rmax, _ = oracle.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, 1.0, vmax)
    a = oracle.filter_mask(s.n_grid, s.axis, eff, rmin, rmax)
    b = oracle.filter_mask_cdist(s.n_grid, s.axis, eff, rmin, rmax)
    np.testing.assert_array_equal(a, b)
    assert 0 < a.sum() < a.size


def test_filter_clipping_box_is_first_voxel(oracle):
    """The reference clips images with the AABB of the step vectors' corners, not of the cell."""
    s = synth.tiny(seed=4, n_atoms=3, n=10, L=8.0)
    box, _ = oracle.box_and_volume(s.n_grid, s.axis)
    eff, rmin, rmax, gmax = oracle.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, 1.0, 1.2)
    lo = np.min(s.axis, axis=0).clip(max=0) - (gmax + 0.1)
    hi = np.max(np.cumsum(s.axis, axis=0), axis=0) + (gmax + 0.1)
    assert np.all(eff >= lo - 1e-12) and np.all(eff <= hi + 1e-12)
    assert len(eff) < 27 * 3


def test_filter_tie_is_excluded(oracle):
    """d == r_min exactly is not kept (strict >); d == r_max is kept (<=)."""
    axis = np.eye(3)
    atom = np.array([[0.0, 0.0, 0.0]])
    pts = oracle.filter_mask([8, 8, 1], axis, atom, np.array([5.0]), np.array([5.0 + 1e-9]))
    keep = pts.reshape(1, 8, 8)[0]
    assert keep[4, 3] == 0 and keep[3, 4] == 0            # (3,4,0): d = 5 exactly -> invalid
    pts2 = oracle.filter_mask([8, 8, 1], axis, atom, np.array([4.9]), np.array([5.0]))
    assert pts2.reshape(8, 8)[4, 3] == 1                   # d = r_max exactly -> near


# ------------------------------------------------------------------ normal equations and solve

@pytest.fixture(scope="module")
def small_fit(oracle):
    s = synth.tiny(seed=7, n_atoms=5, n=12, L=9.0, triclinic=True)
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_flat(), stats=True)
    return s, r


def test_centred_normal_equations_identity(oracle, small_fit):
    s, r = small_fit
    phi = oracle.phi_matrix(r.grid_pos, s.atom_pos, r.ewald)
    pc = phi - phi.mean(axis=0)
    np.testing.assert_allclose(r.A, pc.T @ pc, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(r.b, pc.T @ (r.V - r.V.mean()), rtol=1e-9, atol=1e-12)


def test_kkt_solution_properties(oracle, small_fit):
    s, r = small_fit
    assert abs(r.charges.sum()) < 1e-12
    # bordered system solved: compare with numpy's LAPACK solve of the same matrix
    n = len(r.b)
    K = np.zeros((n + 1, n + 1))
    K[:n, :n] = r.A
    K[:n, n] = 1
    K[n, :n] = 1
    rhs = np.append(r.b, 0.0)
    np.testing.assert_allclose(r.kkt_solution, np.linalg.solve(K, rhs), rtol=1e-10, atol=1e-12)


def test_kkt_lstsq_branch_on_exact_zero_pivot(oracle):
    """solve_kkt takes the reference's lstsq branch (repeat.py:2306-2310) exactly when LAPACK's LU meets a
    zero pivot.  Small integer entries keep the elimination exact: rows 0 and 1 of the bordered matrix
    are equal, the unit pivot eliminates row 1 to exact zeros, the LU raises, and lstsq returns the
    minimum-norm solution, worked by hand: q0 + q1 + lam = 1, q2 + lam = 2, q0 + q1 + q2 = 0 give
    lam = 1.5, q2 = 0.5, q0 + q1 = -0.5, split equally by the minimum norm."""
    A = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    sol, fallback = oracle.solve_kkt(A, np.array([1.0, 1.0, 2.0]), np.ones(3), 0.0)
    assert fallback
    np.testing.assert_allclose(sol, [-0.25, -0.25, 0.5, 1.5], rtol=0, atol=1e-14)
    # a regular system does not take it
    sol2, fallback2 = oracle.solve_kkt(A + np.eye(3), np.array([1.0, 1.0, 2.0]), np.ones(3), 0.0)
    assert not fallback2


def test_self_consistent_recovery_and_rms(oracle):
    """V := Phi q* + c on the kept points -> the fit returns q* and RMS 0."""
    s = synth.tiny(seed=8, n_atoms=4, n=12, L=9.0)
    box, vol = oracle.box_and_volume(s.n_grid, s.axis)
    ew = oracle.ewald_setup(box, vol, 20.0, 1)
    eff, rmin, rmax, _ = oracle.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, 1.0, 1000.0)
    keep = oracle.filter_mask(s.n_grid, s.axis, eff, rmin, rmax)
    nx, ny, nz = s.n_grid
    idx = np.arange(nx * ny * nz)
    k = idx // (nx * ny)
    j = (idx % (nx * ny)) // nx
    i = idx % nx
    allpos = (np.outer(i, s.axis[0]) + np.outer(j, s.axis[1])) + np.outer(k, s.axis[2])
    qstar = np.array([0.3, -0.5, 0.45, -0.25])
    v = np.zeros(nx * ny * nz)
    v[keep == 1] = oracle.phi_matrix(allpos[keep == 1], s.atom_pos, ew) @ qstar + 0.123
    r = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, v, stats=True)
    np.testing.assert_allclose(r.charges, qstar, atol=1e-9)
    assert r.rms < 1e-10
    # invariance to a constant added to V
    r2 = oracle.fit(s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, v + 5.0)
    np.testing.assert_allclose(r2.charges, r.charges, atol=1e-10)


def test_symmetry_reduction_is_projection(oracle):
    rng = np.random.default_rng(2)
    n = 7
    X = rng.normal(size=(30, n))
    A = X.T @ X
    b = rng.normal(size=n)
    trees = [{"base": 1, "linked": [3, 5]}, {"base": 0, "linked": [6]}]
    Ar, br, wr, _ = oracle.apply_symmetry(A, b, np.ones(n), np.arange(n), trees, n)
    ind = [0, 1, 2, 4]
    P = np.zeros((n, len(ind)))
    groups = {0: [0, 6], 1: [1, 3, 5], 2: [2], 4: [4]}
    for c, base in enumerate(ind):
        P[groups[base], c] = 1.0
    np.testing.assert_allclose(Ar, P.T @ A @ P, rtol=1e-12)
    np.testing.assert_allclose(br, P.T @ b, rtol=1e-12)
    np.testing.assert_array_equal(wr, P.sum(axis=0))
    q = oracle.expand_charges(np.array([1.0, 2.0, 3.0, 4.0]), trees, n)
    np.testing.assert_array_equal(q, [1.0, 2.0, 3.0, 2.0, 4.0, 2.0, 1.0])


def test_resp_file_quirk(oracle, tmp_path):
    good = tmp_path / "resp_ok.txt"
    good.write_text("H 0.1 0.0\nO 0.2 -0.5\n")
    assert oracle.read_resp_params(str(good)) == {"H": {"strength": 0.1, "target": 0.0},
                                                  "O": {"strength": 0.2, "target": -0.5}}
    bad = tmp_path / "resp_bad.txt"
    bad.write_text("# Element Strength Target\nH 0.1 0.0\n")
    assert oracle.read_resp_params(str(bad)) == {}
    A = np.eye(2)
    b = np.zeros(2)
    A2, b2 = oracle.apply_resp(A, b, [1, 8], {"H": {"strength": 0.1, "target": 0.2}})
    np.testing.assert_allclose(np.diag(A2), [1.1, 1.0])
    np.testing.assert_allclose(b2, [0.02, 0.0])


def test_parse_atom_ranges(oracle):
    assert oracle.parse_atom_ranges("1-3, 5,3,7-8") == [0, 1, 2, 4, 6, 7]


def test_compaction_order_and_centring(oracle):
    s = synth.tiny(seed=9, n_atoms=3, n=9, L=8.0, triclinic=True)
    box, _ = oracle.box_and_volume(s.n_grid, s.axis)
    eff, rmin, rmax, _ = oracle.effective_atoms(s.atom_pos, s.atomic_numbers, s.axis, box, 1.0, 1000.0)
    keep = oracle.filter_mask(s.n_grid, s.axis, eff, rmin, rmax)
    pos, v, flat = oracle.compact(keep, s.esp_flat(), s.n_grid, s.axis)
    assert np.all(np.diff(flat) > 0)
    assert abs(v.mean()) < 1e-15
    nx, ny, _ = s.n_grid
    f = flat[len(flat) // 2]
    i, j, k = f % nx, (f // nx) % ny, f // (nx * ny)
    np.testing.assert_array_equal(pos[len(flat) // 2], (i * s.axis[0] + j * s.axis[1]) + k * s.axis[2])


# ------------------------------------------------------------------ the reference's CPU filter batches

def _slab_case(oracle, axis, n, n_atoms, vdw_max, seed):
    rng = np.random.default_rng(seed)
    axis = np.asarray(axis, dtype=np.float64)
    box = axis * n
    pos = synth.place_atoms(rng, box, n_atoms, 2.0)
    Z = rng.choice([1, 6, 7, 8], size=n_atoms)
    eff, rmin, rmax, gmax = oracle.effective_atoms(pos, Z, axis, box, 1.0, vdw_max)
    exact = oracle.filter_mask([n, n, n], axis, eff, rmin, rmax)
    slab = {w: oracle.filter_mask_cpu_slab([n, n, n], axis, eff, rmin, rmax, gmax, w) for w in (1, 2)}
    return exact, slab


@pytest.mark.parametrize("triclinic,vdw_max", [(False, 1000.0), (True, 1.6), (True, 3.0)])
def test_reference_cpu_slab_is_exact_on_standard_cells(oracle, triclinic, vdw_max):
    """With the first two grid vectors in the xy plane (every configuration here: cubes written with a along
    x and b in xy) the reference's CPU z-slab prefilter changes nothing: its mask equals the exact predicate
    for one worker and for two (48^3 > 100 000 points, so two z-chunks), and the host's check agrees."""
    from pyrepeat_amd.params import reference_cpu_filter_is_exact
    box = synth.triclinic_box(18.0, 19.0, 17.0, 80.0, 97.0, 104.0) if triclinic else np.eye(3) * 18.0
    axis = box / 48
    assert reference_cpu_filter_is_exact(axis, [48, 48, 48], 1.0, vdw_max)
    exact, slab = _slab_case(oracle, axis, 48, 40, vdw_max, 3)
    assert exact.sum() > 0
    for w, m in slab.items():
        np.testing.assert_array_equal(m, exact)


def test_reference_cpu_slab_differs_on_tilted_cells(oracle):
    """A grid whose first vector points down in z (x_p = i axis[0][2] reaches -21 Bohr) at vdw-max 1.6: with
    one z-chunk the slab keeps every effective atom and the mask is exact; with two chunks (a host with two
    or more cores) the second chunk's slab drops atoms near its tilted points, and the reference's CPU
    mask differs from the exact predicate on hundreds of points (kept inside a dropped atom's r_min, or
    dropped for want of a near atom).  The host's check flags the cell; the device applies the exact
    predicate."""
    from pyrepeat_amd.params import reference_cpu_filter_is_exact
    axis = [[0.5, 0.0, -0.45], [0.05, 0.5, 0.0], [0.0, 0.0, 0.5]]
    assert not reference_cpu_filter_is_exact(axis, [48, 48, 48], 1.0, 1.6)
    exact, slab = _slab_case(oracle, axis, 48, 150, 1.6, 5)
    np.testing.assert_array_equal(slab[1], exact)
    assert (slab[2] != exact).sum() > 100

