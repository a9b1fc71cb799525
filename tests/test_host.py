"""CPU tests of the host side: C ABI exports, parameter files, constraint maps, cube I/O."""
import ctypes
import os
import re

import numpy as np
import pytest

from pyrepeat_amd import _abi, params, synth
from pyrepeat_amd.cube import read_gauss_cube, v_pot_view

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    header = open(os.path.join(ROOT, "include", "repeat_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|double|const char \*)\s*\**(rp_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_abi.EXPORTED)
    for name in declared:
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)


def test_product_has_no_cpu_fallback(tmp_path):
    with pytest.raises(_abi.RepeatHipError):
        _abi._lib_saved = _abi._lib
        try:
            _abi._lib = None
            _abi.load(str(tmp_path / "missing.so"))
        finally:
            _abi._lib = _abi._lib_saved


def _random_trees(rng, n, k):
    trees = []
    for _ in range(k):
        size = int(rng.integers(2, 5))
        ids = [int(x) for x in rng.choice(n, size=size, replace=False)]
        trees.append({"base": ids[0], "linked": ids[1:]})
    return trees


@pytest.mark.parametrize("seed", range(6))
def test_merge_sets_reproduce_sequential_symmetry(oracle, seed):
    """Merge sets + weights == the reference's in-place column/row merges, overlapping groups included."""
    rng = np.random.default_rng(seed)
    n = 12
    X = rng.normal(size=(40, n))
    A = X.T @ X
    b = rng.normal(size=n)
    trees = _random_trees(rng, n, 4)
    Ar, br, wr, _ = oracle.apply_symmetry(A, b, np.ones(n), np.arange(n), trees, n)
    con = params.constraint_arrays(n, np.ones(n, dtype=int), 0.0, {"trees": trees})
    nr = con["n_reduced"]
    M = np.zeros((nr, n))
    for I in range(nr):
        for p in range(con["red_ptr"][I], con["red_ptr"][I + 1]):
            M[I, con["red_src"][p]] += con["red_mult"][p]
    np.testing.assert_allclose(M @ A @ M.T, Ar, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(M @ b, br, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(con["weights"], wr)
    x = rng.normal(size=nr)
    q_ref = oracle.expand_charges(x, trees, n)
    q = np.where(con["expand_src"] >= 0, x[np.maximum(con["expand_src"], 0)], 0.0)
    np.testing.assert_array_equal(q, q_ref)


def test_restraint_terms_match_oracle(oracle):
    types = np.array([1, 8, 6, 1, 7, 30])
    A = np.eye(6) * 3.0
    b = np.zeros(6)
    resp = {"H": {"strength": 0.1, "target": 0.05}, "N": {"strength": 0.3, "target": -0.2}}
    con = params.constraint_arrays(6, types, 0.0, None, resp, None)
    A2, b2 = oracle.apply_resp(A, b, types, resp)
    np.testing.assert_allclose(np.diag(A2) - 3.0, con["diag_add"])
    np.testing.assert_allclose(b2, con["rhs_add"])
    qeq = {"O": {"electroneg": 0.5, "hardness": 0.6}}
    con = params.constraint_arrays(6, types, 0.0, None, None, qeq)
    A3, b3 = oracle.apply_qeq(A, b, types, qeq)
    np.testing.assert_allclose(np.diag(A3) - 3.0, con["diag_add"])
    np.testing.assert_allclose(b3, con["rhs_add"])


def test_parameter_readers_follow_reference(tmp_path, oracle):
    p = tmp_path / "symm.txt"
    p.write_text("# groups\n1-3,5  # first\n\n7,9\n")
    assert params.read_symmetry(str(p)) == oracle.read_symmetry(str(p))
    r = tmp_path / "resp.txt"
    r.write_text("# Element Strength Target\nH 0.1 0.0\n")
    assert params.read_resp_params(str(r)) == {} == oracle.read_resp_params(str(r))


def test_cube_roundtrip_and_reference_layout(tmp_path):
    s = synth.tiny(seed=11, n_atoms=3, n=6, L=7.0, triclinic=True)
    path = tmp_path / "t.cube"
    s.write_cube(str(path))
    c = read_gauss_cube(str(path))
    np.testing.assert_array_equal(c["n_grid"], s.n_grid)
    np.testing.assert_array_equal(c["axis_vector"], s.axis)
    np.testing.assert_array_equal(c["atom_pos"], s.atom_pos)
    np.testing.assert_allclose(c["esp_cube"], s.esp_cube, rtol=1e-10)
    # the reference's V_pot[i + j*nx, k] filled x outer, z inner (repeat.py:797-803)
    nx, ny, nz = s.n_grid
    vp = np.zeros((nx * ny, nz))
    idx = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                vp[i + j * nx, k] = c["esp_cube"][idx]
                idx += 1
    np.testing.assert_array_equal(v_pot_view(c), vp)
    # F-order ravel of V_pot is the flat order i + j*nx + k*nx*ny (repeat.py:2086)
    np.testing.assert_allclose(vp.ravel(order="F"), s.esp_flat(), rtol=1e-10)


def test_request_bound_on_host_and_invalid_requests():
    """rp_request_bytes is host-only: it grows with the grid and the atom count and rejects bad requests."""
    from pyrepeat_amd import params, synth
    small, large = synth.tiny(seed=1, n_atoms=4, n=12), synth.tiny(seed=1, n_atoms=4, n=24)
    opts = _abi.make_options(1, _abi.RP_PREC_OZAKI)
    con = params.constraint_arrays(4, small.atomic_numbers, 0.0, None, None, None)
    args = lambda s: (s.n_grid, s.axis, s.atomic_numbers, s.atom_pos, s.esp_cube)
    b_small = _abi.Server.request_bytes(args(small), con, opts)
    b_large = _abi.Server.request_bytes(args(large), con, opts)
    assert 0 < b_small < b_large
    # at least one 512-row chunk of Phi (fp64) and its 4 int8 slices over 128 padded columns
    assert b_small >= 12 * 128 * 512
    b64 = _abi.Server.request_bytes(args(small), con, _abi.make_options(1, _abi.RP_PREC_FP64))
    assert b64 > 0
    bad = dict(con, n_reduced=0)
    with pytest.raises(_abi.RepeatHipError) as e:
        _abi.Server.request_bytes(args(small), bad, opts)
    assert e.value.code == _abi.RP_EINVAL


@pytest.mark.parametrize("precision", [_abi.RP_PREC_FP64, _abi.RP_PREC_OZAKI, _abi.RP_PREC_OZAKI_REFINED,
                                       _abi.RP_PREC_FP64_FACT])
def test_chunk_plan_row_splits_hold_whole_gram_stages(precision):
    """Every row split of every plan is whole stages of its Gram kernel (128 rows for 3 slices,
    64 for 4, 16 for fp64) and short enough for exact int32 slice groups.  The kept counts of small
    serving cells (10574, 17223, 31782, 47322 for 20-64 atoms) once gave 3-slice plans with
    1088-row splits that the Gram rejected."""
    counts = sorted(set(range(512, 70001, 512)) | set(range(1, 70001, 997)) |
                    {10574, 17223, 31782, 47322, 8704, 97963, 57297, 252456, 943000})
    for n_atoms in (4, 20, 50, 64, 100, 199, 200, 500, 1000, 1600, 5000):
        for m in counts:
            c = _abi.chunk_plan(m, n_atoms, 754, precision)
            rps = c["rows"] // c["splits"]
            assert c["rows"] % (c["splits"] * c["stage_rows"]) == 0, (n_atoms, m, c)
            assert rps >= c["stage_rows"] and c["n_chunks"] * c["rows"] >= m
            assert c["np"] % 128 == 0 and c["np"] >= n_atoms + 2
            if precision == _abi.RP_PREC_FP64_FACT:
                assert c["slices"] == 0 and c["rows"] % 512 == 0 and c["rows"] * 8 * c["np"] <= 2**31
            elif precision != _abi.RP_PREC_FP64:
                p = c["slices"]
                assert p == (3 if precision == _abi.RP_PREC_OZAKI_REFINED else 4)
                assert c["rows"] % 512 == 0
                # a group's per-row sum is at most 2*127*64 + (P-2)*64*64; rps of them fit in int32
                assert (2 * 127 * 64 + (p - 2) * 64 * 64) * rps <= 2**31 - 1
    assert _abi.chunk_plan(10574, 20, 754, _abi.RP_PREC_OZAKI_REFINED)["rows"] % 128 == 0
    with pytest.raises(_abi.RepeatHipError):
        _abi.chunk_plan(0, 20, 754, precision)


def test_stream_splitting_sums_the_same_potential():
    """The serving stream's Ewald splitting (eta 0.6, Rc 8, |k| <= 5) gives the potential of the
    default splitting (eta 0.3, Rc 20, |k| <= 2.5): Ewald sums do not depend on eta once converged."""
    from pyrepeat_amd import synth
    s = synth.tiny(seed=4, n_atoms=6, n=24, L=12.0, triclinic=True)
    a = synth.ewald_esp(s.n_grid, s.axis, s.atom_pos, s.true_charges)
    b = synth.ewald_esp(s.n_grid, s.axis, s.atom_pos, s.true_charges, **synth.STREAM_SPLIT)
    assert np.max(np.abs(a - b)) < 1e-7 * max(1.0, np.max(np.abs(a)))


def test_erf_polynomial_matches_library_erfc():
    """The split path's real-space term erfc(sqrt(a) r)/r = 1/r - sqrt(a) P(a r^2), erf(x) = x P(x^2)
    (csrc/rp_erf_poly.h), against scipy.special.erfc: a = (pi/Rc)^2 keeps a r^2 within [0, pi^2]."""
    from scipy.special import erfc
    text = open(os.path.join(ROOT, "pyrepeat_amd", "csrc", "rp_erf_poly.h")).read()
    deg = int(re.search(r"ERF_DEG = (\d+)", text).group(1))
    u0 = float(re.search(r"ERF_U0 = ([-0-9.e+]+);", text).group(1))
    body = text[text.index("kErfPoly[ERF_DEG + 1] = {") + 25:text.rindex("};")]
    c = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    assert len(c) == deg + 1 and abs(u0 - np.pi ** 2 / 2) < 1e-15

    def P(u):
        s = u - u0
        acc = np.full_like(u, c[deg])
        for d in range(deg - 1, -1, -1):
            acc = acc * s + c[d]
        return acc
    x = np.linspace(0.0, np.pi, 20001)
    assert np.max(np.abs((1.0 - x * P(x * x)) - erfc(x))) < 2e-15
    for rc in (9.0, 12.0, 20.0):
        sa = np.pi / rc
        r = np.linspace(0.5, rc, 5001)
        got = 1.0 / r - sa * P(sa * sa * r * r)
        assert np.max(np.abs(got - erfc(sa * r) / r)) < 4e-15


def test_sphere_enumeration_of_the_image_sum_covers_the_cutoff():
    """The lattice-point ranges of the multi-image sum (rp_fit_common.h, EpiArgs::image_sum_sphere: n2,
    then n1 given n2, then n0 given both, from the cell metric in triangular form) hold every shift
    within Rc of the (point, atom) pair, on skewed cells and cutoffs from half a cell to several cells.
    Brute force over the shift list's box is the reference; a sign or index slip in the metric form
    drops points here."""
    rng = np.random.default_rng(17)
    for trial in range(60):
        a, b, c = rng.uniform(6.0, 14.0, 3)
        al, be, ga = rng.uniform(65.0, 115.0, 3)
        B = synth.triclinic_box(a, b, c, al, be, ga)
        rc = rng.uniform(0.4, 2.5) * min(a, b, c)
        G = B @ B.T
        q00 = G[0, 0]
        a01, a02 = G[0, 1] / q00, G[0, 2] / q00
        q11 = G[1, 1] - G[0, 1] * a01
        a12 = (G[1, 2] - G[0, 2] * a01) / q11
        q22 = G[2, 2] - q00 * a02 ** 2 - q11 * a12 ** 2
        assert min(q00, q11, q22) > 0
        w = rng.uniform(-2, 2, 3)
        form = q22 * w[2] ** 2 + q11 * (w[1] + a12 * w[2]) ** 2 + q00 * (w[0] + a01 * w[1] + a02 * w[2]) ** 2
        assert abs(form - np.sum((w @ B) ** 2)) < 1e-9 * form
        nmax = [int(np.floor(rc / np.linalg.norm(B[i]))) + 1 for i in range(3)]
        R2 = rc * rc

        def root_up(x):
            return np.sqrt(x) * (1 + 1e-6) + 1e-9 if x > 0 else 1e-9
        reach2 = rc / np.sqrt(q22) + 1e-9
        for _ in range(20):
            d = rng.uniform(-1.0, 1.0, 3)
            brute = set()
            for n0 in range(-nmax[0], nmax[0] + 1):
                for n1 in range(-nmax[1], nmax[1] + 1):
                    for n2 in range(-nmax[2], nmax[2] + 1):
                        if np.sum(((d - (n0, n1, n2)) @ B) ** 2) <= R2:
                            brute.add((n0, n1, n2))
            got = set()
            for n2 in range(max(-nmax[2], int(np.ceil(d[2] - reach2))), min(nmax[2], int(np.floor(d[2] + reach2))) + 1):
                w2 = d[2] - n2
                rem2 = R2 * (1 + 1e-12) - q22 * w2 * w2
                if rem2 < 0:
                    continue
                c1, s1 = d[1] + a12 * w2, root_up(rem2 / q11)
                for n1 in range(max(-nmax[1], int(np.ceil(c1 - s1))), min(nmax[1], int(np.floor(c1 + s1))) + 1):
                    w1 = d[1] - n1
                    rem1 = rem2 - q11 * (w1 + a12 * w2) ** 2
                    if rem1 < 0:
                        continue
                    c0, s0 = d[0] + a01 * w1 + a02 * w2, root_up(rem1 / q00)
                    for n0 in range(max(-nmax[0], int(np.ceil(c0 - s0))), min(nmax[0], int(np.floor(c0 + s0))) + 1):
                        got.add((n0, n1, n2))
            assert brute <= got, (trial, brute - got)
            # the ranges follow the sphere, not the whole box of candidates
            assert len(got) <= 2 * len(brute) + 12
