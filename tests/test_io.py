"""Input files (CPU): the library's value parser against Python's float(), the cube reader against a
plain per-token parse of the same file, and the LOCPOT reader against a known-answer file whose values
and positions are fixed by construction (read_vasp_potential, repeat.py:654-768)."""
import bz2
import gzip
import struct

import numpy as np
import pytest

from pyrepeat_amd import _abi, synth
from pyrepeat_amd.cube import (ANG_TO_BOHR, LOCPOT_FACTOR, read_data, read_gauss_cube, read_vasp_potential,
                               v_pot_view, write_gauss_cube)


def _bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def _tokens(rng, n):
    out = []
    for _ in range(n):
        kind = rng.integers(0, 8)
        x = float(rng.standard_normal()) * 10.0 ** int(rng.integers(-30, 30))
        if kind == 0:
            out.append(f"{x:.10E}")
        elif kind == 1:
            out.append(repr(x))
        elif kind == 2:
            out.append(f"{x:.25e}")                      # more digits than the exact fast path takes
        elif kind == 3:
            out.append(f"{x:.3f}")
        elif kind == 4:
            out.append(str(int(rng.integers(-10**6, 10**6))))
        elif kind == 5:
            out.append(f"{x:.6E}")
        elif kind == 6:
            e = int(rng.integers(-330, 320))
            out.append(f"{rng.uniform(1, 10):.15f}e{e}")  # overflow, underflow, subnormals
        else:
            out.append(rng.choice(["0", "-0.0", "+1", ".5", "5.", "0.000123", "007.50", "1E5", "inf", "-inf",
                                   "Infinity", "1e400", "-1e-400", "4.9e-324", "1.7976931348623157e308"]))
    return out


def test_value_parser_is_bitwise_float():
    rng = np.random.default_rng(5)
    toks = _tokens(rng, 20000)
    text = ("\n".join(" ".join(toks[i:i + 6]) for i in range(0, len(toks), 6)) + "\n").encode()
    want = np.array([float(t) for t in toks])
    for nt in (1, 3, 16):
        got, n = _abi.parse_values(text, n_threads=nt)
        assert n == len(toks)
        np.testing.assert_array_equal(_bits(got), _bits(want))


def test_value_parser_thread_cuts_never_split_a_token():
    # several MiB so that the text is cut into many ranges, with ragged whitespace
    rng = np.random.default_rng(6)
    vals = rng.standard_normal(300_000) * 1e-3
    seps = rng.choice([" ", "  ", "\n", "\t", " \r\n"], size=vals.size)
    text = "".join(f"{v:.10E}{s}" for v, s in zip(vals, seps)).encode()
    got, n = _abi.parse_values(text, n_threads=13)
    assert n == vals.size
    np.testing.assert_array_equal(_bits(got), _bits([float(f"{v:.10E}") for v in vals]))
    # cap keeps the first values and still counts all of them
    part, n2 = _abi.parse_values(text, cap=1000, n_threads=13)
    assert n2 == vals.size and np.array_equal(_bits(part), _bits(got[:1000]))
    empty, n3 = _abi.parse_values(b"  \n\t ")
    assert n3 == 0 and empty.size == 0


@pytest.mark.parametrize("tok", ["abc", "1.0.0", "0x10", "1e", "--1", "1.5D+01", "+", ".", "1e5x", "nanx"])
def test_value_parser_rejects_what_float_rejects(tok):
    with pytest.raises(ValueError):
        float(tok)
    with pytest.raises(_abi.RepeatHipError):
        _abi.parse_values(f"1.0 {tok} 2.0".encode())


def _plain_cube_parse(path):
    """repeat.py:770-804 as written: header by lines, then float() on every remaining token."""
    with open(path) as f:
        f.readline()
        f.readline()
        na = int(f.readline().split()[0])
        ng = [int(f.readline().split()[0]) for _ in range(3)]
        for _ in range(na):
            f.readline()
        data = [float(x) for line in f for x in line.split()]
    return np.array(data[:int(np.prod(ng))])


def test_cube_reader_matches_per_token_float(tmp_path):
    s = synth.tiny(seed=3, n_atoms=5, n=14, L=9.0, triclinic=True)
    p = tmp_path / "s.cube"
    s.write_cube(str(p))
    c = read_gauss_cube(str(p), n_threads=4)
    np.testing.assert_array_equal(_bits(c["esp_cube"]), _bits(_plain_cube_parse(p)))
    np.testing.assert_array_equal(c["atom_index"], s.atomic_numbers)
    np.testing.assert_array_equal(c["atom_pos"], s.atom_pos)      # written with 17 digits: exact
    np.testing.assert_array_equal(c["axis_vector"], s.axis)
    # write -> read -> write is a fixed point
    q = tmp_path / "again.cube"
    write_gauss_cube(str(q), c, comment=(f"{s.name} synthetic ESP", "exact Ewald potential plus noise"))
    assert q.read_bytes() == p.read_bytes()


def test_cube_reader_errors(tmp_path):
    s = synth.tiny(seed=3, n_atoms=2, n=6, L=7.0)
    p = tmp_path / "s.cube"
    s.write_cube(str(p))
    lines = p.read_text().splitlines(True)
    short = tmp_path / "short.cube"
    short.write_text("".join(lines[:-1]))
    with pytest.raises(_abi.RepeatHipError, match="values, grid needs"):
        read_gauss_cube(str(short))
    extra = tmp_path / "extra.cube"
    extra.write_text("".join(lines) + "1.0 2.0\n")               # trailing values are ignored
    np.testing.assert_array_equal(read_gauss_cube(str(extra))["esp_cube"], read_gauss_cube(str(p))["esp_cube"])
    neg = tmp_path / "neg.cube"
    neg.write_text("".join(lines[:2]) + "   -2 0.0 0.0 0.0\n" + "".join(lines[3:]))
    with pytest.raises(_abi.RepeatHipError, match="negative atom count"):
        read_gauss_cube(str(neg))
    with pytest.raises(_abi.RepeatHipError, match="cannot open"):
        read_gauss_cube(str(tmp_path / "missing.cube"))


def _write_locpot(path, lattice, scale, symbols, counts, frac, grid, values_ev, opener=open):
    """VASP 5 LOCPOT text: name, scale, lattice (Angstrom), symbols, counts, Direct, positions, blank line,
    grid, values x fastest, five per line."""
    nx, ny, nz = grid
    out = ["synthetic LOCPOT\n", f"{scale:.10f}\n"]
    out += [" ".join(f"{v:.12f}" for v in row) + "\n" for row in lattice]
    out += ["   " + "   ".join(symbols) + "\n", "   " + "   ".join(str(c) for c in counts) + "\n", "Direct\n"]
    out += [" ".join(f"{v:.12f}" for v in f) + "\n" for f in frac]
    out += ["\n", f"   {nx}   {ny}   {nz}\n"]
    flat = values_ev.ravel()                      # already x fastest
    out += [" ".join(f"{v:.11E}" for v in flat[i:i + 5]) + "\n" for i in range(0, flat.size, 5)]
    with opener(str(path), "wt") as f:
        f.write("".join(out))


def test_locpot_known_answer(tmp_path):
    lattice = np.array([[4.0, 0.0, 0.0], [0.5, 4.5, 0.0], [0.25, 0.3, 5.0]])   # Angstrom
    scale = 1.25
    symbols, counts = ["Zn", "O", "H"], [1, 2, 1]
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.5, 0.125], [0.75, 0.25, 0.5], [0.5, 0.5, 0.875]])
    grid = (6, 5, 4)
    i, j, k = np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij")
    v_ev = (0.5 * i - 0.25 * j + 0.125 * k + 0.0625).astype(np.float64)     # exact in binary and in %.11E
    flat = v_ev.transpose(2, 1, 0).ravel()                                   # x fastest
    for name, opener in (("LOCPOT", open), ("LOCPOT.gz", gzip.open), ("LOCPOT.bz2", bz2.open)):
        d = tmp_path / name.split(".")[-1]
        d.mkdir()
        path = d / name
        _write_locpot(path, lattice, scale, symbols, counts, frac, grid, flat, opener)
        # the reference opens LOCPOT, LOCPOT.gz or LOCPOT.bz2 from the bare name
        c = read_vasp_potential(str(d / "LOCPOT"))
        assert c["n_atoms"] == 4 and list(c["atom_index"]) == [30, 8, 8, 1]
        s = scale * ANG_TO_BOHR
        lat = s * lattice
        np.testing.assert_allclose(c["atom_pos"], frac @ lat, rtol=0, atol=1e-13)
        np.testing.assert_allclose(c["axis_vector"], lat / np.array(grid)[:, None], rtol=1e-15)
        assert c["esp_order"] == _abi.RP_ORDER_FLAT
        np.testing.assert_array_equal(_bits(c["esp_cube"]), _bits(LOCPOT_FACTOR * flat))
        # V_pot[ix + iy nx, iz] (repeat.py:745-753)
        vp = v_pot_view(c)
        assert vp.shape == (30, 4)
        assert vp[2 + 3 * 6, 1] == LOCPOT_FACTOR * v_ev[2, 3, 1]
        # detection by name reads it as LOCPOT
        assert read_data(str(path))["esp_order"] == _abi.RP_ORDER_FLAT


def test_locpot_name_falls_back_to_cube(tmp_path, capsys):
    s = synth.tiny(seed=2, n_atoms=2, n=6, L=7.0)
    p = tmp_path / "my_locpot_run.cube"
    s.write_cube(str(p))
    c = read_data(str(p))
    assert "Failed to read as VASP file" in capsys.readouterr().out
    assert c["esp_order"] == _abi.RP_ORDER_CUBE and c["n_atoms"] == 2


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_locpot_fit_equals_cube_order_fit(tmp_path, gpu_ctx, oracle, precision):
    """A LOCPOT goes to the device in its x-fastest order (RP_ORDER_FLAT); the fit must be bitwise the
    fit of the same values in cube order, and match the oracle."""
    from pyrepeat_amd.calculator import ChargeCalculator
    s = synth.tiny(seed=31, n_atoms=6, n=14, L=11.0, triclinic=True)
    nx, ny, nz = (int(v) for v in s.n_grid)
    lat_ang = np.asarray(s.axis) * np.array([nx, ny, nz])[:, None] / ANG_TO_BOHR
    frac = np.linalg.solve(np.asarray(s.axis).T * np.array([nx, ny, nz]), np.asarray(s.atom_pos).T).T % 1.0
    syms = {1: "H", 6: "C", 7: "N", 8: "O"}
    order = np.argsort(s.atomic_numbers, kind="stable")
    types, counts = np.unique(np.asarray(s.atomic_numbers)[order], return_counts=True)
    v_ev = np.asarray(s.esp_flat(), dtype=np.float64) / LOCPOT_FACTOR
    path = tmp_path / "LOCPOT"
    _write_locpot(path, lat_ang, 1.0, [syms[int(z)] for z in types], list(counts), frac[order], (nx, ny, nz), v_ev)
    c_flat = read_vasp_potential(str(path))
    c_cube = dict(c_flat)
    c_cube["esp_cube"] = np.asarray(c_flat["esp_cube"]).reshape(nz, ny, nx).transpose(2, 1, 0).ravel()
    c_cube["esp_order"] = _abi.RP_ORDER_CUBE
    print(f"locpot fit {precision}", flush=True)
    q = []
    for c in (c_flat, c_cube):
        calc = ChargeCalculator(cube=c, precision=precision, context=gpu_ctx)
        q.append(calc.fit_charges())
        calc.close()
    assert np.array_equal(q[0], q[1])
    r = oracle.fit(c_flat["n_grid"], c_flat["axis_vector"], c_flat["atom_index"], c_flat["atom_pos"],
                   np.asarray(c_flat["esp_cube"]))
    np.testing.assert_allclose(q[0], r.charges, atol=1e-9 if precision == "fp64" else 1e-5)
