"""Input files of the fit: Gaussian cube (``read_gauss_cube``, ``repeat.py:770-804``) and VASP LOCPOT
(``read_vasp_potential``, ``repeat.py:654-768``), with the reference's format detection
(``ChargeCalculator._read_data``, ``repeat.py:1790-1812``).

Cube conventions (as the reference reads them):
- line 3: atom count (origin read but ignored by the fit, as the reference ignores it);
- lines 4-6: ``n_i  a_ix a_iy a_iz`` grid step vectors in Bohr;
- atom lines: ``Z  charge  x y z`` (charge column ignored);
- values: x outer, z inner.

The value blocks are parsed by the library's multithreaded parser (``rp_cube_read`` and
``rp_parse_values``), which returns for every token bitwise what ``float()`` returns.  A reader's
result is a dict with the reference's keys; the ESP is kept once, in file order, as ``esp_cube``
with ``esp_order`` naming that order (``RP_ORDER_CUBE`` for cubes, ``RP_ORDER_FLAT``, x fastest, for
LOCPOT), and ``V_pot`` (the reference's ``(nx*ny, nz)`` array) is built on demand by
:func:`v_pot_view`.
"""
from __future__ import annotations

import bz2
import glob
import gzip
import os

import numpy as np

from . import _abi
from .tables import ATOM_SYMBOLS

ANG_TO_BOHR = 1.889725985        # repeat.py:687
LOCPOT_FACTOR = -1 / 27.212      # eV -> Hartree with the sign of the ESP (repeat.py:730)


def read_gauss_cube(filename, n_threads=0):
    c = _abi.read_cube(filename, n_threads)
    c["esp_order"] = _abi.RP_ORDER_CUBE
    return c


def write_gauss_cube(path, cube, comment=("cube", "written by pyrepeat_amd"), origin=(0.0, 0.0, 0.0)):
    """Cube text that :func:`read_gauss_cube` (and the reference's reader) reads back exactly:
    step vectors and positions with 17 significant digits, values ``%.10E``, six per line."""
    nx, ny, nz = (int(v) for v in cube["n_grid"])
    esp = np.asarray(cube["esp_cube"], dtype=np.float64)
    if cube.get("esp_order", _abi.RP_ORDER_CUBE) == _abi.RP_ORDER_FLAT:
        esp = esp.reshape(nz, ny, nx).transpose(2, 1, 0).ravel()
    axis = np.asarray(cube["axis_vector"], dtype=np.float64)
    with open(path, "w") as f:
        f.write(f"{comment[0]}\n{comment[1]}\n")
        f.write(f"{int(cube['n_atoms']):5d} {origin[0]:12.6f} {origin[1]:12.6f} {origin[2]:12.6f}\n")
        for i, n in enumerate((nx, ny, nz)):
            f.write(f"{n:5d} " + " ".join(f"{v:.17g}" for v in axis[i]) + "\n")
        for z, p in zip(cube["atom_index"], np.asarray(cube["atom_pos"], dtype=np.float64)):
            f.write(f"{int(z):5d} {0.0:12.6f} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        rows = esp.reshape(nx * ny, nz)
        for row in rows:
            for s in range(0, nz, 6):
                f.write(" ".join(f"{v:.10E}" for v in row[s:s + 6]) + "\n")


def compressed_open(filename):
    """``compressed_open`` (``repeat.py:654-666``): the file, or its ``.gz`` / ``.bz2`` twin, as binary."""
    names = glob.glob(filename) + glob.glob(filename + ".gz") + glob.glob(filename + ".bz2")
    if not names:
        raise FileNotFoundError(f"File {filename} not found")
    name = names[0]
    if name[-4:] == ".bz2":
        return bz2.BZ2File(name, "rb")
    if name[-3:] == ".gz":
        return gzip.open(name, "rb")
    return open(name, "rb")


def read_vasp_potential(filename, n_threads=0):
    """``read_vasp_potential`` (``repeat.py:668-768``).

    Kept as the reference has it: the lattice is scaled by ``scale * 1.889725985``; up to 20 lines
    after it are searched for ``Direct``, alphabetic tokens are element symbols and numeric ones
    counts; positions are fractional (``(a0 f0 + a1 f1) + a2 f2`` per component); every value after
    the grid line is parsed and multiplied by ``-1/27.212``, the first ``nx*ny*nz`` of them in x-fastest
    order; symbols outside the table map to 0."""
    with compressed_open(filename) as fb:
        def line():
            return fb.readline().decode()
        line()   # system name
        scale = float(line()) * ANG_TO_BOHR
        lattice = [[scale * float(v) for v in line().split()] for _ in range(3)]
        test = ""
        for _ in range(20):
            t = line()
            if "Direct" in t:
                break
            test += t
        tokens = "".join(test.splitlines()).split()
        types = [t for t in tokens if t.isalpha()]
        counts = [int(t) for t in tokens if t.isnumeric()]
        m = lattice
        atoms = []
        for _ in range(sum(counts)):
            c = [float(v) for v in line().split()]
            atoms.append([m[0][0] * c[0] + m[1][0] * c[1] + m[2][0] * c[2],
                          m[0][1] * c[0] + m[1][1] * c[1] + m[2][1] * c[2],
                          m[0][2] * c[0] + m[1][2] * c[1] + m[2][2] * c[2]])
        line()   # blank line
        ngrid = [int(v) for v in line().split()]
        rest = fb.read()
    nx, ny, nz = ngrid
    n_total = nx * ny * nz
    vals, found = _abi.parse_values(rest, cap=n_total, n_threads=n_threads)
    if found < n_total:
        raise ValueError(f"LOCPOT holds {found} values, grid needs {n_total}")
    esp = LOCPOT_FACTOR * vals
    atom_index = []
    for sym, cnt in zip(types, counts):
        atom_index.extend([ATOM_SYMBOLS.index(sym) if sym in ATOM_SYMBOLS else 0] * cnt)
    axis = np.array([[v / ngrid[i] for v in lattice[i]] for i in range(3)], dtype=np.float64)
    return {"n_atoms": len(atom_index), "n_grid": np.array(ngrid, dtype=np.int64),
            "atom_index": np.array(atom_index, dtype=np.int64), "atom_pos": np.array(atoms, dtype=np.float64),
            "axis_vector": axis, "esp_cube": esp, "esp_order": _abi.RP_ORDER_FLAT}


VASP_NAMES = {"locpot", "locpot.gz", "locpot.bz2", "locpot.z"}


def read_data(filename, n_threads=0):
    """``ChargeCalculator._read_data`` (``repeat.py:1790-1812``): a name containing ``locpot`` is read
    as LOCPOT (falling back to the cube reader, with the reference's warnings, if that fails);
    everything else as a Gaussian cube."""
    base = os.path.basename(filename).lower()
    if base in VASP_NAMES or "locpot" in base:
        try:
            return read_vasp_potential(filename, n_threads)
        except Exception as e:  # the reference catches every failure here too
            print(f"Warning: Failed to read as VASP file: {e}")
            print("Trying to read as cube file...")
    return read_gauss_cube(filename, n_threads)


def v_pot_view(cube):
    """The reference's ``V_pot`` array, shape ``(nx*ny, nz)``, ``V_pot[i + j*nx, k]``."""
    nx, ny, nz = (int(v) for v in cube["n_grid"])
    esp = np.asarray(cube["esp_cube"])
    if cube.get("esp_order", _abi.RP_ORDER_CUBE) == _abi.RP_ORDER_FLAT:
        return esp.reshape(nz, nx * ny).T
    return esp.reshape(nx, ny, nz).transpose(1, 0, 2).reshape(nx * ny, nz)


def cube_from_system(system):
    """Cube dict from an in-memory :class:`pyrepeat_amd.synth.System` (no text round trip)."""
    return {"n_atoms": system.n_atoms, "n_grid": np.asarray(system.n_grid, dtype=np.int64),
            "atom_index": np.asarray(system.atomic_numbers, dtype=np.int64),
            "atom_pos": np.asarray(system.atom_pos, dtype=np.float64),
            "axis_vector": np.asarray(system.axis, dtype=np.float64), "esp_cube": system.esp_cube,
            "esp_order": _abi.RP_ORDER_CUBE}


__all__ = ["read_gauss_cube", "write_gauss_cube", "read_vasp_potential", "read_data", "v_pot_view",
           "cube_from_system", "compressed_open"]
