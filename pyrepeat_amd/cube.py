"""Gaussian cube reader with the reference's conventions (``read_gauss_cube``, ``repeat.py:770-804``).

- line 3: atom count (origin ignored, as the reference ignores it);
- lines 4-6: ``n_i  a_ix a_iy a_iz`` grid step vectors in Bohr;
- atom lines: ``Z  charge  x y z`` (charge column ignored);
- values: x outer, z inner.

The values are returned once, in file order (``esp_cube``, what the device
filter reads); ``V_pot`` is the reference's ``(nx*ny, nz)`` view of the same
numbers (``V_pot[i + j*nx, k]``), built on demand.
"""
from __future__ import annotations

import numpy as np


def read_gauss_cube(filename):
    with open(filename, "r") as f:
        f.readline()
        f.readline()
        parts = f.readline().split()
        n_atoms = int(parts[0])
        if n_atoms < 0:
            raise ValueError("cube files with a negative atom count are not supported")
        n_grid = np.zeros(3, dtype=np.int64)
        axis = np.zeros((3, 3))
        for i in range(3):
            parts = f.readline().split()
            n_grid[i] = int(parts[0])
            axis[i] = [float(x) for x in parts[1:4]]
        Z = np.zeros(n_atoms, dtype=np.int64)
        pos = np.zeros((n_atoms, 3))
        for i in range(n_atoms):
            parts = f.readline().split()
            Z[i] = int(parts[0])
            pos[i] = [float(x) for x in parts[2:5]]
        data = np.array(f.read().split(), dtype=np.float64)
    n_total = int(np.prod(n_grid))
    if data.size < n_total:
        raise ValueError(f"cube holds {data.size} values, grid needs {n_total}")
    esp = data[:n_total]
    return {"n_atoms": n_atoms, "n_grid": n_grid, "atom_index": Z, "atom_pos": pos, "axis_vector": axis,
            "esp_cube": esp}


def v_pot_view(cube):
    """The reference's ``V_pot`` array, shape ``(nx*ny, nz)``, ``V_pot[i + j*nx, k]``."""
    nx, ny, nz = (int(v) for v in cube["n_grid"])
    return cube["esp_cube"].reshape(nx, ny, nz).transpose(1, 0, 2).reshape(nx * ny, nz)


def cube_from_system(system):
    """Cube dict from an in-memory :class:`pyrepeat_amd.synth.System` (no text round trip)."""
    return {"n_atoms": system.n_atoms, "n_grid": np.asarray(system.n_grid, dtype=np.int64),
            "atom_index": np.asarray(system.atomic_numbers, dtype=np.int64),
            "atom_pos": np.asarray(system.atom_pos, dtype=np.float64),
            "axis_vector": np.asarray(system.axis, dtype=np.float64), "esp_cube": system.esp_cube}
