"""Parameter files and constraint maps of the REPEAT fit (host side).

File formats and their parse rules follow the reference exactly:
- symmetry file: one group per line, 1-based ids and ranges ``1-5,9``,
  ``#`` comments, first id is the base (``repeat.py:1814-1890``);
- RESP file: ``Sym strength target`` per line; a line with >= 3 tokens that
  does not parse makes the reader return ``{}`` (``repeat.py:1892-1904``);
- QEq file: ``Sym chi hardness`` with the same rule (``repeat.py:1906-1918``).

``constraint_arrays`` turns them into the merge sets the device solver takes
(``rp_constraints`` in ``include/repeat_hip.h``), reproducing the reference's
in-place sequence: columns and ``b`` merged tree by tree, weights incremented
per linked atom, then rows merged in the same order, linked atoms dropped
(``repeat.py:2445-2470``), charges copied base -> linked in order
(``repeat.py:2472-2489``).
"""
from __future__ import annotations

from collections import Counter

import numpy as np

from .tables import DEFAULT_QEQ_PARAMS, element_symbol


def parse_atom_ranges(range_str):
    """'1-5,9,10' -> 0-based ids, duplicates dropped in first-seen order (``repeat.py:1814-1841``)."""
    ids = []
    for part in range_str.split(","):
        part = part.strip()
        if not part:
            continue
        if "-" in part:
            lo, hi = part.split("-", 1)
            ids.extend(range(int(lo.strip()), int(hi.strip()) + 1))
        else:
            ids.append(int(part))
    seen, out = set(), []
    for i in ids:
        z = i - 1
        if z not in seen:
            seen.add(z)
            out.append(z)
    return out


def read_symmetry(filename):
    """``{'trees': [{'n_equiv', 'base', 'linked'}], 'n_independent'}`` or ``None`` on error."""
    try:
        with open(filename) as f:
            lines = [ln.strip() for ln in f if ln.strip() and not ln.startswith("#")]
        trees = []
        for line in lines:
            if "#" in line:
                line = line.split("#")[0].strip()
            if not line:
                continue
            ids = parse_atom_ranges(line)
            if not ids:
                continue
            trees.append({"n_equiv": len(ids), "base": ids[0], "linked": ids[1:]})
        return {"trees": trees, "n_independent": len(trees)}
    except Exception as exc:  # the reference prints a warning and continues without symmetry
        print(f"Warning: Error reading symmetry file: {exc}")
        return None


def _read_keyed(filename, k1, k2):
    out = {}
    try:
        with open(filename) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3:
                    out[parts[0]] = {k1: float(parts[1]), k2: float(parts[2])}
        return out
    except Exception:
        return {}


def read_resp_params(filename):
    return _read_keyed(filename, "strength", "target")


def read_qeq_params(filename):
    return _read_keyed(filename, "electroneg", "hardness")


def symmetry_maps(trees, n_atoms):
    """Merge sets (with multiplicities), weights, independent atoms and the expansion map."""
    sets = [Counter({a: 1}) for a in range(n_atoms)]
    weights = np.ones(n_atoms)
    for t in trees or []:
        base = t["base"]
        for ln in t["linked"]:
            if not (0 <= base < n_atoms and 0 <= ln < n_atoms):
                raise IndexError(f"symmetry atom id out of range (1..{n_atoms})")
            sets[base] = sets[base] + sets[ln]
            weights[base] += 1.0
    linked = set()
    for t in trees or []:
        linked.update(t["linked"])
    independent = [i for i in range(n_atoms) if i not in linked]
    src = np.full(n_atoms, -1, dtype=np.int32)
    for i, a in enumerate(independent):
        src[a] = i
    for t in trees or []:
        for ln in t["linked"]:
            src[ln] = src[t["base"]]
    return independent, [sets[a] for a in independent], weights[independent], src


def constraint_arrays(n_atoms, atom_types, q_tot=0.0, symm=None, resp=None, qeq=None):
    """Arrays for ``rp_fit_solve`` (see ``rp_constraints``)."""
    trees = symm["trees"] if symm else None
    independent, sets, weights, src = symmetry_maps(trees, n_atoms)
    nr = len(independent)
    ptr = np.zeros(nr + 1, dtype=np.int32)
    red_src, red_mult = [], []
    for i, s in enumerate(sets):
        for a in sorted(s):
            red_src.append(a)
            red_mult.append(float(s[a]))
        ptr[i + 1] = len(red_src)
    types = np.asarray(atom_types)[independent]
    diag = np.zeros(nr)
    rhs = np.zeros(nr)
    if resp:
        for i, z in enumerate(types):
            sym = element_symbol(z)
            if sym in resp:
                diag[i] += resp[sym]["strength"]
                rhs[i] += resp[sym]["strength"] * resp[sym]["target"]
    elif qeq:
        for i, z in enumerate(types):
            sym = element_symbol(z)
            if sym in qeq:
                chi, hard = qeq[sym]["electroneg"], qeq[sym]["hardness"]
            elif sym in DEFAULT_QEQ_PARAMS:
                chi, hard = DEFAULT_QEQ_PARAMS[sym]
            else:
                continue
            rhs[i] += chi
            diag[i] += hard
    return {"q_tot": float(q_tot), "n_reduced": nr, "red_ptr": ptr,
            "red_src": np.asarray(red_src, dtype=np.int32), "red_mult": np.asarray(red_mult, dtype=np.float64),
            "weights": np.asarray(weights, dtype=np.float64), "diag_add": diag, "rhs_add": rhs,
            "expand_src": src, "independent": np.asarray(independent, dtype=np.int64)}


def reference_cpu_filter_is_exact(axis, n_grid, vdw_factor, vdw_max):
    """Whether the reference's CPU filter gives the exact keep mask for this grid and these options.

    That filter (``repeat.py:2115-2180``) tests a batch of points only against the effective atoms with
    ``|z_atom - mid| < half + max_r + 1``, where ``mid`` and ``half`` come from the batch's layer indices
    times ``axis[2][2]`` alone, and its batches follow its worker count (the host's cores).  A point's z
    is ``x + k axis[2][2]`` with ``x = i axis[0][2] + j axis[1][2]``.  With ``axis[2][2] >= 0`` every
    ``k axis[2][2]`` of a batch lies within ``half`` of ``mid``, so an atom the test drops is farther than
    ``max_r + 1 - |x|`` from every point of the batch in z alone; if ``|x| < 1`` over the whole grid and
    ``r_min <= r_max <= max_r`` (``vdw_factor <= vdw_max``), it is neither within ``r_min`` nor within
    ``r_max`` of any of them, and dropping it changes nothing (a batch with no atom left keeps nothing, as
    the full test would not either).  Then the CPU mask equals the exact predicate that this build and
    the reference's GPU filter (``repeat.py:349-437``) apply, whatever the worker count.  Otherwise the
    reference's own CPU mask can differ from that, and depends on its worker count.
    """
    axis = np.asarray(axis, dtype=np.float64)
    nx, ny = int(n_grid[0]), int(n_grid[1])
    if not vdw_factor <= vdw_max or axis[2][2] < 0.0:
        return False
    corners = [0.0, (nx - 1) * axis[0][2], (ny - 1) * axis[1][2], (nx - 1) * axis[0][2] + (ny - 1) * axis[1][2]]
    return max(abs(c) for c in corners) < 0.999

