"""``ChargeCalculator``: the reference's fit-path interface on the HIP library.

Same constructor arguments, method names and results as ``repeat.py``'s
``ChargeCalculator`` (``repeat.py:1733-2535``) for the REPEAT fit:

- ``__init__`` reads the cube, runs the vdW filter and the Ewald setup on the
  device (``repeat.py:1736-1786``);
- ``fit_charges()`` streams the basis and solves the constrained least squares
  on the device, returning one charge per atom (``repeat.py:2184-2324``);
- ``compute_stats(charges, need_v_coul)`` returns ``rms_error``, ``avrg_qm``,
  ``avrg_coul`` and optionally ``V_coul`` (``repeat.py:2326-2441``);
- ``compute_qeq_charges(param_set=None)`` is the QEq charge model
  (``repeat.py:2538-2634``): Coulomb matrix and solve on the device.

``precision="fp64"`` is the plain fp64 path (the reference's ``--fp64``);
``precision="split"`` is the split form: the reciprocal sum factorized over the
lattice axes in fp64, the real-space sum in fp64 (erfc as one polynomial on
the argument range the cutoff allows), the Gram of the basis from int8 fixed-point slices on MFMA with exact int32
accumulation, fp64 across chunks and in the solve.
``precision="refined"`` is the split form with 3 slices followed by iterative
refinement of the solve against the fp64 basis (kept in HBM): its charges
converge to those of the fp64 normal equations.
``precision="fp64-factorized"`` is fp64 throughout with the split form's
factorized basis and an fp64 Gram (no slices, no refinement).
Every computation runs in ``librepeat_hip.so``; nothing falls back to the CPU.
"""
from __future__ import annotations

import sys
import time

import numpy as np

from . import _abi
from .cube import read_data, v_pot_view
from .params import (constraint_arrays, read_qeq_params, read_resp_params, read_symmetry,
                     reference_cpu_filter_is_exact)
from .tables import ATOM_SYMBOLS, DEFAULT_QEQ_PARAMS

PRECISIONS = {"fp64": _abi.RP_PREC_FP64, "split": _abi.RP_PREC_OZAKI, "refined": _abi.RP_PREC_OZAKI_REFINED,
              "fp64-factorized": _abi.RP_PREC_FP64_FACT}


class ChargeCalculator:
    def __init__(self, cube_file=None, fit_flag=1, vdw_factor=1.0, vdw_max=1000.0, R_cutoff=20.0, q_tot=0.0,
                 symm_file=None, resp_file=None, qeq_file=None, n_cores=None, block_k=64, use_jit=True,
                 precision="fp64", cube=None, context=None):
        del n_cores, block_k, use_jit  # CPU scheduling knobs of the reference; no meaning on the device
        self.cube_file = cube_file
        self.fit_flag = int(fit_flag)
        self.vdw_factor = float(vdw_factor)
        self.vdw_max = float(vdw_max)
        self.R_cutoff = float(R_cutoff)
        self.q_tot = float(q_tot)
        self.precision = precision
        self.cube = cube if cube is not None else read_data(cube_file)
        self.n_atoms = int(self.cube["n_atoms"])
        self.ctx = context or _abi.default_context()
        self._fit = _abi.Fit(self.ctx, self.cube["n_grid"], self.cube["axis_vector"], self.cube["atom_index"],
                             self.cube["atom_pos"], self.cube["esp_cube"],
                             self.cube.get("esp_order", _abi.RP_ORDER_CUBE), self.fit_flag,
                             PRECISIONS[precision], self.vdw_factor, self.vdw_max, self.R_cutoff)
        if not reference_cpu_filter_is_exact(self.cube["axis_vector"], self.cube["n_grid"], self.vdw_factor,
                                             self.vdw_max):
            print("Note: for this grid the reference's CPU filter prefilters atoms by a z-slab that ignores the "
                  "z components of the first two grid vectors; its mask then depends on its worker count.  This "
                  "build applies the exact keep predicate (as the reference's GPU filter does).", file=sys.stderr)
        self.symm_data = read_symmetry(symm_file) if symm_file else None
        self.resp_data = read_resp_params(resp_file) if resp_file else None
        self.qeq_data = read_qeq_params(qeq_file) if qeq_file else None
        self._grid = None
        self._sum_phi = None
        self._sum_V = None
        self._charges = None

    # ---- filter results (downloaded lazily; the fit itself keeps them on the device)
    def _grid_arrays(self):
        if self._grid is None:
            self._grid = self._fit.grid()
        return self._grid

    @property
    def grid_pos(self):
        return self._grid_arrays()[0]

    @property
    def V_pot(self):
        return self._grid_arrays()[1]

    @property
    def n_grid_points(self):
        return int(self._fit.info.n_keep)

    def V_pot_cube(self):
        return v_pot_view(self.cube)

    # ---- fit
    def constraints(self):
        return constraint_arrays(self.n_atoms, self.cube["atom_index"], self.q_tot, self.symm_data,
                                 self.resp_data, self.qeq_data)

    # A refined fit whose error bound stays above this after its steps is redone in fp64.
    REFINED_BOUND_LIMIT_E = 1e-5

    def fit_charges(self):
        self._fit.accumulate()
        q, sol = self._fit.solve(self.constraints())
        self.solve_report = self._fit.solve_report()
        bound = self.solve_report["error_bound"]
        if self.precision == "refined" and not bound <= self.REFINED_BOUND_LIMIT_E:
            # the refinement did not converge far enough within its steps: the fp64 path's solve
            # (Gram in fp64) of the same system instead
            print(f"Warning: refinement left an error bound of {bound:.2e} e after "
                  f"{self.solve_report['refine_steps']} steps; solving with the fp64 path.", file=sys.stderr)
            self._fit.close()
            self._fit = _abi.Fit(self.ctx, self.cube["n_grid"], self.cube["axis_vector"], self.cube["atom_index"],
                                 self.cube["atom_pos"], self.cube["esp_cube"],
                                 self.cube.get("esp_order", _abi.RP_ORDER_CUBE), self.fit_flag, _abi.RP_PREC_FP64,
                                 self.vdw_factor, self.vdw_max, self.R_cutoff)
            self._grid = None
            self._fit.accumulate()
            q, sol = self._fit.solve(self.constraints())
            self.solve_report = dict(self._fit.solve_report(), refined_bound_e=bound)
        sens = self.solve_report["sensitivity"]
        if sens > _abi.RP_SENSITIVITY_LIMIT_E:
            # ill-conditioned: fp64 rounding of the system alone moves these charges by about sens, so no
            # fp64 implementation (the reference's included) determines them more closely
            print(f"Warning: ill-conditioned fit; rounding the system in fp64 moves the charges by up to "
                  f"{sens:.1e} e.", file=sys.stderr)
        _, _, sum_phi, sum_v = self._fit.normal_equations()
        # The reference scatters the *unreduced* sum_phi as if it were reduced (repeat.py:2317-2320, 2491-2509).
        self._sum_phi = _expand_sum_phi_quirk(sum_phi, self.symm_data, self.n_atoms) if self.symm_data else sum_phi
        self._sum_phi_true = sum_phi
        self._sum_V = sum_v
        self._solution = sol
        self._charges = q
        return q

    def normal_equations(self):
        """``ATA, ATb, sum_phi, sum_V`` of the last fit (the reference's accumulators)."""
        return self._fit.normal_equations()

    def compute_stats(self, charges, need_v_coul=True):
        n = self.n_grid_points
        avrg_coul = float(self._sum_phi @ charges) / n if self._sum_phi is not None else None
        v = self.V_pot
        avrg_qm = float(np.mean(v))
        v_coul = self._fit.phi_matvec(charges)
        if avrg_coul is None:
            avrg_coul = float(np.mean(v_coul))
        diff = (v_coul - avrg_coul) - (v - avrg_qm)
        rms = float(np.sqrt(np.sum(diff ** 2) / n))
        out = {"rms_error": rms, "avrg_qm": avrg_qm, "avrg_coul": avrg_coul}
        if need_v_coul:
            out["V_coul"] = v_coul
        return out

    def rms_from_normal_equations(self, charges, corrected=False):
        """RMS of the fit from the accumulated Gram of [Phi | V | 1], without a second basis pass.

        sum((Phi q - c) - (V - vbar))^2 = q'ATA q - 2 q'ATb + V'V - 2 t (sum_phi'q - sum_V) + n t^2,
        t = c - vbar, c = the reference's average Coulomb potential (``corrected`` uses the true sum_phi).
        """
        ata, atb, sum_phi, sum_v, vtv, n = self._fit_gram_extras()
        q = np.asarray(charges, dtype=np.float64)
        sp = sum_phi if corrected or self._sum_phi is None else self._sum_phi
        c = float(sp @ q) / n
        t = c - sum_v / n
        ss = q @ ata @ q - 2 * (q @ atb) + vtv - 2 * t * (sum_phi @ q - sum_v) + n * t * t
        return float(np.sqrt(max(ss, 0.0) / n))

    def _fit_gram_extras(self):
        ata, atb, sum_phi, sum_v = self._fit.normal_equations()
        v = self.V_pot
        return ata, atb, sum_phi, sum_v, float(v @ v), self.n_grid_points

    # ---- QEq charge model (repeat.py:2538-2634)
    def qeq_parameters(self, param_set=None):
        """Per-atom (electronegativity, idempotential J_ii = 2 * hardness) as compute_qeq_charges
        resolves them (repeat.py:2540-2574): without ``param_set`` the method is QEq on the default
        table, with one it is EQeq (hydrogen then takes chi = (13.598 - 2.0)/2 eV, J = 15.598 eV);
        --qeq-file entries override the table; unknown elements take (0.15, 0.50) with a warning."""
        if param_set is None:
            param_set, method = DEFAULT_QEQ_PARAMS, "QEq"
        else:
            method = "EQeq"
        h_i0, h_i1 = -2.0, 13.598
        ev_to_ha = 1.0 / 27.211386245988
        chi = np.zeros(self.n_atoms)
        jii = np.zeros(self.n_atoms)
        for i, z in enumerate(self.cube["atom_index"]):
            z = int(z)
            sym = ATOM_SYMBOLS[z] if z < len(ATOM_SYMBOLS) else "X"
            if sym == "H" and method == "EQeq":
                chi[i] = 0.5 * (h_i1 + h_i0) * ev_to_ha
                jii[i] = (h_i1 - h_i0) * ev_to_ha
            elif self.qeq_data and sym in self.qeq_data:
                chi[i] = self.qeq_data[sym]["electroneg"]
                jii[i] = 2.0 * self.qeq_data[sym]["hardness"]
            elif sym in param_set:
                c, h = param_set[sym]
                chi[i] = c
                jii[i] = 2.0 * h
            else:
                chi[i] = 0.15
                jii[i] = 0.50
                print(f"Warning: No parameters for {sym}, using defaults")
        return chi, jii, method

    def compute_qeq_charges(self, param_set=None, return_matrix=False):
        """QEq charges on the device (the Coulomb matrix and the N x N solve both run there)."""
        chi, jii, method = self.qeq_parameters(param_set)
        print(f"Computing {method} Coulomb matrix...")
        t0 = time.time()
        out = self._fit.qeq(chi, jii, self.q_tot, want_j=return_matrix)
        print(f"  Coulomb matrix and solve in {time.time() - t0:.2f} seconds")
        return out

    def timings(self):
        return self._fit.timings()

    def close(self):
        self._fit.close()


def _expand_sum_phi_quirk(sum_phi_given, symm, n_atoms):
    trees = symm["trees"]
    linked = set()
    for t in trees:
        linked.update(t["linked"])
    independent = [i for i in range(n_atoms) if i not in linked]
    out = np.zeros(n_atoms)
    for i, idx in enumerate(independent):
        out[idx] = sum_phi_given[i]
    for t in trees:
        for ln in t["linked"]:
            out[ln] = out[t["base"]]
    return out
