"""Command line of the REPEAT fit, flag-compatible with the reference ``repeat.py`` (``repeat.py:2636-2789``).

    python -m pyrepeat_amd system.cube [--fit-type 1] [--vdw-factor 1.0] [--vdw-max 1000]
        [--cutoff 20] [--total-charge 0] [--symm-file F] [--resp-file F] [--qeq-file F]
        [--stats] [--output F] [--fp64] [--precision fp64|split|refined|fp64-factorized] [--charge repeat|qeq]

Every fit runs on the GPU through ``librepeat_hip.so``.  As with the reference's
``--gpu``, the default precision is a split (mixed) form: ``refined`` (int8-sliced
Gram, then fp64 iterative refinement of the solve, so the charges are those of
the fp64 normal equations); ``--fp64`` selects the plain fp64 path and
``--precision split`` the split form without refinement.  ``--gpu``, ``--cores``, ``--block-k`` and
``--no-jit`` are accepted for compatibility and do not change the result.
The charge table, ``Total charge`` and ``RMS error`` lines and the
``--output`` rows use the reference's formats (``repeat.py:2775-2789``).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from .tables import element_symbol


def build_parser():
    p = argparse.ArgumentParser(description="REPEAT charges on AMD MI355X (HIP)")
    p.add_argument("cube_file", nargs="?", default=None)
    p.add_argument("--fit-type", type=int, default=1, choices=[0, 1])
    p.add_argument("--vdw-factor", type=float, default=1.0)
    p.add_argument("--vdw-max", type=float, default=1000.0)
    p.add_argument("--cutoff", type=float, default=20.0)
    p.add_argument("--total-charge", type=float, default=0.0)
    p.add_argument("--symm-file", type=str, default=None)
    p.add_argument("--resp-file", type=str, default=None)
    p.add_argument("--qeq-file", type=str, default=None)
    p.add_argument("--cores", type=int, default=None)
    p.add_argument("--block-k", type=int, default=64)
    p.add_argument("--no-jit", action="store_true")
    p.add_argument("--gpu", nargs="?", const="0", default=None)
    p.add_argument("--fp64", action="store_true", help="plain fp64 path")
    p.add_argument("--precision", choices=["fp64", "split", "refined", "fp64-factorized"], default=None,
                   help="explicit precision form (overrides --fp64)")
    p.add_argument("--output", type=str, default=None)
    p.add_argument("--charge", type=str, default="repeat", choices=["repeat", "qeq"])
    p.add_argument("--stats", action="store_true", default=False)
    return p


def main(argv=None):
    args = build_parser().parse_args(argv)
    if (not args.cube_file) or (not os.path.exists(args.cube_file)):
        print(f"Error: Cube file '{args.cube_file}' not found!")
        return 1
    from .calculator import ChargeCalculator
    precision = args.precision or ("fp64" if args.fp64 else "refined")
    device = 0
    if args.gpu not in (None, "", "all"):
        device = int(str(args.gpu).split(",")[0])
    from . import _abi
    ctx = _abi.Context(device, 0)
    t0 = time.time()
    calc = ChargeCalculator(cube_file=args.cube_file, fit_flag=args.fit_type, vdw_factor=args.vdw_factor,
                            vdw_max=args.vdw_max, R_cutoff=args.cutoff, q_tot=args.total_charge,
                            symm_file=args.symm_file, resp_file=args.resp_file, qeq_file=args.qeq_file,
                            precision=precision, context=ctx)
    print(f"Grid Points: {int(calc._fit.info.n_total)}, Effective Atoms: {int(calc._fit.info.n_eff)}")
    print(f"  Kept {calc.n_grid_points} grid points; read + filter {time.time() - t0:.2f} s ({precision})")
    if args.charge == "qeq":
        # repeat.py:2735-2756
        print("\n========================================")
        print("Using QEq (Charge Equilibration) method")
        print("========================================")
        charges = calc.compute_qeq_charges()
        print("\nQEq Charges:")
        for i, q in enumerate(charges):
            print(f"{i + 1:4d} {element_symbol(calc.cube['atom_index'][i]):2s} {q:12.6f}")
        print(f"\nTotal charge: {np.sum(charges):.6f}")
        if args.output:
            with open(args.output, "w") as f:
                for i, q in enumerate(charges):
                    pos = calc.cube["atom_pos"][i]
                    f.write(f"{i + 1:4d} {element_symbol(calc.cube['atom_index'][i]):2s} {pos[0]:12.6f} {pos[1]:12.6f} "
                            f"{pos[2]:12.6f} {q:12.6f}\n")
            print(f"QEq charges written to {args.output}")
        calc.close()
        ctx.close()
        return 0
    charges = calc.fit_charges()
    stats = None
    if args.stats or args.output:
        stats = calc.compute_stats(charges, need_v_coul=args.output is not None)
    print("\nREPEAT Charges:")
    for i, q in enumerate(charges):
        print(f"{i + 1:4d} {element_symbol(calc.cube['atom_index'][i]):2s} {q:12.6f}")
    print(f"\nTotal charge: {np.sum(charges):.6f}")
    if stats is not None:
        print(f"RMS error: {stats['rms_error']:.6e}")
    if args.output:
        with open(args.output, "w") as f:
            for gp, vqm, vc in zip(calc.grid_pos, calc.V_pot, stats["V_coul"]):
                f.write(f"{gp[0]:12.6f} {gp[1]:12.6f} {gp[2]:12.6f} {vc:16.8e} {vqm:16.8e}\n")
        print(f"Output written to {args.output}")
    t = calc.timings()
    print(f"Device time: filter {t['filter_ms']:.3f} ms, accumulate {t['accumulate_ms']:.3f} ms, "
          f"solve {t['solve_ms']:.3f} ms")
    calc.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
