"""pyrepeat_amd: pyREPEAT's REPEAT charge-fit path on AMD MI355X (gfx950) HIP kernels.

Host interface mirrors the reference (``ChargeCalculator``, ``fit_charges``,
``compute_stats``); compute runs in ``librepeat_hip.so`` (C ABI in
``include/repeat_hip.h``).
"""
from .calculator import ChargeCalculator  # noqa: F401

__all__ = ["ChargeCalculator"]
