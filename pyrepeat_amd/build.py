"""Build the native parts in-tree: ``librepeat_hip.so`` (hipcc, gfx950) and ``libsynth_esp.so`` (gcc)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_hip(jobs: int = 8, force: bool = True) -> str:
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.check_call(["make", "-s", "-C", csrc, "clean"])
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", csrc])
    return os.path.join(HERE, "librepeat_hip.so")


def build_all(force: bool = True) -> None:
    build_hip(force=force)
    from . import synth
    synth.build_helper(force=force)


if __name__ == "__main__":
    build_all()
