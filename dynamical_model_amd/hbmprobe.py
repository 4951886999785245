"""Practical HBM streaming ceilings of one MI355X for the kernels' access
shapes (see probe/hbm_probe.hip): builds the probe with hipcc for gfx950 and
runs it.

    python3 -m dynamical_model_amd.hbmprobe [--build-only]
"""
from __future__ import annotations

import os
import subprocess
import sys

from .build import ARCH, _hipcc

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "probe", "hbm_probe.hip")
EXE = os.path.join(PKG, "probe", "hbm_probe")


def build() -> str:
    cmd = [_hipcc(), "-O3", f"--offload-arch={ARCH}", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", EXE, SRC]
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL)
    return EXE


def main() -> int:
    build()
    if "--build-only" in sys.argv:
        return 0
    return subprocess.run([EXE]).returncode


if __name__ == "__main__":
    sys.exit(main())
