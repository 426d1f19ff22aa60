"""Build ``librlic_hip.so`` in-tree with hipcc for gfx950.

Run as ``python -m rlic_amd._build``.  Compile flags that matter for the
results:
  -ffp-contract=off   no implicit fusion of a*b+c (the reference's Rust never
                      contracts; fused multiply-adds are written explicitly)
  packed-fp32-ops off keep scalar fp32 arithmetic (no v_pk_* fp32 forms)
hipcc's fp32 division is correctly rounded by default and fp32 denormals are
preserved on gfx950, both of which bit-exactness relies on.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
SOURCES = [CSRC / "lic_kernels.hip", CSRC / "lic_scratch.hip", CSRC / "lic_batch.hip",
           CSRC / "lic_serve.hip", CSRC / "lic_generate.hip", CSRC / "lic_equalize.hip",
           CSRC / "lic_api.hip"]
HEADERS = [CSRC / "lic_device.h", CSRC / "lic_launch.h", CSRC / "lic_steps.h", CSRC / "lic_batch.h",
           CSRC / "lic_serve.h", CSRC / "lic_scratch.h", CSRC / "lic_equalize.h",
           CSRC / "lic_streams.h",
           ROOT / "include" / "rlic_hip.h"]
OUTPUT = PKG / "librlic_hip.so"
ARCH = os.environ.get("RLIC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def command(output: Path = OUTPUT) -> list[str]:
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-ffp-contract=off",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-o", str(output),
        *map(str, SOURCES),
    ]


def build(verbose: bool = False) -> Path:
    """Compile the library (always recompiles) and return its path."""
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            "hipcc failed:\n" + " ".join(cmd) + "\n" + proc.stdout + proc.stderr
        )
    os.replace(tmp, OUTPUT)
    if verbose:
        print(" ".join(cmd))
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
