"""Builds the gfx950 library in-tree: cellular_raza_amd/libcrmech.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [os.path.join(HERE, "csrc", "crmech.hip")]
OUTPUT = os.path.join(HERE, "libcrmech.so")

FLAGS = [
    "-std=c++17",
    "-O3",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    # FP64 state arithmetic keeps the reference's separate roundings
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(HERE, "csrc"),
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + ["-o", OUTPUT + ".tmp"] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with status {r.returncode}")
    os.replace(OUTPUT + ".tmp", OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
