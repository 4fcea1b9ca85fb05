"""The force records' voxel decomposition and the voxel index of the binning kernels
use two shortcuts (device_math.hpp): a reciprocal product with the division kept near
integer quotients (axis_index_fast) and a reciprocal quotient with one remainder
correction (udiv_by_inv).  Compiled for the host and checked against the exact
forms on 2e6 random and face-adjacent inputs each (the voxel index must be
bit-exact: cartesian_cuboid_n.rs:469-491).  CPU only."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_index_helpers_exact(tmp_path):
    exe = str(tmp_path / "index_helpers_check")
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "cellular_raza_amd", "csrc"),
           "-x", "hip", os.path.join(HERE, "cpp", "index_helpers_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches; udiv_by_inv: 0 mismatches" in r.stdout
