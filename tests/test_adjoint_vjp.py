"""Host check of the hand-written vector-Jacobian products (voxelyze_amd/csrc/vx_adjoint.h)
against central finite differences: every primitive (quaternion product, rotations,
rotation-vector conversions, FromAngleToPosX), one link in each axis and branch (small or
large angle, with or without damping) and one voxel in each floor/friction state.
Compiled from tests/cpp/adjoint_vjp.cpp with hipcc as a host program (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_vjps_match_finite_differences(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "adjoint_vjp"
    src = os.path.join(ROOT, "tests", "cpp", "adjoint_vjp.cpp")
    r = subprocess.run([hipcc, "-O1", "-std=c++17", "-ffp-contract=off", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "0 failures" in r.stdout, r.stdout[-3000:]
