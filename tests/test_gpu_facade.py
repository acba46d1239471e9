"""The C++ drop-in header: a reference-style program (tests/cpp/facade_cantilever.cpp)
compiled against include/Voxelyze.h runs on the GPU and reproduces the oracle; its
doLinearSolve() matches a direct solve of the reference's linear system."""
import os
import subprocess

import numpy as np
import pytest

import backends as B
import linear_models as LM
from voxelyze_amd import workloads as W

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_cantilever")


@pytest.mark.gpu
def test_facade_program_matches_oracle():
    steps = 300
    assert os.path.exists(EXE), "build it with __graft_entry__.build()"
    r = subprocess.run([EXE, str(steps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    head = lines[0].split()
    rows = [ln.split() for ln in lines[1:]]
    pos = np.array([[float.fromhex(v) for v in row[1:4]] for row in rows])
    quat = np.array([[float.fromhex(v) for v in row[4:8]] for row in rows])
    ref = B.maker(B.oracle_abi())(0.005)
    W.cantilever(ref)
    for _ in range(steps):
        ref.doTimeStep()
    st = ref.state()
    assert np.array_equal(pos, st["pos"]) and np.array_equal(quat, st["orient"])
    assert int(head[1]) == ref.linkCount()
    assert float.fromhex(head[3]) == ref.link(0, 0, 0, 0).axialStrain()


@pytest.mark.gpu
def test_facade_linear_solve_matches_the_reference_system():
    assert os.path.exists(EXE), "build it with __graft_entry__.build()"
    r = subprocess.run([EXE, "0", "linear"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [ln.split() for ln in r.stdout.strip().splitlines()[1:]]
    pos = np.array([[float.fromhex(v) for v in row[1:4]] for row in rows])
    ref = B.maker(B.oracle_abi())(0.005)
    W.cantilever(ref)
    u = LM.direct_solve(LM.linear_system(B.ORACLE_SO, "vxo_", ref))
    disp = pos - ref.state()["pos"]  # the oracle's voxels are still at their nominal positions
    err = np.abs(disp - u[:, :3]).max() / np.abs(u[:, :3]).max()
    print(f"[facade] doLinearSolve: position err {err:.2e}")
    assert err <= 1e-8
