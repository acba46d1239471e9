"""The static linear solve on the device (vxg_linear_solve; CVoxelyze::doLinearSolve,
src/Voxelyze.cpp:243-248) against the reference's own linear system.

The reference forms the system and hands it to Pardiso; the oracle forms the same system bit
for bit (tests/test_linear.py) and a sparse direct solve stands in for Pardiso.  The device
solves it by preconditioned conjugate gradients without storing it.  Bar: positions within
1e-8 of the largest displacement, rotation vectors within 1e-8 of the largest rotation, at
the default relative residual 1e-10; momenta zero afterwards (postResults,
src/VX_LinearSolver.cpp:316-326).  A solve that cannot succeed leaves the state unchanged.
"""
import numpy as np
import pytest

import backends as B
import linear_models as LM

L = 0.005

CASES = {
    "line_tip_force": lambda s: LM.line_cantilever(s, n=10),
    "line_tip_moment": lambda s: LM.line_cantilever(s, n=7, force=(0.0, 2e-4, 0.0), moment=(1e-6, 0.0, 3e-7)),
    "block_cantilever": LM.block_cantilever,
    "random_0": lambda s: LM.random_lattice(s, 0),
    "random_1_stepped": lambda s: LM.random_lattice(s, 1, steps=100),
    "random_3_large": lambda s: LM.random_lattice(s, 3, nx=24, ny=20, nz=12, steps=20),
}


def rotvec(q):
    """Quat3D::ToRotationVector for small angles (include/Quat3D.h:117-122), numpy"""
    w = np.clip(q[:, 0], -1.0, 1.0)
    ang = 2 * np.arccos(w)
    s = np.sqrt(np.maximum(1 - w * w, 1e-300))
    return q[:, 1:] * (ang / s)[:, None]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_solve_matches_the_reference_system(name):
    print(f"[linear] {name}", flush=True)
    ref = B.maker(B.oracle_cr_abi())(L)
    CASES[name](ref)
    system = LM.linear_system(B.ORACLE_CR_SO, "vxo_", ref)
    u = LM.direct_solve(system)
    nominal = ref.state()["pos"] - system["x"].reshape(-1, 6)[:, :3]  # x holds the current displacements
    dev = B.maker(B.product_abi())(L)
    CASES[name](dev)
    assert dev.doLinearSolve(), dev.lastError()
    info = dev.linearSolveInfo
    st = dev.state()
    disp = st["pos"] - nominal
    e_pos = np.abs(disp - u[:, :3]).max() / np.abs(u[:, :3]).max()
    rv = rotvec(st["orient"])
    e_rot = np.abs(rv - u[:, 3:]).max() / max(np.abs(u[:, 3:]).max(), 1e-300)
    print(f"[linear] {name}: {dev.voxelCount()} voxels, {info['iterations']} iterations, residual "
          f"{info['relative_residual']:.2e}, position err {e_pos:.2e}, rotation err {e_rot:.2e}", flush=True)
    assert info["relative_residual"] <= 1e-10
    assert e_pos <= 1e-8 and e_rot <= 1e-8
    assert not st["linMom"].any() and not st["angMom"].any()
    # the state the solve leaves is a state the step continues from
    n, at = dev.step(dev.recommendedTimeStep(), 10)
    assert n == 10 and at < 0


@pytest.mark.gpu
def test_failed_solves_leave_the_state_unchanged():
    dev = B.maker(B.product_abi())(L)
    dev.addMaterial(1e6, 1000.0)
    assert not dev.doLinearSolve() and "no voxels" in dev.lastError()

    def unchanged(sim, call, why):
        before = sim.state()
        assert not call(), why
        assert why in sim.lastError(), sim.lastError()
        after = sim.state()
        assert all(before[k].tobytes() == after[k].tobytes() for k in before)

    # a body with no fixed degree of freedom under a load: no static solution
    free = B.maker(B.product_abi())(L)
    vs = LM.line_cantilever(free, n=4)
    vs[0].external().setFixedAll(False)
    free.step(free.recommendedTimeStep(), 5)
    unchanged(free, lambda: free.doLinearSolve(max_iterations=200), "did not converge")
    # a voxel with free degrees of freedom and no links
    lone = B.maker(B.product_abi())(L)
    LM.line_cantilever(lone, n=3)
    lone.setVoxel(lone.material(0), 10, 10, 10)
    unchanged(lone, lone.doLinearSolve, "singular")
    # per-voxel moduli are not part of the linear model
    mod = B.maker(B.product_abi())(L)
    LM.line_cantilever(mod, n=3)
    mod.setModuli(np.full(mod.voxelCount(), 2e6, np.float32))
    unchanged(mod, mod.doLinearSolve, "per-voxel moduli")
