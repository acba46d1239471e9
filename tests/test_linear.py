"""The static linear solve's system (CVX_LinearSolver, src/VX_LinearSolver.cpp:127-314), CPU.

* The oracle's restatement forms the reference's own system bit for bit: the upper-triangle
  CSR arrays, b and x, read out of the compiled reference (oracle/ref_driver.cpp), on line
  and block cantilevers and on shuffled random lattices with full, partial and prescribed
  fixes, forces and moments, before and after dynamic steps.
* The system is the Euler-Bernoulli beam: a line of n voxels clamped at one end deflects
  under a tip load by F((n-1)L)^3 / 3EI and turns by F((n-1)L)^2 / 2EI (I = L^4/12), and
  stretches by F(n-1)L / EA.
* The oracle's dynamic step (itself bit-exact against the reference) settles, under the
  same small load, to the static solution.
"""
import numpy as np
import pytest

import backends as B
import linear_models as LM

L, E = 0.005, 1e6


def oracle_system(build):
    sim = B.maker(B.oracle_abi())(L)
    build(sim)
    return sim, LM.linear_system(B.ORACLE_SO, "vxo_", sim)


MODELS = {
    "line_tip_force": lambda s: LM.line_cantilever(s, n=10),
    "line_tip_moment": lambda s: LM.line_cantilever(s, n=7, force=(0.0, 2e-4, 0.0), moment=(1e-6, 0.0, 3e-7)),
    "block_cantilever": LM.block_cantilever,
    "random_0": lambda s: LM.random_lattice(s, 0),
    "random_1_stepped": lambda s: LM.random_lattice(s, 1, steps=100),
    "random_2_stepped": lambda s: LM.random_lattice(s, 2, nx=10, ny=4, nz=6, steps=300),
}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_oracle_system_equals_the_reference(name):
    ref = B.ref_abi()
    if ref is None:
        pytest.skip("compiled reference not present (built only where the reference sources are)")
    rs = B.maker(ref)(L)
    MODELS[name](rs)
    want = LM.linear_system(B.REF_SO, "vxr_", rs)
    _, got = oracle_system(MODELS[name])
    for k in want:
        assert want[k].shape == got[k].shape, k
        assert want[k].tobytes() == got[k].tobytes(), f"{name}: {k} differs"


def test_line_cantilever_is_the_euler_bernoulli_beam():
    """Cubic beam elements are exact at the nodes for end loads; the reference's beam
    constants are FP32, and their rounding, amplified by the cancellation b1 - b2^2/2b3 at
    every node, leaves ~1.4e-7 (n-1)^2 relative (2.2e-6 for n = 5, 1.1e-5 for n = 10)."""
    n, F = 5, -1e-3
    _, sys_ = oracle_system(lambda s: LM.line_cantilever(s, n=n, force=(0.0, 0.0, F)))
    u = LM.direct_solve(sys_)
    ell = (n - 1) * L
    I = L**4 / 12
    assert u[-1, 2] == pytest.approx(F * ell**3 / (3 * E * I), rel=1e-5)
    assert u[-1, 4] == pytest.approx(-F * ell**2 / (2 * E * I), rel=1e-5)
    # every station of the beam: w(s) = F s^2 (3 ell - s) / 6EI
    s = np.arange(n) * L
    assert np.allclose(u[:, 2], F * s**2 * (3 * ell - s) / (6 * E * I), rtol=1e-5, atol=0)
    _, sys_ = oracle_system(lambda s: LM.line_cantilever(s, n=n, force=(F, 0.0, 0.0)))
    u = LM.direct_solve(sys_)
    assert u[-1, 0] == pytest.approx(F * ell / (E * L * L), rel=2e-6)


def test_dynamic_step_settles_to_the_static_solution():
    """Six voxels clamped at one end, 1e-6 N at the tip (deflection ~1e-4 of the pitch, so the
    geometric nonlinearity is below the tolerance): the dynamic step, run until it rests,
    lands on the linear solution."""
    build = lambda s: LM.line_cantilever(s, n=6, force=(0.0, 3e-7, -1e-6))
    sim, sys_ = oracle_system(build)
    u = LM.direct_solve(sys_)
    m = sim.material(0)
    m.setGlobalDamping(0.05)
    dt = sim.recommendedTimeStep()
    prev = None
    for _ in range(40):
        assert sim.step(dt, 500) == (500, -1)
        pos = sim.state()["pos"]
        if prev is not None and np.abs(pos - prev).max() < 1e-6 * np.abs(u[:, :3]).max():
            break
        prev = pos
    nominal = np.array([[i * L, 0.0, 0.0] for i in range(6)])
    disp = pos - nominal
    err = np.abs(disp - u[:, :3]).max() / np.abs(u[:, :3]).max()
    print(f"[linear] dynamic rest vs static solution: rel err {err:.2e}")
    assert err < 1e-3
