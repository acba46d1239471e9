"""Mixed precision (VXG_MIXED) against the FP64 step of the same library.

VXG_MIXED evaluates link forces, damping and contacts in FP32 and keeps the
voxel state, force sums, momentum and quaternion integration in FP64.  Its bar
is a relative displacement error of at most 1e-4 against the FP64 step, which
itself reproduces the reference's doTimeStep (tests/test_gpu_parity.py):

    err = max_i |pos_mixed_i - pos_fp64_i| / max_i |pos_fp64_i - nominal_i|

Every test prints the error it measured.
"""
import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W

FP64, MIXED = 0, 1
TOL = 1e-4


def make(build, precision, voxel_size=0.005):
    sim = B.maker(B.product_abi())(voxel_size)
    info = build(sim)
    sim.setPrecision(precision)
    return sim, info


def nominal(sim):
    ijk = np.array([(v.indexX(), v.indexY(), v.indexZ()) for v in map(sim.voxel, range(sim.voxelCount()))],
                   dtype=np.float64)
    return ijk * sim.voxelSize()


def rel_err(ref, out, nom):
    disp = np.abs(ref - nom).max()
    return np.abs(out - ref).max() / max(disp, 1e-300)


def run_pair(name, build, steps, dt_scale=1.0, voxel_size=0.005, drive=None, chunk=None):
    a, info = make(build, FP64, voxel_size)
    b, _ = make(build, MIXED, voxel_size)
    dt = a.recommendedTimeStep() * dt_scale
    assert dt == b.recommendedTimeStep() * dt_scale
    chunk = chunk or steps
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        if drive:
            drive(a, info, done)
            drive(b, info, done)
        na, at_a = a.step(dt, k)
        nb, at_b = b.step(dt, k)
        assert (na, at_a) == (k, -1) and (nb, at_b) == (k, -1), (na, at_a, nb, at_b)
        done += k
    pa, pb = a.state()["pos"], b.state()["pos"]
    err = rel_err(pa, pb, nominal(a))
    print(f"[mixed] {name}: {steps} steps, {a.voxelCount()} voxels, rel disp err {err:.3e}", flush=True)
    assert np.isfinite(pb).all()
    return a, b, err


@pytest.mark.gpu
def test_mixed_cantilever_320():
    """The 20x4x4 cantilever (5 mm pitch, E=1 MPa, -1 N tip load), 20,000 steps."""
    a, b, err = run_pair("cantilever 20x4x4", W.cantilever, 20000, chunk=5000)
    assert err <= TOL
    la, lb = a.link_state(), b.link_state()
    scale = np.abs(la["strain"]).max()
    assert np.abs(la["strain"] - lb["strain"]).max() <= TOL * scale


@pytest.mark.gpu
def test_mixed_drop_lattice_with_floor_friction():
    """32x32x16 Bernoulli(0.9) lattice of two materials dropped on the floor, 3,000 steps."""
    build = lambda s: W.drop_lattice(s, nx=32, ny=32, nz=16)
    _, _, err = run_pair("drop lattice 32x32x16", build, 3000, chunk=1000)
    assert err <= TOL


def porous(sim):
    return W.porous_block(sim, n=16, rmin=1, rmax=3)


def porous_drive(sim, info, step):
    W.porous_block_drive(sim, info["top"], step, 0.005, rate=2.0 / 1000)


@pytest.mark.gpu
def test_mixed_porous_block_contacts_and_plasticity():
    """16^3 porous bilinear block, top face driven down 2 pitches, self-collision on, 1,000 steps:
    links yield and surfaces touch while the FP64 solution is still well conditioned (a 1e-9 N
    interior force moves it by ~5e-9 relative)."""
    a, b, err = run_pair("porous block 16^3", porous, 1000, drive=porous_drive, chunk=100)
    la, lb = a.link_state(), b.link_state()
    ya, yb = ((la["flags"] & 4) != 0).sum(), ((lb["flags"] & 4) != 0).sum()
    print(f"[mixed] porous block: contacts {a.collisionCount()} / {b.collisionCount()}, "
          f"yielded links {ya} / {yb} (fp64 / mixed)")
    assert a.collisionCount() > 0 and ya > 0
    assert err <= TOL


@pytest.mark.gpu
def test_mixed_error_tracks_fp64_sensitivity_when_chaotic():
    """Past ~1,500 steps the same block buckles, yields widely and links fail: the FP64 result then
    moves by percents under a 1e-9 N perturbation of one interior voxel, so no precision can
    be held to 1e-4.  The bar there: the mixed/FP64 difference stays within 10x the FP64
    path's own response to that perturbation."""
    steps = 2000
    a, b, err = run_pair("porous block 16^3 (chaotic)", porous, steps, drive=porous_drive, chunk=100)
    c, info = make(porous, FP64)
    z = c.state()["pos"][:, 2]
    c.voxel(int(np.argmin(np.abs(z - z.max() / 2)))).external().setForce(1e-9, 0.0, 0.0)
    dt = c.recommendedTimeStep()
    for done in range(0, steps, 100):
        porous_drive(c, info, done)
        assert c.step(dt, 100) == (100, -1)
    sens = rel_err(a.state()["pos"], c.state()["pos"], nominal(a))
    print(f"[mixed] chaotic porous block: mixed err {err:.3e}, fp64 response to 1e-9 N {sens:.3e}")
    assert sens > 1e-4  # the regime is chaotic
    assert err <= 10 * sens


@pytest.mark.gpu
def test_mixed_divergence_matches_fp64():
    """A strain above 100 stops a mixed launch at the same step as the FP64 one."""
    def build(sim):
        m = sim.addMaterial(1e6, 1e3)
        sim.setVoxel(m, 0, 0, 0).external().setFixedAll()
        sim.setVoxel(m, 1, 0, 0).external().setForce(10.0, 0, 0)

    a, _ = make(build, FP64, 0.001)
    b, _ = make(build, MIXED, 0.001)
    dt = a.recommendedTimeStep() * 40
    ra, rb = a.step(dt, 40), b.step(dt, 40)
    print(f"[mixed] divergence: fp64 {ra}, mixed {rb}")
    assert ra == rb and ra[1] >= 0
    pa, pb = a.state()["pos"], b.state()["pos"]
    assert np.allclose(pa, pb, rtol=1e-4, atol=0)
    assert a.currentTime() == b.currentTime()


@pytest.mark.gpu
def test_mixed_divergence_in_single_step_launches():
    """doTimeStep until it returns false: the diverging step is the last of its launch, and
    the voxels and the clock stay as before it, as in the FP64 step."""
    def build(sim):
        m = sim.addMaterial(1e6, 1e3)
        sim.setVoxel(m, 0, 0, 0).external().setFixedAll()
        sim.setVoxel(m, 1, 0, 0).external().setForce(10.0, 0, 0)

    a, _ = make(build, FP64, 0.001)
    b, _ = make(build, MIXED, 0.001)
    dt = a.recommendedTimeStep() * 40
    steps = []
    for sim in (a, b):
        k = 0
        while k < 100 and sim.doTimeStep(dt):
            k += 1
        steps.append(k)
    print(f"[mixed] single-step divergence after {steps}")
    assert steps[0] == steps[1] < 100
    assert np.allclose(a.state()["pos"], b.state()["pos"], rtol=1e-4, atol=0)
    assert a.currentTime() == b.currentTime()


@pytest.mark.gpu
def test_precision_switch_mid_run():
    """FP64 -> mixed -> FP64 keeps the state: the result stays within the mixed bar of an FP64-only run."""
    a, _ = make(W.cantilever, FP64)
    b, _ = make(W.cantilever, FP64)
    dt = a.recommendedTimeStep()
    a.step(dt, 3000)
    b.step(dt, 1000)
    b.setPrecision(MIXED)
    b.step(dt, 1000)
    b.setPrecision(FP64)
    b.step(dt, 1000)
    err = rel_err(a.state()["pos"], b.state()["pos"], nominal(a))
    print(f"[mixed] fp64/mixed/fp64 switch: rel disp err {err:.3e}")
    assert err <= TOL
    assert a.currentTime() == b.currentTime()


@pytest.mark.gpu
def test_mixed_single_steps_equal_one_launch():
    """doTimeStep x n and vxg_step(n) give identical mixed results."""
    a, _ = make(W.cantilever, MIXED)
    b, _ = make(W.cantilever, MIXED)
    dt = a.recommendedTimeStep()
    a.step(dt, 300)
    for _ in range(300):
        assert b.doTimeStep(dt)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert sa[k].tobytes() == sb[k].tobytes(), k
