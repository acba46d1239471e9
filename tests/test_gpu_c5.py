"""The differentiable batch at its stated size: 256 robots, each the largest connected
component of a seeded 70 % random fill of a 32x32x16 box (2.93M voxels), three linear
materials (passive 1 MPa; actuated 1 MPa with CTE +-0.01 per degC), gravity, floor friction,
+-39 degC at 2 Hz, 2,000 mixed-precision steps with checkpoints every 100.  Loss: minus the
summed centre-of-mass x-travel of the robots; gradients with respect to per-voxel log E and
actuation phase.

Checks:
  * the mixed step's adjoint against the FP64 fused step's adjoint (the FP64 reference of
    the mixed step);
  * central differences along 16 seeded directions for the mixed step, and the first 8 of them
    for the FP64 fused step.  With floor friction the loss is piecewise smooth (voxels switch between static and
    kinetic friction on force comparisons), so a difference quotient over h = 1e-3 does not
    converge to the derivative; FP64's own discrepancy is the yardstick the mixed step is held
    to;
  * in a smooth setting at the same horizon on 32 of the robots -- no floor and no gravity (no
    contact or friction switches) and 5 degC actuation (at +-39 degC and CTE 0.01 the links
    stretch by up to 39 % and switch between the small- and large-angle forms of
    CVX_Link::orientLink, another piecewise definition) -- a weighted-position loss: Richardson-extrapolated central differences
    within 1e-6 (FP64 fused) and 1e-5 (mixed) of the adjoint.  The mixed step forms its beam
    constants in FP32 from the FP32 moduli (the reference's types), which leaves ~1e-7
    relative noise on the loss, so its difference steps are 4e-2 and 2e-2 (noise/h below
    1e-5); the FP64 fused step uses 1e-2 and 5e-3.
"""
import math

import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W

STEPS, EVERY, ROBOTS = 2000, 100, 256


def build(precision, robots=ROBOTS, floor=True):
    sim = B.maker(B.product_abi())(0.005)
    W.differentiable_batch(sim, robots)
    if not floor:  # the smooth setting: no contact, and strains small enough to keep every link small-angle
        sim.enableFloor(False)
        sim.setGravity(0.0)
        sim.setActuation(True, 0.0, 5.0, 2.0)
    sim.setPrecision(precision)
    return sim


def params(n, seed=9):
    u = W.uniform(seed, np.arange(2 * n, dtype=np.uint64))
    return np.log(1e6) + 0.2 * (u[:n] - 0.5), 2 * math.pi * u[n:]


def com_weights(sim):
    dl = np.zeros((sim.voxelCount(), 3))
    for r in range(sim.robotCount()):
        f, c = sim.robotVoxels(r)
        dl[f:f + c, 0] = -1.0 / c  # equal voxel masses: the centre of mass is the mean
    return dl


def rollout(sim, E, phi, w, nsteps=STEPS, grad=True):
    """loss sum(w * (final - initial positions)) with the FP32 moduli and phases as uploaded"""
    sim.resetTime()
    sim.setModuli(E)
    sim.setPhases(phi)
    pos0 = sim.state()["pos"]
    sim.adjointRun(nsteps, EVERY)
    loss = float((w * (sim.state()["pos"] - pos0)).sum())
    if not grad:
        return loss, None
    return loss, sim.adjointBackward(w)


@pytest.fixture(scope="module")
def c5():
    out = {}
    for prec in (1, 2):
        print(f"[c5] precision {prec}: 256 robots, 2000 steps", flush=True)
        sim = build(prec)
        theta, phi = params(sim.voxelCount())
        w = com_weights(sim)
        E, ph = np.exp(theta).astype(np.float32), phi.astype(np.float32)
        loss, (ge, gp) = rollout(sim, E, ph, w)
        sim.adjointFree()
        out[prec] = dict(sim=sim, theta=theta, phi=phi, w=w, loss=loss, ge=ge, gp=gp)
        print(f"[c5] precision {prec}: {sim.voxelCount()} voxels, loss {loss:.9e}, max |dL/dlogE| "
              f"{np.abs(ge).max():.3e}, max |dL/dphase| {np.abs(gp).max():.3e}", flush=True)
    return out


@pytest.mark.gpu
def test_c5_gradients_are_finite_and_mixed_matches_fp64(c5):
    m, f = c5[1], c5[2]
    for g in (m["ge"], m["gp"], f["ge"], f["gp"]):
        assert np.isfinite(g).all()
    et = np.abs(m["ge"] - f["ge"]).max() / np.abs(f["ge"]).max()
    ep = np.abs(m["gp"] - f["gp"]).max() / np.abs(f["gp"]).max()
    el = abs(m["loss"] - f["loss"]) / abs(f["loss"])
    print(f"[c5] mixed vs FP64: loss {el:.2e}, gradient dlogE {et:.2e}, dphase {ep:.2e}", flush=True)
    assert el <= 1e-3 and et <= 5e-2 and ep <= 5e-2


def fd_discrepancies(d, ndir=16, h=1e-3):
    sim, theta, phi, w, ge, gp = d["sim"], d["theta"], d["phi"], d["w"], d["ge"], d["gp"]
    n = len(theta)
    out = []
    for k in range(ndir):
        u = W.uniform(100 + k, np.arange(2 * n, dtype=np.uint64)) * 2 - 1
        Ep, Em = np.exp(theta + h * u[:n]).astype(np.float32), np.exp(theta - h * u[:n]).astype(np.float32)
        pp, pm = (phi + h * u[n:]).astype(np.float32), (phi - h * u[n:]).astype(np.float32)
        lp = rollout(sim, Ep, pp, w, grad=False)[0]
        lm = rollout(sim, Em, pm, w, grad=False)[0]
        dt = np.log(Ep.astype(np.float64)) - np.log(Em.astype(np.float64))
        dp = pp.astype(np.float64) - pm.astype(np.float64)
        size = np.abs(ge) @ np.abs(dt) + np.abs(gp) @ np.abs(dp)
        out.append(abs((lp - lm) - (ge @ dt + gp @ dp)) / size)
    return np.array(out)


@pytest.mark.gpu
def test_c5_central_differences_fp64(c5):
    r = fd_discrepancies(c5[2], ndir=8)
    c5["fd_fp64"] = r
    print(f"[c5] FP64 fused: |FD - adjoint| / size over 8 directions: median {np.median(r):.2e}, worst "
          f"{r.max():.2e}", flush=True)
    assert r.max() <= 0.2


@pytest.mark.gpu
def test_c5_central_differences_mixed(c5):
    r = fd_discrepancies(c5[1])
    ref = c5.get("fd_fp64")
    print(f"[c5] mixed: |FD - adjoint| / size over 16 directions: median {np.median(r):.2e}, worst {r.max():.2e}"
          + ("" if ref is None else f" (FP64 fused over 8: median {np.median(ref):.2e}, worst {ref.max():.2e})"),
          flush=True)
    if ref is not None:
        assert np.median(r) <= 2 * np.median(ref) + 1e-5
    assert r.max() <= 0.2


def fd_rel(sim, theta, phi, w, ge, gp, u, h):
    """((L+ - L-) - adjoint . applied perturbation) / sum |gradient| |perturbation|"""
    n = len(theta)
    Ep, Em = np.exp(theta + h * u[:n]).astype(np.float32), np.exp(theta - h * u[:n]).astype(np.float32)
    pp, pm = (phi + h * u[n:]).astype(np.float32), (phi - h * u[n:]).astype(np.float32)
    lp = rollout(sim, Ep, pp, w, grad=False)[0]
    lm = rollout(sim, Em, pm, w, grad=False)[0]
    dt = np.log(Ep.astype(np.float64)) - np.log(Em.astype(np.float64))
    dp = pp.astype(np.float64) - pm.astype(np.float64)
    size = np.abs(ge) @ np.abs(dt) + np.abs(gp) @ np.abs(dp)
    return ((lp - lm) - (ge @ dt + gp @ dp)) / size


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [2, 1], ids=["fp64-fused", "mixed"])
def test_c5_horizon_smooth_setting_central_differences(precision):
    """32 of the robots, no floor, no gravity, 5 degC, 2,000 steps, a seeded weighted-position
    loss; 4 seeded directions, Richardson-extrapolated central differences.

    Steps 5e-3 and 2.5e-3 (at 1e-2 one of the directions already moves a link across the
    small/large-angle switch of CVX_Link::orientLink within the 2,000 steps; below 1e-3 the FP64
    rounding of a 2,000-step loss shows).  Bar 3e-5 of the derivative's size for both steps
    (measured 1e-6 to 1.2e-5: over 2,000 steps the difference quotients of the FP64 fused step
    carry ~1e-5 of rounding, against 1e-7 over 100 steps, where test_gpu_adjoint.py holds it to
    1e-6).  For the mixed step the test also measures the FP32 noise floor of its loss (its beam
    constants are FP32 values of the FP32 moduli, the reference's types): the discrepancy of a
    difference over h = 1e-6, which that floor dominates, times 1e-6."""
    sim = build(precision, robots=32, floor=False)
    n = sim.voxelCount()
    theta, phi = params(n)
    w = (W.uniform(77, np.arange(3 * n, dtype=np.uint64)) * 2 - 1).reshape(n, 3)
    _, (ge, gp) = rollout(sim, np.exp(theta).astype(np.float32), phi.astype(np.float32), w)
    h1, h2 = 5e-3, 2.5e-3
    bar = 3e-5
    if precision == 1:
        u0 = W.uniform(99, np.arange(2 * n, dtype=np.uint64)) * 2 - 1
        floor_rel = abs(fd_rel(sim, theta, phi, w, ge, gp, u0, 1e-6)) * 1e-6
        bar = max(bar, 3 * floor_rel / h2)
        print(f"[c5] smooth, mixed: FP32 noise floor of the loss {floor_rel:.2e} of the derivative's size per unit "
              f"step, {floor_rel / h2:.2e} at h={h2}: bar {bar:.2e}", flush=True)
    worst = 0.0
    for k in range(4):
        u = W.uniform(100 + k, np.arange(2 * n, dtype=np.uint64)) * 2 - 1
        r1, r2 = fd_rel(sim, theta, phi, w, ge, gp, u, h1), fd_rel(sim, theta, phi, w, ge, gp, u, h2)
        err = abs((4 * r2 - r1) / 3)
        worst = max(worst, err)
        print(f"[c5] smooth, precision {precision}, direction {k}: {r1:.2e} (h={h1}) {r2:.2e} (h={h2}) "
              f"extrapolated {err:.2e}", flush=True)
    assert worst <= bar, worst
