"""The adjoint of self-collision (vx_adjoint.h pair_force / pair_vjp, vx_adjoint.hip
k_adj_contacts): the penalty contact of two voxels (src/VX_Collision.cpp:42-56) transposed per
voxel over its partner list, with the partner lists each recomputed step used.

Model: two robots, each two 3x3x3 blocks of the +0.01/degC actuated material side by side with
one empty lattice site between them, no gravity and no floor. At +-30 degC and phases near
pi/2 the blocks swell by ~30 % and press into each other across the gap (contact at 1.25
pitch between voxel centres), then push apart; the two blocks are not linked, so every facing
pair is a listed contact. Per-voxel log modulus theta and phase phi as in the rest of the
adjoint tests; loss = sum(w * final positions).

Checks:
  * the contacts act: the loss moves by more than its own size when collisions are off;
  * FP64 fused step, 200 and 1,000 steps, with and without collision damping: central
    differences at h = 1e-5 and 3e-6 (the better of the two, as a contact onset is a kink, or
    with damping a jump, of the force) along 8 seeded directions agree with the adjoint within
    1e-5 of the derivative's size (measured <= 3.3e-7);
  * the mixed step's adjoint agrees with the FP64 fused step's;
  * the checkpoint interval does not change the gradient, bit for bit (watch lists restored
    from checkpoints and rebuilt mid-segment reproduce the forward's).
"""
import math

import numpy as np
import pytest

import backends as B
from test_gpu_adjoint import weighted_loss_f
from voxelyze_amd import workloads as W

AMP = 30.0


def build(precision, collide=True, damping=0.0, n=3, gap=1):
    sim = B.maker(B.product_abi())(0.005)
    mats = W.robot_materials(sim)
    for m in mats:
        m.setCollisionDamping(damping)
    ijk = np.asarray([(x + s * (n + gap), y, z) for x in range(n) for y in range(n) for z in range(n) for s in (0, 1)],
                     np.int32)
    for r in range(2):
        if r:
            sim.addRobot()
        sim.setVoxels(np.full(len(ijk), mats[2].index), ijk)
    sim.setGravity(0.0)
    sim.enableFloor(False)
    sim.setActuation(True, 0.0, AMP, 2.0)
    if collide:
        sim.enableCollisions()
    sim.setPrecision(precision)
    return sim


def params(n, seed=3):
    u = W.uniform(seed, np.arange(2 * n, dtype=np.uint64))
    theta = np.log(1e6) + 0.3 * (u[:n] - 0.5)
    phi = math.pi / 2 + 0.4 * (u[n:] - 0.5)
    return np.exp(theta).astype(np.float32), phi.astype(np.float32)


def weights(n):
    return (W.uniform(78, np.arange(3 * n, dtype=np.uint64)) * 2 - 1).reshape(n, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps,damping", [(200, 0.0), (200, 0.5), (1000, 0.0)])
def test_contact_adjoint_matches_central_differences(nsteps, damping):
    sim = build(2, damping=damping)
    n = sim.voxelCount()
    E, phi = params(n)
    theta = np.log(E.astype(np.float64))
    w = weights(n)
    loss, ge, gp = weighted_loss_f(sim, E, phi, w, nsteps)
    free, _, _ = weighted_loss_f(build(2, collide=False, damping=damping), E, phi, w, nsteps, grad=False)
    print(f"[contact adjoint] {nsteps} steps, damping {damping}: loss {loss:.9e}, without contacts {free:.9e}")
    assert abs(loss - free) > abs(loss)  # the blocks press into each other
    worst = 0.0
    for d in range(8):
        u = W.uniform(200 + d, np.arange(2 * n, dtype=np.uint64)) * 2 - 1

        def rel(h):
            Ep, Em = np.exp(theta + h * u[:n]).astype(np.float32), np.exp(theta - h * u[:n]).astype(np.float32)
            pp, pm = (phi + h * u[n:]).astype(np.float32), (phi - h * u[n:]).astype(np.float32)
            lp, _, _ = weighted_loss_f(sim, Ep, pp, w, nsteps, grad=False)
            lm, _, _ = weighted_loss_f(sim, Em, pm, w, nsteps, grad=False)
            dt = np.log(Ep.astype(np.float64)) - np.log(Em.astype(np.float64))
            dp = pp.astype(np.float64) - pm.astype(np.float64)
            size = np.abs(ge) @ np.abs(dt) + np.abs(gp) @ np.abs(dp)
            return abs((lp - lm) - (ge @ dt + gp @ dp)) / size

        r1, r2 = rel(1e-5), rel(3e-6)
        worst = max(worst, min(r1, r2))
        print(f"[contact adjoint] direction {d}: h=1e-5 {r1:.2e}, h=3e-6 {r2:.2e}", flush=True)
    assert worst <= 1e-5, worst


@pytest.mark.gpu
def test_mixed_contact_adjoint_equals_fp64_adjoint():
    a, b = build(2), build(1)
    n = a.voxelCount()
    E, phi = params(n, seed=4)
    w = weights(n)
    la, gea, gpa = weighted_loss_f(a, E, phi, w, 300)
    lb, geb, gpb = weighted_loss_f(b, E, phi, w, 300)
    et = np.abs(geb - gea).max() / np.abs(gea).max()
    ep = np.abs(gpb - gpa).max() / np.abs(gpa).max()
    print(f"[contact adjoint] loss fp64 {la:.9e} mixed {lb:.9e}; gradient difference theta {et:.2e} phi {ep:.2e}")
    assert abs(la - lb) <= 1e-5 * abs(la)
    assert et <= 1e-3 and ep <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 2])
def test_contact_gradient_does_not_depend_on_the_checkpoint_interval(precision):
    sim = build(precision)
    n = sim.voxelCount()
    E, phi = params(n, seed=5)
    w = weights(n)
    out = []
    for every in (1000, 37):
        sim.resetTime()
        sim.setModuli(E)
        sim.setPhases(phi)
        sim.adjointRun(1000, every)
        out.append(sim.adjointBackward(w))
    assert out[0][0].tobytes() == out[1][0].tobytes() and out[0][1].tobytes() == out[1][1].tobytes()
