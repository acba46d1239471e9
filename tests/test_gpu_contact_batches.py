"""Contact steps in batches without host round trips (vx_mixed_host.cpp, vx_fused_host.cpp):
a multi-step launch runs up to 64 contact steps between host synchronisations and halts on
the device where the watch list goes stale (the step kernels test it on the positions they
write), then rebuilds it and resumes.
The watch list must be tested and rebuilt exactly where the step-by-step loop does it
(src/Voxelyze.cpp:686-698): one launch of N steps equals N one-step launches bit for bit, on
every precision path, while a soft cube falls onto a fixed plate (moving several half
pitches, so the list goes stale many times inside batches) and lands on it.
"""
import numpy as np
import pytest

import backends as B


def drop(precision):
    sim = B.maker(B.product_abi())(0.001)
    soft = sim.addMaterial(2e4, 1000.0)
    soft.setCollisionDamping(0.2)
    for x in range(4):
        for y in range(4):
            sim.setVoxel(soft, x, y, 0).external().setFixedAll()
    for x in range(1, 3):
        for y in range(1, 3):
            for z in range(5, 7):
                sim.setVoxel(soft, x, y, z)
    sim.setGravity(1.0)
    sim.enableCollisions(True)
    sim.setPrecision(precision)
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_one_launch_equals_single_steps_with_contacts(precision):
    a, b = drop(precision), drop(precision)
    dt = a.recommendedTimeStep()
    n = 900
    z0 = a.state()["pos"][:, 2].max()
    done, _ = a.step(dt, n)
    assert done == n
    for _ in range(n):
        b.step(dt, 1)
    sa, sb = a.state(), b.state()
    fall = z0 - sa["pos"][:, 2].max()
    print(f"[contact batches] precision {precision}: top fell {fall / 0.001:.2f} pitch, "
          f"{a.collisionCount()} pairs listed")
    assert fall > 2.0 * 0.001  # the cube fell onto the plate: the list went stale many times
    assert a.collisionCount() == b.collisionCount()
    for k in ("pos", "orient", "linMom", "angMom"):
        if k in sa:
            assert sa[k].tobytes() == sb[k].tobytes(), k


def robots_with_contacts(precision):
    """four robots with self-collision; robot 1 is blown apart by a 1e4 N force on one voxel, so
    its watch list goes stale every few steps and it diverges inside a batch"""
    from voxelyze_amd import workloads as W
    sim = B.maker(B.product_abi())(0.005)
    W.robot_population(sim, W.robot_shapes(51, 4, vmin=40, vmax=220, box=8))
    sim.enableCollisions(True)
    sim.setPrecision(precision)
    f1, _ = sim.robotVoxels(1)
    sim.voxel(f1).external().setForce(1e4, 0.0, 0.0)
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batches_with_a_diverging_robot_equal_single_steps(precision):
    """Halts of contact batches and a robot's divergence inside them: one launch of N robot
    steps equals N one-step launches bit for bit (state, clocks, steps done per robot)."""
    a, b = robots_with_contacts(precision), robots_with_contacts(precision)
    n = 300
    done_a = a.robotStep(n)
    done_b = np.zeros(b.robotCount(), np.int64)
    for _ in range(n):
        done_b += b.robotStep(1)
    print(f"[contact batches] precision {precision}: steps done {done_a} (one launch), {done_b} (single steps)")
    assert done_a[1] < n and all(done_a[r] == n for r in (0, 2, 3))
    assert np.array_equal(done_a, done_b)
    sa, sb = a.state(), b.state()
    for k in ("pos", "orient", "linMom", "angMom"):
        assert sa[k].tobytes() == sb[k].tobytes(), k
    for r in range(a.robotCount()):
        assert a.robotTime(r) == b.robotTime(r)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batches_when_every_robot_stops(precision):
    """Every robot diverges inside the first batch: the launch ends with the step that settles
    the ping-pong buffers, which runs outside any batch (a halt the batch recorded must not make
    it return early).  One launch equals one-step launches bit for bit."""
    def blown(prec):
        sim = robots_with_contacts(prec)
        for r in range(sim.robotCount()):
            f, _ = sim.robotVoxels(r)
            sim.voxel(f).external().setForce(1e4, 0.0, 0.0)
        return sim
    a, b = blown(precision), blown(precision)
    n = 200
    done_a = a.robotStep(n)
    done_b = np.zeros(b.robotCount(), np.int64)
    for _ in range(n):
        done_b += b.robotStep(1)
    print(f"[contact batches] precision {precision}: every robot stops: {done_a} (one launch), {done_b} (single steps)")
    assert (done_a < n).all() and np.array_equal(done_a, done_b)
    sa, sb = a.state(), b.state()
    for k in ("pos", "orient", "linMom", "angMom"):
        assert sa[k].tobytes() == sb[k].tobytes(), k
    for r in range(a.robotCount()):
        assert a.robotTime(r) == b.robotTime(r)
