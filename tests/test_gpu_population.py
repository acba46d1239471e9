"""The robot population at its benchmark size: 4,096 robots grown to 100-8,000 voxels in a
20^3 box (16.5M voxels), four materials, +-39 degC at 2 Hz actuated on the device, gravity and
floor friction, every robot with its own recommended dt, mixed precision, 10,000 steps.

The reference has no batch (each robot behaves as a CVoxelyze of its own), so at this size
the checks are properties:
  * every robot completes the 10,000 steps, and every voxel's pose is finite;
  * each robot's clock is its own dt summed 10,000 times in float (currentTime += dt,
    src/Voxelyze.cpp:278);
  * no robot leaves its neighbourhood (centre-of-mass travel well under the 0.1 m box);
  * robots are independent: sampled robots stepped in a batch of their own (first, second,
    middle, last and largest) end with the same bits as inside the full population.
"""
import math

import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W

STEPS = 10000


def population(shp, phases):
    sim = B.maker(B.product_abi())(0.005)
    W.robot_population(sim, shp, phases=phases)
    sim.setPrecision(1)
    return sim


@pytest.mark.gpu
def test_full_population_properties():
    shp = W.robot_shapes(4, 4096)
    sizes = np.array([len(s[0]) for s in shp])
    ph = (W.uniform(9, np.arange(sizes.sum(), dtype=np.uint64)) * 2 * math.pi).astype(np.float32)
    sim = population(shp, ph)
    R = sim.robotCount()
    assert R == 4096 and sim.voxelCount() == sizes.sum()
    dts = [sim.robotRecommendedTimeStep(r) for r in range(R)]
    com0 = sim.robotCom()
    done = np.zeros(R, np.int64)
    for _ in range(STEPS // 500):
        done += sim.robotStep(500)
    print(f"[population] {R} robots, {sizes.sum()} voxels, {STEPS} steps", flush=True)
    assert (done == STEPS).all()
    st = sim.state()
    pos = st["pos"]
    assert np.isfinite(pos).all() and np.isfinite(st["orient"]).all()
    for r in (0, 1, R // 2, R - 1):
        t = np.float32(0.0)
        for _ in range(STEPS):
            t = np.float32(t + np.float32(dts[r]))
        assert sim.robotTime(r) == float(t), r
    travel = sim.robotCom()[:, 0] - com0[:, 0]
    print(f"[population] CoM x-travel: median |dx| {np.median(np.abs(travel)):.2e} m, max {np.abs(travel).max():.2e} m")
    assert np.isfinite(travel).all() and np.abs(travel).max() < 0.05
    firsts = np.concatenate([[0], np.cumsum(sizes)])
    sample = sorted({0, 1, R // 2, R - 1, int(np.argmax(sizes))})
    sub = population([shp[r] for r in sample], np.concatenate([ph[firsts[r]:firsts[r + 1]] for r in sample]))
    for _ in range(STEPS // 500):
        assert (sub.robotStep(500) == 500).all()
    ps = sub.state()["pos"]
    for k, r in enumerate(sample):
        f, c = sub.robotVoxels(k)
        same = ps[f:f + c].tobytes() == pos[firsts[r]:firsts[r] + c].tobytes()
        print(f"[population] robot {r} ({c} voxels) alone in a small batch: bit-identical {same}", flush=True)
        assert same
