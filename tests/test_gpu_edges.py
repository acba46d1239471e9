"""Edge cases of doTimeStep on the device, each run through the C ABI next to the CPU oracle
built with correctly rounded trig, and compared bit for bit on both FP64 paths:

* a model with no voxels: a step only advances the clock, and recommendedTimeStep is 0
  (src/Voxelyze.cpp:251-311);
* doTimeStep(0) does nothing; doTimeStep(dt < 0) steps with recommendedTimeStep()
  (src/Voxelyze.cpp:253-254);
* one voxel with no links (its recommendedTimeStep from the voxel alone,
  src/Voxelyze.cpp:302-307) falling onto the floor;
* ragged topology: voxels removed and added between steps, then stepped on;
* the library's own limits: lattice indices beyond 16 bits are refused with an error;
* a lattice eight times the benchmark's (256x256x128 sites, 7.5M voxels, 20M links): the
  fused and split FP64 steps give the same bits, mixed precision stays within 1e-4.
"""
import numpy as np
import pytest

import backends as B
import scenarios as S
from voxelyze_amd import workloads as W

PATHS = [0, 1]
IDS = ["fused", "split"]


def same(a, b):
    sa, sb = S.snapshot(a), S.snapshot(b)
    bad = [k for k in sa if sa[k].tobytes() != sb[k].tobytes()]
    assert not bad, f"differs from the oracle in {bad}"
    assert np.float32(a.currentTime()).tobytes() == np.float32(b.currentTime()).tobytes()


def pair(path, build):
    ref = B.maker(B.oracle_cr_abi())(0.005)
    dev = B.maker(B.product_abi(), path)(0.005)
    for sim in (ref, dev):
        build(sim)
    return ref, dev


@pytest.mark.gpu
@pytest.mark.parametrize("path", PATHS, ids=IDS)
def test_empty_model_advances_the_clock_only(path):
    print(f"[edges] empty path {path}", flush=True)
    ref, dev = pair(path, lambda sim: sim.addMaterial(1e6, 1000.0))
    assert dev.voxelCount() == 0 and dev.linkCount() == 0
    assert dev.recommendedTimeStep() == ref.recommendedTimeStep() == 0.0
    for sim in (ref, dev):
        assert sim.step(1e-4, 25) == (25, -1)
        assert sim.doTimeStep(-1.0)  # recommendedTimeStep() = 0: a step of length 0
    same(ref, dev)
    assert dev.currentTime() > 0


def small_lattice(sim):
    W.drop_lattice(sim, 6, 5, 4, seed=7, lift=0)


@pytest.mark.gpu
@pytest.mark.parametrize("path", PATHS, ids=IDS)
def test_zero_and_negative_dt(path):
    print(f"[edges] dt 0 / dt < 0 path {path}", flush=True)
    ref, dev = pair(path, small_lattice)
    dt = dev.recommendedTimeStep()
    assert dt == ref.recommendedTimeStep() > 0
    for sim in (ref, dev):
        sim.step(dt, 10)
    before = S.snapshot(dev)
    t0 = dev.currentTime()
    assert dev.doTimeStep(0.0) and ref.doTimeStep(0.0)
    after = S.snapshot(dev)
    assert all(before[k].tobytes() == after[k].tobytes() for k in before)
    assert dev.currentTime() == t0
    for sim in (ref, dev):
        for _ in range(20):
            assert sim.doTimeStep(-1.0)
        assert sim.step(-1.0, 30) == (30, -1)
    same(ref, dev)
    # negative dt is recommendedTimeStep(): the same bits as stepping with it explicitly
    dev2 = B.maker(B.product_abi(), path)(0.005)
    small_lattice(dev2)
    dev2.step(dt, 10)
    dev2.step(dt, 50)
    same(dev, dev2)


@pytest.mark.gpu
@pytest.mark.parametrize("path", PATHS, ids=IDS)
def test_single_voxel_falls_onto_the_floor(path):
    print(f"[edges] single voxel path {path}", flush=True)

    def build(sim):
        m = sim.addMaterial(1e6, 1000.0)
        m.setStaticFriction(1.0)
        m.setKineticFriction(0.5)
        sim.setGravity(1.0)
        sim.enableFloor(True)
        sim.setVoxel(m, 0, 0, 2)

    ref, dev = pair(path, build)
    assert dev.linkCount() == 0
    dt = dev.recommendedTimeStep()
    assert dt == ref.recommendedTimeStep() > 0
    lowest = 1.0
    for _ in range(30):
        for sim in (ref, dev):
            assert sim.step(dt, 100) == (100, -1)
        lowest = min(lowest, dev.state()["pos"][0, 2])
    same(ref, dev)
    assert lowest < 0.0025  # it reached the floor (and bounced back up)


@pytest.mark.gpu
@pytest.mark.parametrize("path", PATHS, ids=IDS)
def test_ragged_topology_between_steps(path):
    print(f"[edges] ragged path {path}", flush=True)
    ref, dev = pair(path, lambda sim: W.drop_lattice(sim, 8, 8, 4, seed=11, lift=1))
    dt = dev.recommendedTimeStep()
    rng = np.random.default_rng(5)
    for rnd in range(3):
        for sim in (ref, dev):
            assert sim.step(dt, 100) == (100, -1)
        same(ref, dev)
        ijk = S.snapshot(dev)["vox_ijk"]
        gone = ijk[rng.choice(len(ijk), 6, replace=False)]
        new = [(int(x), int(y), 9) for x, y in rng.integers(0, 8, size=(3, 2))]
        for sim in (ref, dev):
            m = sim.material(0)
            for x, y, z in gone:
                sim.setVoxel(None, int(x), int(y), int(z))
            for x, y, z in new:
                sim.setVoxel(m, x, y, z)
        assert dev.voxelCount() == ref.voxelCount() and dev.linkCount() == ref.linkCount()
        dt = dev.recommendedTimeStep()
        assert dt == ref.recommendedTimeStep()
    for sim in (ref, dev):
        assert sim.step(dt, 100) == (100, -1)
    same(ref, dev)


@pytest.mark.gpu
def test_lattice_index_limit():
    sim = B.maker(B.product_abi())(0.005)
    m = sim.addMaterial(1e6, 1000.0)
    assert sim.setVoxel(m, 32767, -32768, 0) is not None
    assert sim.setVoxel(m, 40000, 0, 0) is None
    assert "-32768..32767" in sim.lastError()
    assert sim.voxelCount() == 1


@pytest.mark.gpu
def test_lattice_eight_times_the_benchmark():
    """256x256x128 sites: the fused and split FP64 steps give the same bits after 30 steps;
    mixed precision stays within 1e-4 relative displacement of them."""
    print("[edges] 256x256x128 lattice", flush=True)
    sims = []
    for path, prec in ((0, 0), (1, 0), (0, 1)):
        sim = B.maker(B.product_abi(), path)(0.005)
        W.drop_lattice(sim, 256, 256, 128, seed=2, lift=0)
        sim.setPrecision(prec)
        sims.append(sim)
    nominal = sims[0].state()["pos"]
    V, L = sims[0].voxelCount(), sims[0].linkCount()
    print(f"[edges] {V} voxels, {L} links", flush=True)
    assert V > 7_000_000 and L > 19_000_000
    dt = sims[0].recommendedTimeStep()
    for sim in sims:
        assert sim.step(dt, 30) == (30, -1)
    a, b, c = (s.state() for s in sims)
    assert all(a[k].tobytes() == b[k].tobytes() for k in a)
    la, lb = sims[0].link_state(), sims[1].link_state()
    assert all(la[k].tobytes() == lb[k].tobytes() for k in la)
    err = np.abs(c["pos"] - a["pos"]).max() / np.abs(a["pos"] - nominal).max()
    print(f"[edges] 7.5M voxels: fused == split bitwise, mixed rel disp err {err:.3e}", flush=True)
    assert err <= 1e-4
