"""The fused FP64 step (one kernel per step, vx_fused.hip) against the split FP64 step (a link
kernel and a voxel kernel) and the oracle: the same bits in every state the C ABI exposes,
end forces included, across the cases where the fused step has its own machinery --
divergence in the last step of a launch and before it (the saved link state and the
recomputed end forces), robots that do not step, device actuation, self-collision, switching
between the two paths, and replayed step graphs.

Every scenario of tests/scenarios.py is also compared with the correctly rounded oracle on
both paths (test_gpu_parity.py).
"""
import math

import numpy as np
import pytest

import backends as B
import scenarios as S
from voxelyze_amd import workloads as W

FUSED, SPLIT = 0, 1


def bits_equal(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def same_state(a, b, what):
    sa, sb = S.snapshot(a), S.snapshot(b)
    bad = [k for k in sa if not bits_equal(sa[k], sb[k])]
    assert not bad, f"{what}: differs in {bad[:8]}"


def two_bar(make):
    """two voxels, one fixed, one pulled hard enough to diverge at a large dt"""
    sim = make(0.001)
    m = sim.addMaterial(1e6, 1e3)
    sim.setVoxel(m, 0, 0, 0).external().setFixedAll()
    sim.setVoxel(m, 1, 0, 0).external().setForce(10.0, 0, 0)
    return sim


@pytest.mark.gpu
def test_divergence_in_single_step_launches():
    """doTimeStep until it returns false: the diverging step is the launch's last one."""
    sims = [two_bar(B.maker(B.product_abi(), p)) for p in (FUSED, SPLIT)]
    ref = two_bar(B.maker(B.oracle_cr_abi()))
    dt = ref.recommendedTimeStep() * 40
    steps = []
    for sim in sims + [ref]:
        k = 0
        while k < 100 and sim.doTimeStep(dt):
            k += 1
        steps.append(k)
    print(f"[fused64] single-step divergence after {steps}", flush=True)
    assert steps[0] == steps[1] == steps[2] < 100
    same_state(sims[0], sims[1], "fused vs split")
    same_state(sims[0], ref, "fused vs oracle")
    assert sims[0].currentTime() == sims[1].currentTime() == ref.currentTime()


@pytest.mark.gpu
def test_divergence_inside_a_launch():
    """One launch of 40 steps that diverges early: voxels as before the diverging step, links
    as after it, end forces of the diverging step (recomputed after the launch)."""
    sims = [two_bar(B.maker(B.product_abi(), p)) for p in (FUSED, SPLIT)]
    ref = two_bar(B.maker(B.oracle_cr_abi()))
    dt = ref.recommendedTimeStep() * 40
    res = [s.step(dt, 40) for s in sims + [ref]]
    print(f"[fused64] divergence inside a launch: {res}", flush=True)
    assert res[0] == res[1] == res[2] and 0 <= res[0][1] < 38
    same_state(sims[0], sims[1], "fused vs split")
    same_state(sims[0], ref, "fused vs oracle")
    assert sims[0].currentTime() == sims[1].currentTime() == ref.currentTime()


def robots(path, actuated=True, collisions=False, seed=31):
    sim = B.maker(B.product_abi(), path)(0.005)
    shp = W.robot_shapes(seed, 4, vmin=40, vmax=160, box=8)
    tot = sum(len(s[0]) for s in shp)
    ph = (W.uniform(7, np.arange(tot, dtype=np.uint64)) * 2 * math.pi).astype(np.float32)
    W.robot_population(sim, shp, 39.0 if actuated else 0.0, 2.0, phases=ph)
    if collisions:
        sim.enableCollisions(True)
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("collisions", [False, True], ids=["free", "contacts"])
def test_actuated_robots_match_split(collisions):
    a, b = robots(FUSED, collisions=collisions), robots(SPLIT, collisions=collisions)
    for _ in range(3):
        da, db = a.robotStep(120), b.robotStep(120)
        assert np.array_equal(da, db)
    same_state(a, b, "actuated robots")
    for r in range(a.robotCount()):
        assert a.robotTime(r) == b.robotTime(r)


@pytest.mark.gpu
def test_diverging_and_idle_robots_match_split():
    """Robot 1 blows up mid-launch, robot 2 does not step (dt 0); the others run on."""
    a, b = robots(FUSED), robots(SPLIT)
    for sim in (a, b):
        f1, _ = sim.robotVoxels(1)
        sim.voxel(f1).external().setForce(1e4, 0.0, 0.0)
    dt = np.array([a.robotRecommendedTimeStep(r) for r in range(a.robotCount())], np.float32)
    dt[2] = 0.0
    da, db = a.robotStep(150, dt), b.robotStep(150, dt)
    print(f"[fused64] robots done {da} / {db}", flush=True)
    assert np.array_equal(da, db) and da[1] < 150 and da[0] == 150
    same_state(a, b, "diverging and idle robots")
    for r in range(a.robotCount()):
        assert a.robotTime(r) == b.robotTime(r)


@pytest.mark.gpu
def test_switching_paths_mid_run():
    a = B.maker(B.product_abi(), SPLIT)(0.005)
    b = B.maker(B.product_abi(), SPLIT)(0.005)
    W.cantilever(a)
    W.cantilever(b)
    dt = a.recommendedTimeStep()
    a.step(dt, 301)
    b.step(dt, 100)
    b.setFp64Path(FUSED)
    b.step(dt, 101)
    b.setFp64Path(SPLIT)
    b.step(dt, 100)
    same_state(a, b, "split/fused/split")
    assert a.currentTime() == b.currentTime()


@pytest.mark.gpu
def test_fused_graph_replays_equal_plain_launches():
    import ctypes as C
    a = B.maker(B.product_abi(), FUSED)(0.005)
    b = B.maker(B.product_abi(), FUSED)(0.005)
    for sim in (a, b):
        W.cantilever(sim)
    b.abi.set_graphs(b.handle, 0)
    dt = a.recommendedTimeStep()
    for rep in range(4):
        a.step(dt, 25)  # odd length: the ping-pong parity flips on every replay
        b.step(dt, 25)
        same_state(a, b, f"replay {rep}")
    cap, rep = C.c_int(0), C.c_int(0)
    a.abi.graph_stats(a.handle, C.byref(cap), C.byref(rep))
    assert cap.value >= 1 and rep.value >= 1


@pytest.mark.gpu
@pytest.mark.parametrize("path", [FUSED, SPLIT], ids=["fused", "split"])
def test_random_poisson_models_match_the_oracle(path):
    """Random small models whose materials mostly have a Poisson's ratio, stepped in blocks
    with the state-dependent recommended step: bit for bit the correctly rounded oracle
    (end forces and transverse info included)."""
    import test_oracle as T
    for seed in range(4):
        a = T.random_model(B.maker(B.product_abi(), path), seed, poisson=True)
        b = T.random_model(B.maker(B.oracle_cr_abi()), seed, poisson=True)
        for k in range(4):
            dt = b.recommendedTimeStep()
            assert np.float32(a.recommendedTimeStep()) == np.float32(dt)
            assert a.step(dt, 40) == b.step(dt, 40)
            same_state(a, b, f"seed {seed} block {k}")
            ta, tb = a.link_transverse(), b.link_transverse()
            assert all(bits_equal(ta[key], tb[key]) for key in ta), f"seed {seed} block {k}"


@pytest.mark.gpu
def test_poisson_runs_in_the_fp64_step_only():
    sim = B.maker(B.product_abi())(0.001)
    m = sim.addMaterial(1e6, 1e3)
    m.setPoissonsRatio(0.3)
    for i in range(3):
        sim.setVoxel(m, i, 0, 0)
    sim.setPrecision(1)
    with pytest.raises(RuntimeError):
        sim.step(sim.recommendedTimeStep(), 5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,path,collisions", [(0, FUSED, False), (0, SPLIT, False), (1, FUSED, False),
                                                        (0, FUSED, True), (1, FUSED, True)],
                         ids=["fp64-fused", "fp64-split", "mixed", "fp64-contacts", "mixed-contacts"])
def test_checkpoint_restore_replays_the_same_steps(precision, path, collisions):
    """Save the dynamic state, step, restore, step again: the same bits, the same clocks."""
    sim = robots(path, collisions=collisions)
    sim.setPrecision(precision)
    sim.robotStep(60)
    assert sim.checkpointSave(3)
    sim.robotStep(45)
    first, t1 = S.snapshot(sim), [sim.robotTime(r) for r in range(sim.robotCount())]
    assert sim.checkpointRestore(3)
    sim.robotStep(45)
    second, t2 = S.snapshot(sim), [sim.robotTime(r) for r in range(sim.robotCount())]
    bad = [k for k in first if not bits_equal(first[k], second[k])]
    assert not bad and t1 == t2, bad[:8]


@pytest.mark.gpu
def test_checkpoint_refuses_another_topology():
    sim = B.maker(B.product_abi())(0.005)
    W.cantilever(sim)
    sim.step(sim.recommendedTimeStep(), 10)
    assert sim.checkpointSave(0)
    m = sim.addMaterial(1e6, 1e3)
    sim.setVoxel(m, 30, 0, 0)
    with pytest.raises(RuntimeError):
        sim.checkpointRestore(0)


@pytest.mark.gpu
@pytest.mark.parametrize("edit", ["remove-readd", "replace"])
def test_checkpoint_refuses_a_count_preserving_topology_edit(edit):
    """Removing voxel 0 and adding it again keeps the voxel and link counts but shifts every
    index; replacing a voxel's material keeps them too.  Both must refuse the restore."""
    sim = B.maker(B.product_abi())(0.005)
    W.cantilever(sim)
    sim.step(sim.recommendedTimeStep(), 10)
    nv, nl = sim.voxelCount(), sim.linkCount()
    v0 = sim.voxel(0)
    ijk, mat = (v0.indexX(), v0.indexY(), v0.indexZ()), sim.material(0)
    assert sim.checkpointSave(0)
    if edit == "remove-readd":
        sim.setVoxel(None, *ijk)
        sim.setVoxel(mat, *ijk)
    else:
        sim.setVoxel(sim.addMaterial(2e6, 1e3), *ijk)
    assert (sim.voxelCount(), sim.linkCount()) == (nv, nl)
    with pytest.raises(RuntimeError):
        sim.checkpointRestore(0)
    # a checkpoint saved after the edit restores
    assert sim.checkpointSave(1)
    sim.step(sim.recommendedTimeStep(), 3)
    assert sim.checkpointRestore(1)


def bimorph(make):
    """a 6-voxel-long thermal bimorph (one expanding layer), fixed at x = 0"""
    sim = make(0.001)
    m1 = sim.addMaterial(1e6, 1e3)
    m1.setInternalDamping(1.0)
    m1.setCte(0.01)
    m2 = sim.addMaterial(1e7, 1e3)
    m2.setInternalDamping(1.0)
    for i in range(6):
        a = sim.setVoxel(m1, i, 0, 0)
        b = sim.setVoxel(m2, i, 0, 1)
        if i == 0:
            a.external().setFixedAll()
            b.external().setFixedAll()
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("path", [FUSED, SPLIT], ids=["fused", "split"])
def test_temperatures_set_between_multi_step_launches(path):
    """Temperatures set on the host change rest lengths (CVX_Voxel::setTemperature,
    VX_Voxel.cpp:108-114); a multi-step launch without actuation keeps them in both halves of
    the fused step's link state instead of rewriting them every step.  Launches of 40, 41, 7,
    1, 1, 1 and 10 steps with ambient and per-voxel temperature changes between them give the
    bits of the oracle stepped one doTimeStep at a time."""
    dev = bimorph(B.maker(B.product_abi(), path))
    ref = bimorph(B.maker(B.oracle_cr_abi()))
    dt = ref.recommendedTimeStep()
    assert dev.recommendedTimeStep() == dt

    def edit(sim, phase):
        if phase == 0:
            sim.setAmbientTemperature(5, True)
        elif phase == 1:
            sim.voxel(3).setTemperature(12.0)
            sim.voxel(5).setTemperature(-4.0)
        elif phase == 2:
            sim.setAmbientTemperature(-3, True)
        elif phase == 4:
            sim.voxel(4).setTemperature(20.0)

    for phase, n in enumerate([40, 41, 7, 1, 1, 1, 10]):
        edit(dev, phase)
        edit(ref, phase)
        done, _ = dev.step(dt, n)
        assert done == n
        for _ in range(n):
            assert ref.doTimeStep(dt)
        same_state(dev, ref, f"after launch {phase} ({n} steps)")
    print(f"[fused64] path {path}: temperatures between launches, bit-exact", flush=True)
