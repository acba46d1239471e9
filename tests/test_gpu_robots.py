"""Robots: many independent bodies stepped together (vxg_robot_*), with device-side
actuation (vxg_set_actuation).

The reference has no batch: a robot is defined to behave exactly as a CVoxelyze of its
own, driven by a host loop that sets every voxel's temperature before each doTimeStep
(CVX_Voxel::setTemperature, VX_Voxel.cpp:108-114).  So the checks are:
  * FP64 batch == each robot alone in its own simulation of the same library, driven
    by that host loop: bit-exact (same device arithmetic, robot-local coordinates);
  * the same host loop on the CPU oracle: within 1e-9 relative displacement (the device
    evaluates correctly rounded sin/cos/acos, glibc does not always);
  * mixed-precision batch vs FP64 batch: within 1e-4 relative displacement;
  * a diverging robot stops alone; the centre of mass equals numpy's.
"""
import math

import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W
from voxelyze_amd.api import _dptr

AMP, FREQ = 39.0, 2.0


def shapes(n=5, seed=21, vmin=40, vmax=220):
    return W.robot_shapes(seed, n, vmin=vmin, vmax=vmax, box=8)


def phases_for(shp, seed=5):
    tot = sum(len(s[0]) for s in shp)
    return (W.uniform(seed, np.arange(tot, dtype=np.uint64)) * 2 * math.pi).astype(np.float32)


def batch(shp, precision=0, phases=None, collisions=False):
    sim = B.maker(B.product_abi())(0.005)
    W.robot_population(sim, shp, AMP, FREQ, phases=phases)
    if collisions:
        sim.enableCollisions(True)
    sim.setPrecision(precision)
    return sim


def alone(abi, ijk, mat, collisions=False):
    """One robot as a plain simulation, materials in the same order."""
    sim = B.maker(abi)(0.005)
    mats = W.robot_materials(sim)
    sim.setGravity(1.0)
    sim.enableFloor(True)
    for (x, y, z), m in zip(ijk, mat):
        sim.setVoxel(mats[int(m)], int(x), int(y), int(z))
    if collisions:
        sim.enableCollisions(True)
    return sim


def host_drive(sim, dt, nsteps, phase):
    """for each step: setTemperature(T(t)) on every voxel, then doTimeStep(dt)."""
    n = sim.voxelCount()
    vox = [sim.voxel(i) for i in range(n)]
    for k in range(nsteps):
        t = float(np.float32(sim.currentTime()))
        for i in range(n):
            T = np.float32(0.0 + AMP * math.sin(6.283185307179586 * FREQ * t + float(phase[i])))
            vox[i].setTemperature(float(T))
        if not sim.doTimeStep(dt):
            return k
    return nsteps


def rel(ref, out, nominal):
    return np.abs(out - ref).max() / max(np.abs(ref - nominal).max(), 1e-300)


@pytest.mark.gpu
def test_fp64_batch_equals_robots_alone():
    shp = shapes()
    ph = phases_for(shp)
    sim = batch(shp, phases=ph)
    R = sim.robotCount()
    assert R == len(shp)
    dts = np.array([sim.robotRecommendedTimeStep(r) for r in range(R)], np.float32)
    nsteps = 150
    done = sim.robotStep(nsteps)
    assert (done == nsteps).all()
    pos = sim.state()["pos"]
    off = 0
    for r, (ijk, mat) in enumerate(shp):
        first, count = sim.robotVoxels(r)
        assert (first, count) == (off, len(ijk))
        ref = alone(B.product_abi(), ijk, mat)
        assert ref.recommendedTimeStep() == dts[r]
        assert host_drive(ref, float(dts[r]), nsteps, ph[off:off + count]) == nsteps
        p_ref = ref.state()["pos"]
        same = pos[first:first + count].tobytes() == p_ref.tobytes()
        print(f"[robots] robot {r}: {count} voxels, dt {dts[r]:.4g}, bit-exact {same}", flush=True)
        assert same
        assert sim.robotTime(r) == ref.currentTime()
        off += count


@pytest.mark.gpu
def test_fp64_batch_against_oracle_host_loop():
    shp = shapes(n=3, seed=31)
    ph = phases_for(shp, seed=6)
    sim = batch(shp, phases=ph)
    nsteps = 120
    sim.robotStep(nsteps)
    pos = sim.state()["pos"]
    for r, (ijk, mat) in enumerate(shp):
        first, count = sim.robotVoxels(r)
        ref = alone(B.oracle_abi(), ijk, mat)
        dt = sim.robotRecommendedTimeStep(r)
        nominal = ref.state()["pos"]
        assert host_drive(ref, dt, nsteps, ph[first:first + count]) == nsteps
        err = rel(ref.state()["pos"], pos[first:first + count], nominal)
        print(f"[robots] robot {r} vs oracle: rel disp err {err:.3e}", flush=True)
        assert err <= 1e-9


@pytest.mark.gpu
def test_mixed_batch_within_1e4_of_fp64():
    shp = shapes(n=6, seed=41, vmin=100, vmax=500)
    a = batch(shp, 0)
    b = batch(shp, 1)
    nominal = a.state()["pos"]
    ca0 = a.robotCom()
    for _ in range(4):
        assert (a.robotStep(500) == 500).all()
        assert (b.robotStep(500) == 500).all()
    pa, pb = a.state()["pos"], b.state()["pos"]
    for r in range(a.robotCount()):
        f, c = a.robotVoxels(r)
        err = rel(pa[f:f + c], pb[f:f + c], nominal[f:f + c])
        print(f"[robots] robot {r}: mixed vs fp64 rel disp err {err:.3e}", flush=True)
        assert err <= 1e-4
    dxa = a.robotCom()[:, 0] - ca0[:, 0]
    dxb = b.robotCom()[:, 0] - ca0[:, 0]
    print(f"[robots] CoM x-travel fp64 {dxa}, mixed {dxb}")
    assert np.all(np.abs(dxa - dxb) <= 1e-4 * np.abs(nominal).max())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_a_diverging_robot_stops_alone(precision):
    shp = shapes(n=4, seed=51)
    quiet = batch(shp, precision)
    loud = batch(shp, precision)
    f1, c1 = loud.robotVoxels(1)
    loud.voxel(f1).external().setForce(1e4, 0.0, 0.0)  # robot 1 blows up
    nsteps = 200
    done_q = quiet.robotStep(nsteps)
    done_l = loud.robotStep(nsteps)
    print(f"[robots] precision {precision}: steps done {done_l}")
    assert (done_q == nsteps).all()
    assert done_l[1] < nsteps and all(done_l[r] == nsteps for r in (0, 2, 3))
    pq, pl = quiet.state()["pos"], loud.state()["pos"]
    for r in (0, 2, 3):
        f, c = quiet.robotVoxels(r)
        assert pq[f:f + c].tobytes() == pl[f:f + c].tobytes()
        assert quiet.robotTime(r) == loud.robotTime(r)
    # the stopped robot: as the same robot alone, stepped until doTimeStep returns false
    ref = batch([shp[1]], precision)
    ref.voxel(0).external().setForce(1e4, 0.0, 0.0)
    d = ref.robotStep(nsteps)
    assert d[0] == done_l[1]
    pr = ref.state()["pos"]
    assert np.array_equal(pr, pl[f1:f1 + c1])
    assert ref.robotTime(0) == loud.robotTime(1)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_a_robot_with_dt_zero_keeps_its_state(precision):
    """doTimeStep(0) does nothing (Voxelyze.cpp:253): a robot given dt 0 in a batch launch
    keeps its state and clock bit for bit while the others step."""
    shp = shapes(n=4, seed=71)
    sim = batch(shp, precision)
    sim.robotStep(30)
    f, c = sim.robotVoxels(2)
    before = sim.state(f, c)
    t2, t0 = sim.robotTime(2), sim.robotTime(0)
    dt = np.array([sim.robotRecommendedTimeStep(r) for r in range(sim.robotCount())], np.float32)
    dt[2] = 0.0
    done = sim.robotStep(41, dt)
    print(f"[robots] precision {precision}: dt 0 robot, steps done {done}")
    after = sim.state(f, c)
    assert all(before[k].tobytes() == after[k].tobytes() for k in before)
    assert sim.robotTime(2) == t2
    assert sim.robotTime(0) > t0 and (done == 41).all()


@pytest.mark.gpu
def test_robot_com_is_the_mass_weighted_mean():
    shp = shapes(n=5, seed=61)
    sim = batch(shp, 1)
    sim.robotStep(300)
    com = sim.robotCom()
    pos = sim.state()["pos"]
    rho, L = 1000.0, 0.005
    mass = float(np.float32(L * L * L * rho))
    for r in range(sim.robotCount()):
        f, c = sim.robotVoxels(r)
        m = np.full(c, mass)  # every material here has rho = 1000
        ref = (pos[f:f + c] * m[:, None]).sum(0) / m.sum()
        assert np.allclose(com[r], ref, rtol=0, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_partial_state_reads_equal_the_full_read(precision):
    """After steps, a read of some fields copies only those device arrays (positions() for the
    whole population, orientations of a range); a read of every field refreshes the host
    mirror.  Both give the same bits."""
    sim = batch(shapes(n=4, seed=71), precision, collisions=True)
    sim.robotStep(200)
    n = sim.voxelCount()
    pos = sim.positions()
    q = np.zeros((n - 7, 4))
    sim.abi.voxel_get_state(sim.handle, 7, n - 7, None, _dptr(q), None, None, None)
    mid = sim.positions(n // 3, n // 2)
    full = sim.state()
    assert pos.tobytes() == full["pos"].tobytes()
    assert q.tobytes() == full["orient"][7:].tobytes()
    assert mid.tobytes() == full["pos"][n // 3:n // 3 + n // 2].tobytes()
    assert sim.positions().tobytes() == full["pos"].tobytes()  # from the refreshed mirror


@pytest.mark.gpu
def test_robots_with_self_collision_match_robots_alone():
    """Robots share lattice coordinates (they overlap in space) but only self-contact."""
    shp = [(np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [2, 1, 0], [2, 2, 0], [1, 2, 0], [0, 2, 0],
                      [0, 2, 1], [0, 2, 2], [1, 2, 2], [2, 2, 2], [2, 1, 2], [1, 1, 2]], np.int32),
            np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], np.int32))] * 2
    sim = batch(shp, 0, collisions=True)
    sim.setActuation(False)
    sim.robotStep(400)
    pos = sim.state()["pos"]
    ref = alone(B.product_abi(), *shp[0], collisions=True)
    dt = sim.robotRecommendedTimeStep(0)
    ref.step(dt, 400)
    pr = ref.state()["pos"]
    for r in range(2):
        f, c = sim.robotVoxels(r)
        assert pos[f:f + c].tobytes() == pr.tobytes()
    assert sim.collisionCount() == 2 * ref.collisionCount()


@pytest.mark.gpu
def test_robot_state_info_matches_state_info():
    """stateInfo against the oracle (pinned bit for bit to the reference's stateInfo by the
    golden vectors): the host entry point vxg_state_info equals the correctly rounded oracle's
    stateInfo bit for bit on a stepped cantilever; the per-robot device reduction has the same
    MIN and MAX exactly and TOTAL/AVERAGE within float rounding of the reference's float sum
    (the device sums in FP64); signed sums (stress, strain, pressure) against the FP64 sum of
    the oracle's own per-element values."""
    from voxelyze_amd import workloads as W
    import scenarios as S
    sim = B.maker(B.product_abi())(0.005)
    ref = B.maker(B.oracle_cr_abi())(0.005)
    for x in (sim, ref):
        W.cantilever(x)
        x.step(x.recommendedTimeStep(), 500)
    got, want = S.state_infos(sim), S.state_infos(ref)
    assert got.tobytes() == want.tobytes(), np.argwhere(got != want)[:8]
    ls = ref.link_state()
    signed = {5: ls["stress"].astype(np.float64), 6: ls["strain"].astype(np.float64)}
    for info in range(10):
        for typ in range(4):
            dev = sim.robotStateInfo(info, typ)[0]
            host = float(want[info, typ])
            if typ < 2:
                assert np.float32(dev) == np.float32(host), (info, typ)
            elif info in signed:  # sums with cancellation: against the FP64 sum of the values
                v = signed[info]
                exact = v.sum() / (len(v) if typ == 3 else 1)
                assert dev == pytest.approx(exact, rel=1e-12, abs=1e-30), (info, typ)
            elif info != 8:
                assert dev == pytest.approx(host, rel=1e-5, abs=1e-30), (info, typ)
    # several robots (no actuation: temperatures stay at ambient), each against itself alone
    # on the oracle
    shp = shapes(n=4, seed=81)
    many = batch(shp, 0)
    many.setActuation(False)
    many.robotStep(200)
    dt = [many.robotRecommendedTimeStep(r) for r in range(many.robotCount())]
    for r, (ijk, mat) in enumerate(shp):
        one = alone(B.oracle_cr_abi(), ijk, mat)
        one.step(dt[r], 200)
        want = S.state_infos(one)
        for info in range(10):
            for typ in range(4):
                dev = many.robotStateInfo(info, typ)[r]
                if typ < 2:
                    assert np.float32(dev) == want[info, typ], (r, info, typ)
                elif info not in (5, 6, 8):
                    assert dev == pytest.approx(float(want[info, typ]), rel=1e-5, abs=1e-30), (r, info, typ)
    for info in (0, 2, 5, 6, 7, 9):
        per = many.robotStateInfo(info, 2)
        assert per.sum() == pytest.approx(many.stateInfo(info, 2), rel=1e-5, abs=1e-30), info
    rho, L = 1000.0, 0.005
    counts = np.array([many.robotVoxels(r)[1] for r in range(many.robotCount())])
    assert np.array_equal(many.robotStateInfo(9, 2), counts * float(np.float32(L * L * L * rho)))


@pytest.mark.gpu
def test_strain_energy_needs_the_fp64_step():
    """The fused mixed step keeps no end forces, so STRAIN_ENERGY refuses there instead of
    reporting stale ones."""
    from voxelyze_amd import workloads as W
    sim = B.maker(B.product_abi())(0.005)
    W.cantilever(sim)
    sim.setPrecision(1)
    sim.step(sim.recommendedTimeStep(), 20)
    assert math.isnan(sim.stateInfo(7, 2))
    assert "STRAIN_ENERGY" in sim.lastError()
    assert sim.stateInfo(0, 1) > 0
