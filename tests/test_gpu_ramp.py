"""Prescribed-displacement ramps evaluated on the device (vxg_ext_set_displacement_ramp)
against the oracle driven by the host loop they replace: before every doTimeStep,
setDisplacementAll(base + rate * t) on each driven voxel, t the current time
(CVX_External::setDisplacementAll, src/VX_External.cpp:68-75; the fixed voxel's pose,
src/VX_Voxel.cpp:167-172, 210-228).  Bit for bit on both FP64 paths, with the steps in one
multi-step launch (a replayed graph where there is no self-collision) or one by one."""
import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W


def bits_equal(a, b):
    return all(a[k].tobytes() == b[k].tobytes() for k in a)


def host_loop(sim, voxels, base, rate, dt, nsteps):
    for _ in range(nsteps):
        t = float(np.float32(sim.currentTime()))
        d = tuple(base[c] + rate[c] * t for c in range(3))
        for i in voxels:
            sim.voxel(int(i)).external().setDisplacementAll(d)
        sim.doTimeStep(dt)


def cantilever(make):
    sim = make(0.005)
    info = W.cantilever(sim, nx=12, ny=3, nz=3, total_force=0.0)
    return sim, info["tips"]


def porous(make):
    sim = make(0.005)
    info = W.porous_block(sim, n=12, seed=5, porosity=0.2, rmin=1, rmax=3)
    return sim, info["top"]


def porous_contacts(make):
    """the porous block with self-collision (contact steps in batches on the device)"""
    sim, top = porous(make)
    sim.enableCollisions(True)
    return sim, top


@pytest.mark.gpu
@pytest.mark.parametrize("path", [0, 1], ids=["fused", "split"])
@pytest.mark.parametrize("model", ["cantilever", "porous", "porous-contacts"])
def test_device_ramp_equals_host_loop(model, path):
    print(f"[ramp] {model} path {path}", flush=True)
    build = {"cantilever": cantilever, "porous": porous, "porous-contacts": porous_contacts}[model]
    base, rate = (0.0, 0.0, 1e-5), (0.0, 0.0, -0.05) if model != "cantilever" else (0.02, 0.0, 0.04)
    ref, rv = build(B.maker(B.oracle_cr_abi()))
    dt = ref.recommendedTimeStep()
    host_loop(ref, rv, base, rate, dt, 240)
    for chunks in ([240], [1] * 240):
        dev, dvox = build(B.maker(B.product_abi(), path))
        assert dev.recommendedTimeStep() == dt
        for i in dvox:
            dev.voxel(int(i)).external().setDisplacementRamp(base, rate)
        for c in chunks:
            n, _ = dev.step(dt, c)
            assert n == c
        sa, sb = dev.state(), ref.state()
        assert bits_equal(sa, sb), f"{model}, launches of {chunks[0]}: differs from the host loop"
        la, lb = dev.link_state(), ref.link_state()
        assert bits_equal(la, lb)
        assert dev.currentTime() == ref.currentTime()


@pytest.mark.gpu
def test_mixed_ramp_tracks_fp64_ramp():
    a, av = porous(B.maker(B.product_abi()))
    b, bv = porous(B.maker(B.product_abi()))
    b.setPrecision(1)
    nominal = a.state()["pos"]
    dt = a.recommendedTimeStep()
    for sim, vox in ((a, av), (b, bv)):
        for i in vox:
            sim.voxel(int(i)).external().setDisplacementRamp((0.0, 0.0, 0.0), (0.0, 0.0, -0.05))
        sim.step(dt, 300)
    pa, pb = a.state()["pos"], b.state()["pos"]
    err = np.abs(pb - pa).max() / np.abs(pa - nominal).max()
    print(f"[ramp] mixed vs FP64 after 300 steps: {err:.2e}", flush=True)
    assert err <= 1e-4


@pytest.mark.gpu
def test_any_other_displacement_call_ends_the_ramp():
    sim, vox = cantilever(B.maker(B.product_abi()))
    ref, rvox = cantilever(B.maker(B.oracle_cr_abi()))
    dt = sim.recommendedTimeStep()
    v = sim.voxel(int(vox[0])).external()
    v.setDisplacementRamp((0.0, 0.0, 0.0), (0.0, 0.0, 0.1))
    v.setDisplacementAll((0.0, 0.0, 2e-5))
    ref.voxel(int(rvox[0])).external().setDisplacementAll((0.0, 0.0, 2e-5))
    sim.step(dt, 50)
    ref.step(dt, 50)
    assert bits_equal(sim.state(), ref.state())
