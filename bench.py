"""Throughput of the explicit step (CVoxelyze::doTimeStep) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
                    [--fp64-path fused|split] [--workload lattice|cantilever]
                    [--no-cpu-baseline] [--no-compare]

Default workload: the 128x128x64-site Bernoulli(0.9) lattice of two linear
materials (E 1 MPa / 10 MPa) dropped under gravity onto a friction floor,
~943k voxels and ~2.52M links, 5 mm pitch, dt = recommendedTimeStep().  A step
is one doTimeStep over the whole lattice; state stays resident in HBM.

Prints one JSON line: voxel-steps/s of the selected precision (FP64 plain by
default, run as one fused kernel per step), the roofline of its dominant kernel
(algorithmic bytes per launch / HIP-event-timed launch duration, against the
8 TB/s HBM peak), the other FP64 path (a link and a voxel kernel per step) with a
bitwise comparison of the two, the other precision measured the same way on an
identical copy of the workload with the relative displacement error between the
two after the same steps, and, on N=1, the CPU oracle timed on this host's cores
on a bounded sample of the workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-steps/s at 1M voxels (FP64 plain vs mixed), %HBM peak, adjoint/fwd"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes (per element per step) of each kernel, FP64 plain:
#   link kernel : 288 B per link  (endpoints 8, meta+flags 8, local pose 72 R + 72 W, rest length 8,
#                                  strain state 8 R + 16 W, end forces 96 W) + 56 B per voxel (pose gather)
#   voxel kernel: 244 B per voxel (state 104 R + 104 W, neighbour table 24, metadata 12) + 96 B per link
#                                  (end-force gather)
#   fused FP64 step: the fixed FP64 byte model of the whole step, 300 B per voxel + 384 B per link (the
#                 two kernels above together), which the fused kernel moves less of (no force hand-off)
#   mixed step  : the fixed mixed-precision byte model of the path, 300 B per voxel (pose gather 56,
#                 state 104 R + 104 W, neighbour table 24, metadata 12) + 212 B per link (FP32 local
#                 state 36 R + 36 W, rest length 4, end forces 48 W + 48 R, meta 40), against which the
#                 fused kernel (which has no force hand-off) is also reported
ALG_BYTES = {
    "link": lambda V, L: 288.0 * L + 56.0 * V,
    "voxel": lambda V, L: 244.0 * V + 96.0 * L,
    "step_mixed": lambda V, L: 300.0 * V + 212.0 * L,
    "step_fp64": lambda V, L: 300.0 * V + 384.0 * L,
    # reverse pass of the mixed step, actuated (per voxel: its partials, six adjoint slots, the
    # recorded state and link sums, flags, temperature, indices, gradients; per link: the
    # positive end's pose, adjoint sums and indices, the link's recorded FP32 pose and flags,
    # its pose adjoint read and written; every voxel's six slots are written once)
    "adj_voxels": lambda V, L: 920.0 * V,
    "adj_links": lambda V, L: 576.0 * V + 304.0 * L,
}
# profile classes of vxg_profile_read
PROFILE_CLASSES = ["link", "voxel", "contact", "step_mixed", "step_fp64", "adj_voxels", "adj_links", "adj_contacts"]
KERNEL_NAMES = {"link": "k_links_fp64", "voxel": "k_voxels_fp64", "step_mixed": "k_step_mixed",
                "step_fp64": "k_step_fp64", "adj_voxels": "k_adj_voxels", "adj_links": "k_adj_links"}
FP64_PATHS = {"fused": 0, "split": 1}


def build_workload(sim, workload):
    from voxelyze_amd import workloads as W
    if workload == "lattice":
        info = W.drop_lattice(sim, 128, 128, 64, seed=2)
        return {"workload": "drop lattice 128x128x64 sites p=0.9, 2 linear materials, gravity+floor friction",
                "sites": info["n_sites"]}
    if workload == "cantilever":
        W.cantilever(sim)
        return {"workload": "cantilever 20x4x4, E=1MPa, fixed face + tip load"}
    raise SystemExit(f"unknown workload {workload}")


def traffic_from_profiles(kernel, workload):
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(workload, budget_s=15.0, max_steps=2000):
    """The reference's own multi-core CPU path on the same workload, timed on this host: its
    sources compiled with USE_OMP (oracle/_ref/libvxref_omp.so: the OpenMP link and voxel
    loops of src/Voxelyze.cpp:260-279), driven through oracle/ref_driver.cpp.  Where that
    library was not built, the oracle (oracle/vx_oracle.c, OpenMP over the same loops)."""
    from voxelyze_amd._abi import bind
    from voxelyze_amd.api import CVoxelyze
    ref = os.path.join(ROOT, "oracle", "_ref", "libvxref_omp.so")
    port = os.path.join(ROOT, "oracle", "liboracle.so")
    if os.path.exists(ref):
        abi, kind, what = bind(ref, "vxr_", "reference"), "reference", "the reference built with USE_OMP"
    elif os.path.exists(port):
        abi, kind, what = bind(port, "vxo_", "oracle"), "port", "oracle/vx_oracle.c with OpenMP"
    else:
        return None
    sim = CVoxelyze(0.005, abi)
    build_workload(sim, workload)
    dt = sim.recommendedTimeStep()
    sim.step(dt, 2)
    n, t0 = 0, time.perf_counter()
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        sim.step(dt, 5)
        n += 5
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    V = sim.voxelCount()
    return {"value": V * n / el, "unit": "voxel-steps/s", "cores": threads, "kind": kind, "cpu_model": cpu_model(),
            "sample": f"{what}, {n} steps after 2 warm-up steps of the same workload ({V} voxels), {el:.1f} s"}


# the library's launch-configuration kernel ids
KERNEL_IDS = {"k_step_mixed": 0, "k_links_fp64": 1, "k_voxels_fp64": 2, "k_step_fp64": 3}


def parse_launch(text):
    """--launch "k_step_fp64=128:1,k_links_fp64=64:0" -> {"k_step_fp64": (128, 1), ...}"""
    out = {}
    for item in filter(None, (text or "").split(",")):
        name, cfg = item.split("=")
        if name not in KERNEL_IDS:
            raise SystemExit(f"bench: unknown kernel {name!r} in --launch (one of {sorted(KERNEL_IDS)})")
        blk, xcd = cfg.split(":")
        out[name] = (int(blk), int(xcd))
    return out


def tuned_launch(precision, workload, fp64_path):
    """The launch configurations the autotuner settles on for this model size, found on a
    scratch copy of the workload (the tuner times every candidate on its first steps), so that
    the measured copy starts settled and its warm-up and timed steps are plain steps."""
    import ctypes as C
    from voxelyze_amd.api import CVoxelyze, FP64_PLAIN, MIXED
    sim = CVoxelyze(0.005)
    build_workload(sim, workload)
    sim.setPrecision(MIXED if precision == "mixed" else FP64_PLAIN)
    sim.setFp64Path(FP64_PATHS[fp64_path])
    dt = sim.recommendedTimeStep()
    for _ in range(4):
        sim.step(dt, 10)
    out = {}
    for kname, kk in KERNEL_IDS.items():
        b, x, st = C.c_int(), C.c_int(), C.c_int()
        sim.abi.launch_config(sim.handle, kk, C.byref(b), C.byref(x), C.byref(st))
        if st.value:
            out[kname] = (b.value, x.value)
    return out


def chunked(sim, dt, n, chunk):
    """n steps as launches of `chunk` steps (the last one shorter); returns steps done and the
    first diverging step or -1"""
    done = 0
    while done < n:
        k = min(chunk, n - done)
        d, at = sim.step(dt, k)
        done += d
        if d != k:
            return done, at
    return done, -1


def measure(precision, workload, steps, warmup, fp64_path="fused", launch=None, prof_steps=40):
    """Build the workload with the autotuner's launch shapes pinned (or the given ones), step
    it `warmup` times, then time exactly `steps` steps on the library's production path: one
    vxg_step launch of `steps` steps, recorded as a HIP graph (the recording is inside the timed
    region: on this lattice it costs less than splitting the steps into launches of a length the
    warm-up could record, each of which ends with a synchronisation), no per-kernel events.  Kernel durations come from a
    separate profiled pass of `prof_steps` further steps (HIP events around every kernel).
    Returns the sim (kept alive for comparisons) and the figures."""
    import ctypes as C
    import math
    import torch
    from voxelyze_amd.api import CVoxelyze, FP64_PLAIN, MIXED

    if not launch:
        launch = tuned_launch(precision, workload, fp64_path)
        launch = {k: v for k, v in launch.items() if (k == "k_step_mixed") == (precision == "mixed")}
    t_build = time.perf_counter()
    sim = CVoxelyze(0.005)
    cfg = build_workload(sim, workload)
    sim.setPrecision(MIXED if precision == "mixed" else FP64_PLAIN)
    sim.setFp64Path(FP64_PATHS[fp64_path])
    for kname, (blk, xcd) in (launch or {}).items():
        if not sim.abi.set_launch_config(sim.handle, KERNEL_IDS[kname], int(blk), int(xcd)):
            raise SystemExit(f"bench: bad launch configuration {kname}={blk}:{xcd}")
    V, L = sim.voxelCount(), sim.linkCount()
    dt = sim.recommendedTimeStep()
    nominal = sim.state()["pos"]  # before the first step every voxel sits at its lattice position
    t_build = time.perf_counter() - t_build
    print(f"[bench] {precision}: built {V} voxels, {L} links in {t_build:.1f}s, dt={dt:.7g}", file=sys.stderr,
          flush=True)
    if warmup:
        sim.step(dt, warmup)
    chunk = steps
    torch.cuda.synchronize()
    cap0, rep0 = C.c_int(), C.c_int()
    sim.abi.graph_stats(sim.handle, C.byref(cap0), C.byref(rep0))
    t0 = time.perf_counter()
    done, diverged = chunked(sim, dt, steps, chunk)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cap1, rep1 = C.c_int(), C.c_int()
    sim.abi.graph_stats(sim.handle, C.byref(cap1), C.byref(rep1))
    if done != steps:
        print(f"[bench] {precision}: diverged at step {warmup + diverged}", file=sys.stderr)
        sys.exit(1)
    # kernel durations: a separate pass with HIP events around every kernel (graphs off meanwhile)
    sim.abi.profile(sim.handle, 1)
    sim.abi.profile_read(sim.handle, None, None, 0)
    chunked(sim, dt, prof_steps, chunk)
    torch.cuda.synchronize()
    sim.abi.profile(sim.handle, 0)
    nc = len(PROFILE_CLASSES)
    ms = (C.c_double * nc)()
    cnt = (C.c_int * nc)()
    sim.abi.profile_read(sim.handle, ms, cnt, nc)
    kernels = {}
    for i, name in enumerate(PROFILE_CLASSES):
        if cnt[i] and name in ALG_BYTES:
            avg_s = ms[i] / cnt[i] / 1e3
            b = ALG_BYTES[name](V, L)
            kernels[name] = {"avg_us": avg_s * 1e6, "alg_bytes": b, "gbs": b / avg_s / 1e9, "launches": cnt[i]}
    print(f"[bench] {precision}: {el / steps * 1e3:.4f} ms/step ({rep1.value - rep0.value} graph replays, "
          f"{cap1.value - cap0.value} captures in the timed region), kernels "
          + ", ".join(f"{k} {v['avg_us']:.1f} us {v['gbs']:.0f} GB/s" for k, v in kernels.items()),
          file=sys.stderr, flush=True)
    cfg.update({"voxels": V, "links": L, "dt": dt, "voxel_pitch_m": 0.005})
    tuned = {}
    for kname, kk in KERNEL_IDS.items():
        b, x, st = C.c_int(), C.c_int(), C.c_int()
        sim.abi.launch_config(sim.handle, kk, C.byref(b), C.byref(x), C.byref(st))
        if st.value:
            tuned[kname] = {"block": b.value, "xcd_mapping": bool(x.value)}
    cfg["autotuned_launch"] = tuned
    cfg["timed_launches"] = {"steps_per_launch": chunk, "graph_replays": rep1.value - rep0.value,
                             "graph_captures": cap1.value - cap0.value}
    return sim, {"value": V * steps / el, "el": el, "kernels": kernels, "cfg": cfg, "V": V, "L": L,
                 "nominal": nominal, "after_steps": warmup + steps + prof_steps}


def roofline(kernels, workload):
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"])
    d = kernels[dom]
    traffic = traffic_from_profiles(KERNEL_NAMES[dom], workload)
    return {"bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": KERNEL_NAMES[dom],
            "alg_bytes_per_launch": d["alg_bytes"],
            "kernels": {KERNEL_NAMES[k]: {"avg_us": round(v["avg_us"], 2), "GB/s": round(v["gbs"], 1),
                                          "alg_bytes": v["alg_bytes"]} for k, v in kernels.items()}}


def rel_disp_error(ref_sim, sim, nominal):
    """max_i |pos_i - pos_ref_i| / max_i |pos_ref_i - nominal_i|"""
    pr = ref_sim.state()["pos"]
    po = sim.state()["pos"]
    return float(np.abs(po - pr).max() / max(np.abs(pr - nominal).max(), 1e-300))


def profiled(sim, run):
    """kernel classes -> (avg ms, launches) over run() with HIP events around every kernel"""
    import ctypes as C
    import torch
    sim.abi.profile(sim.handle, 1)
    sim.abi.profile_read(sim.handle, None, None, 0)
    run()
    torch.cuda.synchronize()
    sim.abi.profile(sim.handle, 0)
    nc = len(PROFILE_CLASSES)
    ms, cnt = (C.c_double * nc)(), (C.c_int * nc)()
    sim.abi.profile_read(sim.handle, ms, cnt, nc)
    return {PROFILE_CLASSES[i]: (ms[i] / cnt[i], cnt[i]) for i in range(nc) if cnt[i]}


def bench_robots(n_robots, steps, fp64_value, subset=64):
    """The batched population: n_robots shapes grown to 100-8,000 voxels in a 20^3 box,
    four materials, +-39 degC at 2 Hz actuated on the device, gravity and floor friction,
    every robot with its own recommended dt, mixed precision, one launch per step replayed from
    a graph (100-step launches).  Beside it: the mixed step's error against the FP64 fused step
    on the first `subset` robots over the same steps, with FP64's own response to 1e-9 N on one
    voxel, and the roofline of the step kernel from a separate profiled pass."""
    import torch
    from voxelyze_amd import workloads as W
    from voxelyze_amd.api import CVoxelyze, FP64_PLAIN, MIXED
    t0 = time.perf_counter()
    shapes = W.robot_shapes(4, n_robots)
    sim = CVoxelyze(0.005)
    W.robot_population(sim, shapes)
    sim.setPrecision(MIXED)
    V, L = sim.voxelCount(), sim.linkCount()
    com0 = sim.robotCom()
    chunk = min(100, steps)
    for _ in range(3):
        sim.robotStep(chunk)  # warm-up: the autotuner settles, the graph of a launch is recorded
    torch.cuda.synchronize()
    print(f"[bench] robots: {n_robots} robots, {V} voxels, {L} links, built in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    done = np.zeros(n_robots, np.int64)
    left = steps
    while left > 0:
        k = min(chunk, left)
        done += sim.robotStep(k)
        left -= k
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    travel = sim.robotCom()[:, 0] - com0[:, 0]
    value = V * steps / el
    prof = profiled(sim, lambda: sim.robotStep(20))
    after = 3 * chunk + steps + 20
    avg_ms = prof["step_mixed"][0]
    b = ALG_BYTES["step_mixed"](V, L)
    roof = {"bound": "hbm", "kernel": "k_step_mixed", "avg_us": avg_ms * 1e3, "alg_bytes_per_launch": b,
            "achieved": round(b / (avg_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    print(f"[bench] robots: {el / steps * 1e3:.3f} ms/step, {value:.3e} voxel-steps/s, step kernel "
          f"{avg_ms * 1e3:.0f} us ({roof['frac'] * 100:.1f} % of HBM peak on the mixed byte model)", file=sys.stderr,
          flush=True)
    del sim
    # mixed against FP64 on a subset of the robots, same steps
    pos = {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64_PLAIN), ("fp64_pushed", FP64_PLAIN)):
        sub = CVoxelyze(0.005)
        W.robot_population(sub, shapes[:subset])
        sub.setPrecision(prec)
        nominal = sub.state()["pos"]
        firsts = [sub.robotVoxels(r) for r in range(sub.robotCount())]
        if name == "fp64_pushed":  # 1e-9 N on the first voxel of every robot
            for f, _ in firsts:
                sub.voxel(int(f)).external().setForce(1e-9, 0.0, 0.0)
        left = after
        while left > 0:
            k = min(500, left)
            sub.robotStep(k)
            left -= k
        pos[name] = sub.state()["pos"]
        del sub
    scale = max(np.abs(pos["fp64"] - nominal).max(), 1e-300)

    def per_robot(name):
        d = np.abs(pos[name] - pos["fp64"]).max(axis=1)
        return np.array([d[f:f + c].max() / scale for f, c in firsts])

    pr_mixed, pr_push = per_robot("mixed"), per_robot("fp64_pushed")
    err, resp = float(pr_mixed.max()), float(pr_push.max())
    print(f"[bench] robots: mixed vs FP64 on {subset} robots after {after} steps {err:.3e} (median robot "
          f"{np.median(pr_mixed):.2e}, {(pr_mixed > 1e-4).sum()} robots above 1e-4); FP64's response to 1e-9 N on one "
          f"voxel of every robot {resp:.3e} (median robot {np.median(pr_push):.2e})", file=sys.stderr, flush=True)
    return {"workload": f"{n_robots} robots grown to 100-8000 voxels in a 20^3 box, 4 materials, +-39 degC 2 Hz "
                        "actuation on the device, gravity + floor friction, per-robot dt, mixed precision",
            "voxels": V, "links": L, "steps": steps, "value": value, "unit": "voxel-steps/s",
            "ms_per_step": el / steps * 1e3, "vs_fp64_plain_lattice": value / fp64_value,
            "all_robots_completed": bool((done == steps).all()), "roofline": roof,
            "com_x_travel_m": {"mean": float(travel.mean()), "min": float(travel.min()), "max": float(travel.max()),
                               "median_abs": float(np.median(np.abs(travel))), "after_steps": after - 20},
            "mixed_vs_fp64": {"robots": subset, "after_steps": after, "rel_disp_err": err,
                              "median_robot_rel_disp_err": float(np.median(pr_mixed)),
                              "robots_above_1e-4": int((pr_mixed > 1e-4).sum()),
                              "fp64_response_to_1e-9N_on_every_robot_rel_disp": resp,
                              "fp64_response_median_robot": float(np.median(pr_push))}}


def bench_mixed_error(steps=5000, every=1000):
    """Mixed against FP64 on the drop lattice over the steps its configuration names (it falls
    for ~4,600 steps, then lands on the floor with friction): the relative displacement error,
    the same for the deformation alone (each state minus its mean displacement, which is the
    free fall), and FP64's own response to 1e-9 N on one bottom voxel beside both."""
    from voxelyze_amd.api import CVoxelyze, FP64_PLAIN, MIXED
    sims = {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64_PLAIN), ("fp64_pushed", FP64_PLAIN)):
        sim = CVoxelyze(0.005)
        build_workload(sim, "lattice")
        sim.setPrecision(prec)
        sims[name] = sim
    nominal = sims["fp64"].state()["pos"]
    low = int(np.argmin(nominal[:, 2]))
    sims["fp64_pushed"].voxel(low).external().setForce(1e-9, 0.0, 0.0)
    dt = sims["fp64"].recommendedTimeStep()
    rows = []
    for k in range(every, steps + 1, every):
        for sim in sims.values():
            sim.step(dt, every)
        p = {n: s.state()["pos"] for n, s in sims.items()}
        d = {n: q - nominal for n, q in p.items()}
        dm = {n: x - x.mean(0) for n, x in d.items()}
        scale, dscale = np.abs(d["fp64"]).max(), max(np.abs(dm["fp64"]).max(), 1e-300)
        rows.append({"after_steps": k,
                     "rel_disp_err": float(np.abs(d["mixed"] - d["fp64"]).max() / scale),
                     "fp64_response_to_1e-9N": float(np.abs(d["fp64_pushed"] - d["fp64"]).max() / scale),
                     "deformation_rel_err": float(np.abs(dm["mixed"] - dm["fp64"]).max() / dscale),
                     "deformation_fp64_response_to_1e-9N": float(np.abs(dm["fp64_pushed"] - dm["fp64"]).max() / dscale)})
        print(f"[bench] mixed error after {k} steps: displacement {rows[-1]['rel_disp_err']:.2e} (FP64 response "
              f"{rows[-1]['fp64_response_to_1e-9N']:.2e}), deformation {rows[-1]['deformation_rel_err']:.2e} "
              f"(FP64 response {rows[-1]['deformation_fp64_response_to_1e-9N']:.2e})", file=sys.stderr, flush=True)
    return {"workload": "drop lattice 128x128x64 sites p=0.9 (as the headline), mixed vs FP64 fused, 1e-9 N on the "
                        "lowest voxel for FP64's own response", "rows": rows}


def bench_contacts(steps):
    """Nonlinear material and self-contact: the 64^3-site porous block (bilinear material with
    yield and failure, 20 % spherical voids, bottom face fixed, top face driven down 0.1 pitch
    per 1,000 steps by a ramp the device evaluates before every step, surface self-collision
    on), in mixed precision and in FP64, over the configuration's steps.  Timed: the step
    launches (grid broadphase with the per-step staleness test, contact forces, link and voxel
    updates) in launches of 1,000 steps; kernel times from a separate profiled pass."""
    import ctypes as C
    import torch
    from voxelyze_amd import workloads as W
    from voxelyze_amd.api import CVoxelyze, FP64_PLAIN, MIXED
    out, states = {}, {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64_PLAIN), ("fp64_pushed", FP64_PLAIN)):
        t0 = time.perf_counter()
        sim = CVoxelyze(0.005)
        info = W.porous_block(sim, 64, connected=True)
        sim.setPrecision(prec)
        V, L = sim.voxelCount(), sim.linkCount()
        nominal = sim.state()["pos"]
        if name == "fp64_pushed":  # 1e-9 N on the voxel nearest the middle height: FP64's own sensitivity
            z = nominal[:, 2]
            sim.voxel(int(np.argmin(np.abs(z - z.max() / 2)))).external().setForce(1e-9, 0.0, 0.0)
        dt = sim.recommendedTimeStep()
        W.porous_block_ramp(sim, info["top"], 0.005, dt)  # evaluated on the device before every step
        sim.step(dt, 50)  # warm-up, during which the autotuner settles
        torch.cuda.synchronize()
        print(f"[bench] contacts {name}: porous block {V} voxels, {L} links, built in "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        done = 50
        t0 = time.perf_counter()
        while done - 50 < steps:
            k = min(1000, steps - (done - 50))
            n, at = sim.step(dt, k)
            if n != k:
                raise SystemExit(f"bench: the porous block diverged at step {done + at}")
            done += k
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        timed = done - 50
        prof = profiled(sim, lambda: sim.step(dt, 100))
        done += 100
        states[name] = (sim.state()["pos"], nominal)
        if name == "fp64_pushed":
            del sim
            continue
        kernels = {k: round(v[0] * v[1] / 100, 4) for k, v in prof.items()}
        out[name] = {"value": V * timed / el, "unit": "voxel-steps/s", "ms_per_step": el / timed * 1e3,
                     "kernel_ms_per_step": kernels, "contact_pairs": sim.collisionCount()}
        print(f"[bench] contacts {name}: {el / timed * 1e3:.4f} ms/step, {V * timed / el:.3e} voxel-steps/s, "
              f"{sim.collisionCount()} contact pairs, kernels " + ", ".join(f"{k} {v:.4f} ms" for k, v in kernels.items()),
              file=sys.stderr, flush=True)
        del sim
    (pm, nominal), (pf, _), (pp, _) = states["mixed"], states["fp64"], states["fp64_pushed"]
    scale = max(np.abs(pf - nominal).max(), 1e-300)
    out["mixed_vs_fp64_rel_disp_err"] = float(np.abs(pm - pf).max() / scale)
    # the same lattice in FP64 with 1e-9 N on one voxel: where this is of the order of the
    # mixed error, the trajectory (contacts, link branch switches) is sensitive and no precision
    # reproduces it closer
    out["fp64_response_to_1e-9N_rel_disp"] = float(np.abs(pp - pf).max() / scale)
    print(f"[bench] contacts: mixed vs fp64 rel disp err {out['mixed_vs_fp64_rel_disp_err']:.3e}, fp64 response "
          f"to 1e-9 N on one voxel {out['fp64_response_to_1e-9N_rel_disp']:.3e}", file=sys.stderr, flush=True)
    out.update({"workload": "porous block 64^3 sites, bilinear E=5 MPa -> 0.5 MPa at 0.05 strain, failure at "
                            "0.3 strain (3.75e5 Pa), 20 % voids r=2-6 (largest connected component), bottom fixed, "
                            "top driven -0.1 pitch/1000 steps (device ramp), self-collision", "voxels": V, "links": L,
                "steps": timed,
                "after_steps": done})
    return out


def bench_linear(nx=128, ny=128, nz=64):
    """The static linear solve (CVoxelyze::doLinearSolve) on a filled nx*ny*nz block: bottom
    face fixed, a lateral load spread over the top face.  The second of two identical solves is
    timed (the first also uploads the new model); it includes forming the right-hand side on
    the host, the conjugate-gradient iterations and the write-back."""
    import torch
    from voxelyze_amd.api import CVoxelyze
    t0 = time.perf_counter()
    sim = CVoxelyze(0.005)
    m = sim.addMaterial(1e6, 1000.0)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    ijk = g[np.lexsort((g[:, 0], g[:, 1], g[:, 2]))].astype(np.int32)
    sim.setVoxels(np.zeros(len(ijk), np.int32), ijk)
    for i in np.nonzero(ijk[:, 2] == 0)[0]:
        sim.voxel(int(i)).external().setFixedAll()
    top = np.nonzero(ijk[:, 2] == nz - 1)[0]
    for i in top:
        sim.voxel(int(i)).external().setForce(1e-3 / len(top), 0.0, 0.0)
    V = sim.voxelCount()
    if not sim.doLinearSolve():
        raise SystemExit(f"bench: linear solve failed: {sim.lastError()}")
    torch.cuda.synchronize()
    print(f"[bench] linear: {nx}x{ny}x{nz} block, {V} voxels ({6 * V} degrees of freedom), built and solved "
          f"once in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ok = sim.doLinearSolve()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    info = sim.linearSolveInfo
    print(f"[bench] linear: {el:.3f} s, {info['iterations']} iterations, relative residual "
          f"{info['relative_residual']:.2e}", file=sys.stderr, flush=True)
    return {"workload": f"filled {nx}x{ny}x{nz} block, E=1 MPa, bottom face fixed, 1 mN lateral load on the top face",
            "voxels": V, "dof": 6 * V, "converged": bool(ok), "iterations": info["iterations"],
            "relative_residual": info["relative_residual"], "seconds": el,
            "ms_per_iteration": el / max(info["iterations"], 1) * 1e3}


def bench_adjoint(n_robots, steps):
    """Adjoint cost: the 32x32x16, 70%-fill robots of three materials, mixed precision; the
    forward is `steps` plain steps, the backward is vxg_adjoint_backward (segment recompute
    from checkpoints every 100 steps + reverse kernels) for dL/d(final positions) of the
    CoM x-travel loss."""
    import torch
    from voxelyze_amd import workloads as W
    from voxelyze_amd.api import CVoxelyze, MIXED
    t0 = time.perf_counter()
    sim = CVoxelyze(0.005)
    W.differentiable_batch(sim, n_robots)
    sim.setPrecision(MIXED)
    V = sim.voxelCount()
    sim.robotStep(20)  # warm-up, during which the autotuner settles
    dl = np.zeros((V, 3))
    for r in range(sim.robotCount()):
        f, c = sim.robotVoxels(r)
        dl[f:f + c, 0] = -1.0 / c
    # the adjoint's device buffers (checkpoints of the whole horizon, the segment tape) are
    # allocated by a first rollout and kept for the next, as in a training loop
    sim.resetTime()
    sim.adjointRun(steps, 100)
    sim.resetTime()
    sim.adjointRun(min(steps, 100), 100)
    sim.adjointBackward(dl)
    sim.resetTime()
    sim.robotStep(steps)  # records the forward's step graph outside the timing
    # resetTime rewrites the state on the host; robotCom uploads it (outside the timing) so that
    # both timed forwards start from a resident state, as the backward does
    sim.resetTime()
    sim.robotCom()
    torch.cuda.synchronize()
    print(f"[bench] adjoint: {n_robots} robots, {V} voxels, built in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    sim.robotStep(steps)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    sim.resetTime()
    sim.robotCom()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.adjointRun(steps, 100)
    torch.cuda.synchronize()
    fwd_ckpt = time.perf_counter() - t0
    t0 = time.perf_counter()
    ge, gp = sim.adjointBackward(dl)
    torch.cuda.synchronize()
    bwd = time.perf_counter() - t0
    print(f"[bench] adjoint: forward {fwd:.3f}s, forward with checkpoints {fwd_ckpt:.3f}s, backward {bwd:.3f}s, "
          f"gradient/forward {(fwd_ckpt + bwd) / fwd:.2f}", file=sys.stderr, flush=True)
    # kernel durations of the backward: a separate profiled backward of the same rollout (HIP
    # events around every recompute step and reverse kernel; the link kernel's cost grows as
    # more links take the large-angle branch, so the whole horizon is averaged)
    import ctypes as C
    L = sim.linkCount()
    sim.resetTime()
    sim.adjointRun(steps, 100)
    sim.abi.profile(sim.handle, 1)
    sim.abi.profile_read(sim.handle, None, None, 0)
    sim.adjointBackward(dl)
    torch.cuda.synchronize()
    sim.abi.profile(sim.handle, 0)
    nc = len(PROFILE_CLASSES)
    ms, cnt = (C.c_double * nc)(), (C.c_int * nc)()
    sim.abi.profile_read(sim.handle, ms, cnt, nc)
    sim.adjointFree()
    kernels = {}
    for cls in ("step_mixed", "adj_voxels", "adj_links"):
        i = PROFILE_CLASSES.index(cls)
        if cnt[i]:
            avg_s = ms[i] / cnt[i] / 1e3
            b = ALG_BYTES[cls](V, L)
            kernels[KERNEL_NAMES[cls]] = {"avg_us": round(avg_s * 1e6, 1), "alg_bytes": b,
                                          "GB/s": round(b / avg_s / 1e9, 1), "frac": round(b / avg_s / 1e9 / HBM_PEAK_GBS, 4)}
    print("[bench] adjoint kernels: " + ", ".join(f"{k} {v['avg_us']:.1f} us {v['GB/s']:.0f} GB/s"
                                                   for k, v in kernels.items()), file=sys.stderr, flush=True)
    return {"workload": f"{n_robots} robots 32x32x16 70% fill, 3 materials, +-39 degC 2 Hz actuation, gravity + floor "
                        "friction, mixed precision, checkpoints every 100 steps",
            "voxels": V, "steps": steps, "forward_s": fwd, "forward_with_checkpoints_s": fwd_ckpt, "backward_s": bwd,
            # gradient cost over plain forward cost: (forward with checkpoints + segment
            # recompute + reverse) / forward; the backward call does the recompute and reverse
            "adjoint_over_forward": (fwd_ckpt + bwd) / fwd,
            "backward_over_forward": bwd / fwd,
            "links": L,
            "kernels_per_step": kernels,
            "gradient_finite": bool(np.isfinite(ge).all() and np.isfinite(gp).all())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--fp64-path", default="fused", choices=["fused", "split"],
                    help="FP64 step as one fused kernel per step (default) or a link and a voxel kernel")
    ap.add_argument("--workload", default="lattice", choices=["lattice", "cantilever"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="measure the selected precision only")
    ap.add_argument("--no-robots", action="store_true", help="skip the 4,096-robot batch")
    ap.add_argument("--no-contacts", action="store_true", help="skip the porous block with self-collision")
    ap.add_argument("--no-linear", action="store_true", help="skip the static linear solve")
    ap.add_argument("--contact-steps", type=int, default=20000)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost")
    ap.add_argument("--robots", type=int, default=4096)
    ap.add_argument("--robot-steps", type=int, default=10000)
    ap.add_argument("--adjoint-robots", type=int, default=256)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB -> 1 GiB elementwise copy (known bytes) to scale PMC byte counters")
    ap.add_argument("--launch", default="",
                    help="pin launch configurations instead of autotuning them, e.g. k_step_fp64=128:1 (block:xcd_mapping)")
    a = ap.parse_args()
    launch = parse_launch(a.launch)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    if a.calibrate_copy:
        src = torch.ones(256 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        print("[bench] calibration copy: 3 x (1 GiB read + 1 GiB write)", file=sys.stderr, flush=True)

    sim, r = measure(a.precision, a.workload, a.steps, a.warmup, a.fp64_path, launch)
    cfg = r["cfg"]
    cfg["precision"] = "fp64_plain" if a.precision == "fp64" else "mixed"
    if a.precision == "fp64":
        cfg["fp64_path"] = a.fp64_path
    line = {
        "metric": METRIC, "value": r["value"], "unit": "voxel-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["el"] / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if a.precision == "fp64" else "f32/f64",
        "data": "synthetic (seeded SplitMix64 lattice)", "config": cfg, "roofline": roofline(r["kernels"], a.workload),
    }
    if not a.no_compare and a.precision == "fp64":
        # the other FP64 path on an identical copy: the same bits after the same steps
        other_path = "split" if a.fp64_path == "fused" else "fused"
        sim3, r3 = measure("fp64", a.workload, a.steps, a.warmup, other_path, launch)
        sa, sb = sim.state(), sim3.state()
        same = all(sa[k].tobytes() == sb[k].tobytes() for k in sa)
        la, lb = sim.link_state(), sim3.link_state()
        same = same and all(la[k].tobytes() == lb[k].tobytes() for k in la)
        line["fp64_" + other_path] = {"value": r3["value"], "ms_per_step": r3["el"] / a.steps * 1e3,
                                      "roofline": roofline(r3["kernels"], a.workload),
                                      "autotuned_launch": r3["cfg"]["autotuned_launch"],
                                      "bit_identical_to_" + a.fp64_path: same}
        print(f"[bench] fp64 {other_path}: {r3['el'] / a.steps * 1e3:.4f} ms/step, bit-identical state: {same}",
              file=sys.stderr, flush=True)
        del sim3
    if not a.no_compare:
        other = "mixed" if a.precision == "fp64" else "fp64"
        sim2, r2 = measure(other, a.workload, a.steps, a.warmup, launch=launch)
        ref, test = (sim, sim2) if a.precision == "fp64" else (sim2, sim)
        err = rel_disp_error(ref, test, r["nominal"])
        fp64_v, mixed_v = (r["value"], r2["value"]) if a.precision == "fp64" else (r2["value"], r["value"])
        line[other] = {"value": r2["value"], "ms_per_step": r2["el"] / a.steps * 1e3,
                       "roofline": roofline(r2["kernels"], a.workload),
                       "autotuned_launch": r2["cfg"]["autotuned_launch"]}
        line["mixed_vs_fp64"] = {"speedup": mixed_v / fp64_v, "rel_disp_err": err,
                                 "after_steps": r["after_steps"]}
        print(f"[bench] mixed/fp64 speedup {mixed_v / fp64_v:.2f}x, rel disp err {err:.3e}", file=sys.stderr,
              flush=True)
        del sim2
    if not a.no_compare and a.workload == "lattice":
        line["mixed_error_5000"] = bench_mixed_error()
    if not a.no_robots:
        line["robots"] = bench_robots(a.robots, a.robot_steps, r["value"])
    if not a.no_contacts:
        line["contacts"] = bench_contacts(a.contact_steps)
    if not a.no_linear:
        line["linear"] = bench_linear()
    if not a.no_adjoint:
        line["adjoint"] = bench_adjoint(a.adjoint_robots, a.adjoint_steps)
    if not a.no_cpu_baseline:
        print("[bench] timing the CPU baseline on a bounded sample", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(a.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
