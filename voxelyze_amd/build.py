"""Builds libvxg.so (the HIP library) in-tree for gfx950.

    python -m voxelyze_amd.build

FP64 bit-faithfulness needs: no FMA contraction (-ffp-contract=off, plus the
pragma in every source), IEEE division and square root in both precisions
(-fhip-fp32-correctly-rounded-divide-sqrt, no -ffast-math), and packed FP32
instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libvxg.so")
SOURCES = ["vx_kernels.hip", "vx_fused.hip", "vx_contact.hip", "vx_mixed.hip", "vx_adjoint.hip", "vx_model.cpp",
           "vx_fused_host.cpp", "vx_mixed_host.cpp", "vx_adjoint_host.cpp", "vx_checkpoint.cpp", "vx_linear.hip", "vx_linear_host.cpp", "vx_capi.cpp"]
HEADERS = ["vx_device.h", "vx_math.h", "vx_model.h", "vx_step_fp64.h", "vx_contact_force.h", "vx_adjoint.h", "vx_json.h"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm)")


FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-x", "hip",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose=False):
    cmd = [hipcc()] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libvxg.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


WORK_LIB = os.path.join(HERE, "libvxwork.so")


def build_workloads():
    """libvxwork.so: the native robot-growth generator used by workloads.robot_shape."""
    cmd = ["gcc", "-O2", "-shared", "-fPIC", os.path.join(CSRC, "vx_workloads.c"), "-o", WORK_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed building libvxwork.so")
    os.replace(WORK_LIB + ".tmp", WORK_LIB)
    return WORK_LIB


def build_facade_example():
    """tests/cpp/facade_cantilever: a reference-style C++ program built against the
    header-only facade include/Voxelyze.h and linked with libvxg.so."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_cantilever.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "facade_cantilever")
    cmd = ["g++", "-O2", "-std=c++11", "-I", os.path.join(ROOT, "include"), src, "-o", out,
           "-L", HERE, "-lvxg", "-Wl,-rpath,$ORIGIN/../../voxelyze_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the facade example failed")
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_workloads())
    print(build_facade_example())
