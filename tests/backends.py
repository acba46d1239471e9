"""Loads the libraries that export the vxg C ABI: the product (HIP), the CPU
oracle and, where it was built, the compiled reference."""
import functools
import os
import subprocess

from voxelyze_amd._abi import bind
from voxelyze_amd.api import CVoxelyze

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_CR_SO = os.path.join(ROOT, "oracle", "liboracle_cr.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvxref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


@functools.lru_cache(None)
def oracle_abi():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    return bind(ORACLE_SO, "vxo_", "oracle")


@functools.lru_cache(None)
def oracle_cr_abi():
    """The oracle with correctly rounded sin/cos/acos (what the device evaluates)."""
    if not os.path.exists(ORACLE_CR_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    return bind(ORACLE_CR_SO, "vxo_", "oracle_cr")


@functools.lru_cache(None)
def ref_abi():
    return bind(REF_SO, "vxr_", "reference") if os.path.exists(REF_SO) else None


REF_OMP_SO = os.path.join(ROOT, "oracle", "_ref", "libvxref_omp.so")


@functools.lru_cache(None)
def ref_omp_abi():
    """The reference built with its USE_OMP loops (the CPU baseline bench.py times)."""
    return bind(REF_OMP_SO, "vxr_", "reference_omp") if os.path.exists(REF_OMP_SO) else None


@functools.lru_cache(None)
def product_abi():
    from voxelyze_amd import _lib
    return _lib.abi()


def maker(abi, fp64_path=None):
    """Simulation factory; fp64_path selects the HIP library's FP64 step (0 fused, 1 split)."""
    if fp64_path is None:
        return lambda sz: CVoxelyze(sz, abi)

    def make(sz):
        sim = CVoxelyze(sz, abi)
        sim.setFp64Path(fp64_path)
        return sim
    return make


def load_golden(name):
    import numpy as np
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}
