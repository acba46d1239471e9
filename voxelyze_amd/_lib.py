"""Loads the HIP library libvxg.so (built in-tree by voxelyze_amd/build.py).

There is no fallback: without the library, or without a GPU, creating a
simulation raises.
"""
import functools
import os

from ._abi import bind

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvxg.so")


@functools.lru_cache(None)
def abi():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: build it with `python -m voxelyze_amd.build`")
    a = bind(LIB, "vxg_", "hip")
    _check_device(a)
    return a


def _check_device(a):
    h = a.create(0.001)
    if not h:
        raise RuntimeError(a.last_error(None).decode() or "no GPU")
    a.destroy(h)
