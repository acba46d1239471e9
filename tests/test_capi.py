"""The product library loads and exports every entry point include/vxg.h
declares, with the signature table the Python binding uses (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from voxelyze_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "vxg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vxg_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    names = declared()
    assert len(names) >= 50
    assert sorted("vxg_" + k for k in list(_abi.SIGNATURES) + list(_abi.EXTENSIONS)) == names
    assert not set(_abi.SIGNATURES) & set(_abi.EXTENSIONS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB):
        pytest.fail("libvxg.so not built: run python -m voxelyze_amd.build")
    lib = C.CDLL(_lib.LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing


def test_oracle_and_reference_export_the_same_set():
    import backends as B
    for abi, prefix in ((B.oracle_abi(), "vxo_"), (B.ref_abi(), "vxr_")):
        if abi is None:
            continue
        for k in _abi.SIGNATURES:
            assert hasattr(abi.lib, prefix + k)


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    lib = C.CDLL(_lib.LIB)
    lib.vxg_create.restype = C.c_void_p
    assert not lib.vxg_create(C.c_double(0.001))
    with pytest.raises(RuntimeError):
        _lib.abi.__wrapped__()
