"""The .vxl.json reader and writer (CVoxelyze::loadJSON/saveJSON, src/Voxelyze.cpp:61-241).

Fixtures: tests/golden/<name>.vxl.json written by the compiled reference's saveJSON and
<name>.npz, the state the reference reaches after stepping the model it wrote
(tests/golden/make_json_golden.py).  The json_open fixture is a model without externals,
which the reference writes without closing the top-level object.
"""
import json
import os

import numpy as np
import pytest

import backends as B
import scenarios as S
from test_gpu_parity import bits_equal, rel_disp_error

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(S.JSON_CASES)


def fixture(name):
    return os.path.join(GOLDEN, name + ".vxl.json")


def read_lenient(path):
    """json.loads of a reference file, closing what it left open at the end"""
    text = open(path).read().rstrip()
    for tail in ("", "}", "]}", "}]}"):
        try:
            return json.loads(text + tail)
        except json.JSONDecodeError:
            continue
    raise AssertionError(f"{path}: not JSON")


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_loaded_model_steps_like_the_reference(name):
    """Load the reference's file, step: the correctly rounded oracle's bits for the model the
    reference built, and the reference's own trajectory within the trigonometric tolerance."""
    sim = B.maker(B.product_abi())(0.001)
    assert sim.loadJSON(fixture(name))
    out = S.json_run(sim, name)
    ref = S.json_run(S.json_model(B.maker(B.oracle_cr_abi()), name), name)
    bad = [k for k in ref if not bits_equal(ref[k], out[k])]
    assert not bad, f"{name}: differs from the oracle in {bad[:8]}"
    gold = B.load_golden(name)
    bad = [k for k in gold if not bits_equal(gold[k], out[k])]
    err = rel_disp_error(gold, out)
    print(f"[json] {name}: {len(gold) - len(bad)}/{len(gold)} arrays bit-exact vs the reference, rel disp err {err:.2e}")
    assert err <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_save_load_round_trip(name, tmp_path):
    """saveJSON writes the reference's document (closed), and reading it back gives the same
    model: the same bits after the same steps."""
    a = B.maker(B.product_abi())(0.001)
    assert a.loadJSON(fixture(name))
    p = tmp_path / (name + ".vxl.json")
    assert a.saveJSON(p)
    mine, theirs = json.loads(p.read_text()), read_lenient(fixture(name))
    assert mine == theirs
    b = B.maker(B.product_abi())(0.001)
    assert b.loadJSON(p)
    sa, sb = S.json_run(a, name), S.json_run(b, name)
    assert all(bits_equal(sa[k], sb[k]) for k in sa)


@pytest.mark.gpu
def test_load_replaces_the_model_and_keeps_settings():
    sim = B.maker(B.product_abi())(0.001)
    m = sim.addMaterial(1e6, 1e3)
    for i in range(4):
        sim.setVoxel(m, i, 0, 0)
    sim.setFp64Path(1)
    assert sim.loadJSON(fixture("json_mixed"))
    assert sim.fp64Path() == 1
    doc = read_lenient(fixture("json_mixed"))
    assert sim.voxelCount() == len(doc["voxels"]) // 4
    assert sim.voxelSize() == doc["voxelSize"]
    assert sim.gravity() == 0.0  # the reader applies no environment key, as the reference's
    with pytest.raises(RuntimeError):
        sim.loadJSON(os.path.join(GOLDEN, "no_such_file.vxl.json"))
