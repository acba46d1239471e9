"""The .vxl.json fixtures and the C++ reader/writer (voxelyze_amd/csrc/vx_json.h) on the host.

* every fixture parses with Python's json once what the reference left open is closed;
* tests/cpp/json_roundtrip.cpp (built here with g++ against vx_json.h) reads every fixture,
  writes it back and reads that again: same values, the written text is valid JSON;
* where the compiled reference is present, it loads its own linear-material files and steps
  them to the recorded state bit for bit (its reader skips stress/strain point lists).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import backends as B
import scenarios as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)


def test_fixtures_are_json():
    from test_gpu_json import read_lenient
    for name in S.JSON_CASES:
        doc = read_lenient(os.path.join(GOLDEN, name + ".vxl.json"))
        assert doc["voxelSize"] > 0 and len(doc["voxels"]) % 4 == 0 and doc["materials"]


@pytest.fixture(scope="module")
def roundtrip_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("json") / "json_roundtrip")
    src = os.path.join(HERE, "cpp", "json_roundtrip.cpp")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "voxelyze_amd", "csrc"), src, "-o", out],
                   check=True)
    return out


@pytest.mark.parametrize("name", sorted(S.JSON_CASES))
def test_cpp_reader_writer_round_trip(name, roundtrip_bin, tmp_path):
    from test_gpu_json import read_lenient
    dst = tmp_path / "out.json"
    subprocess.run([roundtrip_bin, os.path.join(GOLDEN, name + ".vxl.json"), str(dst)], check=True)
    assert json.loads(dst.read_text()) == read_lenient(os.path.join(GOLDEN, name + ".vxl.json"))


def test_cpp_reader_rejects_broken_text(roundtrip_bin, tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('{"voxelSize": 0.001, "materials": [ {"youngsModulus": 1e6,, } ] }')
    r = subprocess.run([roundtrip_bin, str(bad), str(tmp_path / "o.json")], capture_output=True, text=True)
    assert r.returncode != 0 and "parse error" in r.stderr


@pytest.mark.parametrize("name", ["json_cantilever"])
def test_reference_reads_its_own_file(name):
    ref = B.ref_abi()
    if ref is None:
        pytest.skip("compiled reference not present (built only where the reference sources are)")
    sim = B.maker(ref)(0.001)
    assert sim.loadJSON(os.path.join(GOLDEN, name + ".vxl.json"))
    out = S.json_run(sim, name)
    gold = B.load_golden(name)
    assert all(np.asarray(gold[k]).tobytes() == np.asarray(out[k]).tobytes() for k in gold)
