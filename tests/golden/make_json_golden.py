"""Records the .vxl.json fixtures and their trajectories from the compiled reference.

Run in the build container (where /root/reference exists):
    make -C oracle ref && python tests/golden/make_json_golden.py
For each model of tests/scenarios.py JSON_CASES the reference builds the model, writes it with
CVoxelyze::saveJSON (tests/golden/<name>.vxl.json), applies the environment and steps it; the
final state is <name>.npz.  (The trajectory comes from the model the reference built, not
from reloading the file: the reference's reader does not read stress/strain point lists.)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from voxelyze_amd._abi import bind  # noqa: E402
from voxelyze_amd.api import CVoxelyze  # noqa: E402
import scenarios as S  # noqa: E402


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libvxref.so")
    if not os.path.exists(lib):
        raise SystemExit("build the reference driver first: make -C oracle ref")
    ref = bind(lib, "vxr_", "reference")
    for name in S.JSON_CASES:
        sim = S.json_model(lambda sz: CVoxelyze(sz, ref), name)
        path = os.path.join(HERE, name + ".vxl.json")
        assert sim.saveJSON(path)
        out = S.json_run(sim, name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **{k: np.asarray(v) for k, v in out.items()})
        print(name, os.path.getsize(path), "bytes", len(out), "arrays", flush=True)


if __name__ == "__main__":
    main()
