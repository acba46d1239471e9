"""Records the golden vectors under tests/golden/ from the compiled reference.

Run in the build container (where /root/reference exists):
    make -C oracle ref && python tests/golden/make_golden.py
It loads oracle/_ref/libvxref.so (the reference's own sources plus the C-ABI
driver oracle/ref_driver.cpp), runs every scenario of tests/scenarios.py and
writes one compressed .npz per scenario: inputs are the scenario code itself,
outputs are the reference's probe traces and final voxel/link state.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from voxelyze_amd._abi import bind  # noqa: E402
from voxelyze_amd.api import CVoxelyze  # noqa: E402
import scenarios as S  # noqa: E402


def main(names=None):
    lib = os.path.join(ROOT, "oracle", "_ref", "libvxref.so")
    if not os.path.exists(lib):
        raise SystemExit("build the reference driver first: make -C oracle ref")
    ref = bind(lib, "vxr_", "reference")
    for name in names or S.SCENARIOS:
        t = time.time()
        out = S.SCENARIOS[name](lambda sz: CVoxelyze(sz, ref))
        arrays = {k: np.asarray(v) for k, v in out.items()}
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(f"{name:28s} {len(arrays):4d} arrays  {time.time() - t:.2f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
