"""Workload generators: the native robot growth is the Python definition, row for row."""
import numpy as np
import pytest

from voxelyze_amd import workloads as W


@pytest.mark.parametrize("seed,n,box", [(4, 100, 20), (5, 3000, 20), (6, 8000, 20), (9, 700, 9)])
def test_native_growth_equals_definition(seed, n, box):
    a = W.grow_robot(seed, n, box)
    b = W.grow_robot_native(seed, n, box)
    assert np.array_equal(a, b)
    assert len({tuple(r) for r in a}) == len(a)  # no voxel twice


def test_robot_shapes_are_connected_and_sized():
    for r in range(6):
        ijk, mat = W.robot_shape(11, r, vmin=100, vmax=400)
        assert 100 <= len(ijk) <= 400 and ijk[:, 2].min() == 0
        assert set(np.unique(mat)) <= {0, 1, 2, 3}
        cells = {tuple(p) for p in ijk}
        seen, stack = {tuple(ijk[0])}, [tuple(ijk[0])]
        while stack:  # face connectivity by flood fill
            x, y, z = stack.pop()
            for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                q = (x + d[0], y + d[1], z + d[2])
                if q in cells and q not in seen:
                    seen.add(q)
                    stack.append(q)
        assert len(seen) == len(cells)
