"""vxg_robot_com_noalloc: the centre-of-mass reduction staged in the library's static device
array instead of a library allocation, for callers that cannot allocate.  Checked against
vxg_robot_com bit for bit on two calls in a row, the second on a smaller batch, and on its
refusals (too many robots, state not resident, pending edits)."""
import numpy as np
import pytest

import backends as B
from voxelyze_amd import workloads as W


def population(n, seed):
    sim = B.maker(B.product_abi())(0.005)
    W.robot_population(sim, W.robot_shapes(seed, n, vmin=30, vmax=150, box=7), 39.0, 2.0)
    return sim


@pytest.mark.gpu
def test_noalloc_com_equals_robot_com_on_two_calls():
    print("[no-alloc] 6 robots then 2 robots", flush=True)
    big = population(6, 31)
    big.robotStep(120)
    got = big.robotComNoAlloc()
    want = big.robotCom()
    assert got.shape == (6, 3) and got.tobytes() == want.tobytes()
    small = population(2, 32)
    small.robotStep(80)
    got2 = small.robotComNoAlloc()
    want2 = small.robotCom()
    assert got2.shape == (2, 3) and got2.tobytes() == want2.tobytes()
    assert not np.array_equal(got[:2], got2)
    big.robotStep(10)
    assert big.robotComNoAlloc().tobytes() == big.robotCom().tobytes()


@pytest.mark.gpu
def test_noalloc_com_refuses_what_it_cannot_do_without_allocating():
    sim = population(2, 33)
    with pytest.raises(RuntimeError, match="resident"):
        sim.robotComNoAlloc()
    sim.robotStep(5)
    sim.robotComNoAlloc()
    m = sim.material(0)
    sim.setVoxel(m, 40, 40, 0)  # pending edit of the newest robot
    with pytest.raises(RuntimeError, match="pending"):
        sim.robotComNoAlloc()
    many = B.maker(B.product_abi())(0.005)
    for _ in range(65536):
        many.addRobot()
    assert many.robotCount() == 65537
    with pytest.raises(RuntimeError, match="65536"):
        many.robotComNoAlloc()
