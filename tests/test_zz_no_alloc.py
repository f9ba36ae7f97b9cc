"""lpmg_spmv_noalloc, the SpMV entry point that allocates no device memory
(its staging lives in a static device array), against lpmg_spmv on the same
assembled stiffness: the same kernels on the same data, so bit for bit, in
FP64 and in the mixed-precision FP32 path.  Two systems in a row, the second
smaller, so the second call runs over what the first left in the shared array;
and a system above the array's 2^20 unknowns is refused with a message."""
import numpy as np
import pytest

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P

pytestmark = pytest.mark.gpu

MAX_DOFS = 1 << 20


def _assembled(prob):
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.apply_step(s, gids, prob.steps[0])
    s.forward()
    s.assemble(precision=1)   # FP64 blocks and their FP32 copy
    return s


def test_noalloc_spmv_matches_spmv_on_two_calls_second_smaller():
    systems = [_assembled(P.sc_cube(n=16, steps=2)), _assembled(P.hex_plate(nx=30, ny=30, steps=2))]
    assert systems[0].dim * systems[0].N > systems[1].dim * systems[1].N
    rng = np.random.default_rng(11)
    for s in systems:
        n = s.dim * s.N
        print(f"noalloc spmv: dim={s.dim} N={s.N} unknowns={n}", flush=True)
        x = rng.standard_normal(n)
        for precision in (0, 1):
            y_ref = s.spmv(x, precision)
            y = s.spmv_noalloc(x, precision)
            assert np.isfinite(y_ref).all() and np.abs(y_ref).max() > 0
            np.testing.assert_array_equal(y, y_ref)
    for s in systems:
        s.close()


def test_noalloc_spmv_refuses_more_unknowns_than_its_scratch():
    xyz = lat.sc_block(71, 71, 71, 0.1)        # 357,911 particles, 1,073,733 unknowns
    assert 3 * len(xyz) > MAX_DOFS
    s = lpmg.System(lat.SC3D, 0.1, lat.ELASTIC, xyz)
    with pytest.raises(RuntimeError, match="exceed"):
        s.spmv_noalloc(np.zeros(3 * len(xyz)))
    s.close()
