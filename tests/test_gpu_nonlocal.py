"""Nonlocal neighbour lists (Assembly::searchNonlocalNeighbors,
assembly.h:117-130) of the GPU cell-grid search against the oracle's O(N^2)
loop, bit for bit: the same particles, each list in id order with the particle
itself included, on jittered hexagonal, simple-cubic and BCC lattices whose
particles carry different nonlocal lengths, and after a second search with
another cutoff ratio (the lists are rebuilt, not appended to)."""
import numpy as np
import pytest

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from oracle import Oracle

pytestmark = pytest.mark.gpu

R = 0.1


def _jittered(gen, seed, amp=0.06):
    xyz = gen()
    rng = np.random.default_rng(seed)
    j = rng.uniform(-amp * R, amp * R, xyz.shape)
    if np.all(xyz[:, 2] == 0):
        j[:, 2] = 0
    return xyz + j


CASES = [
    ("hex", lat.HEXAGON2D, lambda: lat.hex_plate(30, 30, R), [0.22, 0.35, 0.5]),
    ("sc", lat.SC3D, lambda: lat.sc_block(11, 11, 11, R), [0.2, 0.3, 0.45]),
    ("bcc", lat.BCC3D, lambda: lat.bcc_block(7, 7, 7, R), [0.25, 0.31, 0.4]),
]


@pytest.mark.parametrize("name,L,gen,lengths", CASES, ids=[c[0] for c in CASES])
def test_nonlocal_lists_bit_exact_against_oracle(name, L, gen, lengths):
    xyz = _jittered(gen, seed=len(name))
    systems = [lpmg.System(L, R, lat.ELASTIC_DAMAGE, xyz), Oracle(L, R, lat.ELASTIC_DAMAGE, xyz)]
    n = len(xyz)
    # three interleaved subsets with different nonlocal lengths
    # (ParticleElasticDamage::setParticleProperty, particle_elastic_damage.h:175-250)
    for k, nl in enumerate(lengths):
        idx = np.arange(k, n, len(lengths), dtype=np.int32)
        for s in systems:
            s.set_material(P.MAT_DAMAGE, [nl, 1 if L < 2 else 0, 200e3, 0.3, 1e-3, 1.0, 1.0, 1.0], idx=idx)
    for ratio in (1.5, 3.0):
        for s in systems:
            s.search_nonlocal(ratio)
        (gp, gi), (op, oi) = (s.nonlocal_neighbors() for s in systems)
        np.testing.assert_array_equal(gp, op)
        np.testing.assert_array_equal(gi, oi)
        for i in range(0, n, max(1, n // 97)):
            row = gi[gp[i]:gp[i + 1]]
            assert i in row                       # self included (distance 0 < ratio * L)
            assert np.all(np.diff(row) > 0)       # id order, no repeats
        lens = np.diff(gp)
        assert lens.max() > lens.min()            # the lengths differ per particle
    for s in systems:
        s.close()
