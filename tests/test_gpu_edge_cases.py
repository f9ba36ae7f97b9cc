"""Edge cases of the HIP path against the oracle.

* rows beyond one wavefront: a hub particle with more than 64 bonds (read from
  a bond list), so rows of more than 64 bonds and more than 64 connections --
  the SpMV's multi-pass rows in FP64 and FP32 and the tangent's path-walking
  kernel;
* a row of exactly 32 bonds (the largest the lane-per-bond forward pass takes);
* a bond listed twice in a bond list (the tangent's scan over path ends);
* the smallest systems: two particles and one bond; a particle without bonds
  is an input error (the reference's stiffness has no diagonal for it), reported
  by the creation call;
* an empty boundary group and a load step that changes nothing (Newton returns
  at once, nothing moves).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from oracle import Oracle

pytestmark = pytest.mark.gpu

R = 0.1
GPU_OPTS = dict(precond=1, rtol=1e-24, atol=0.0)
ORC_OPTS = dict(precond=0, rtol=1e-24, atol=0.0)
MAT = (0, [0, 0, 70e3, 0.33, 1.0])


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _hub_bonds(xyz, n_extra=70):
    """The lattice's bonds plus layer-1 bonds (both directions) from the particle
    nearest the centre to its n_extra nearest particles it is not bonded to."""
    o = Oracle(lat.SC3D, R, lat.ELASTIC, xyz)
    bptr, nb0, p2 = o.bonds()
    n = len(xyz)
    p1 = np.repeat(np.arange(n), np.diff(bptr))
    layer = np.concatenate([np.r_[np.zeros(nb0[i]), np.ones(bptr[i + 1] - bptr[i] - nb0[i])]
                            for i in range(n)])
    c = int(np.argmin(np.linalg.norm(xyz - xyz.mean(0), axis=1)))
    bonded = set(p2[bptr[c]:bptr[c + 1]].tolist()) | {c}
    order = np.argsort(np.linalg.norm(xyz - xyz[c], axis=1), kind="stable")
    far = [int(j) for j in order if int(j) not in bonded][:n_extra]
    ep1 = np.r_[np.full(len(far), c), far]
    ep2 = np.r_[far, np.full(len(far), c)]
    bonds = (np.r_[p1, ep1].astype(np.int32), np.r_[p2, ep2].astype(np.int32),
             np.r_[layer, np.ones(2 * len(far))].astype(np.int32))
    return bonds, c


def _csr(o):
    Kp, IK, JK, K = o.csr()
    n = o.dim * o.N
    rows = np.repeat(np.arange(n), np.diff(IK))
    A = sp.csr_matrix((K, (rows, JK - 1)), shape=(n, n))
    return A + sp.triu(A, 1).T


def test_rows_beyond_one_wavefront_match_oracle():
    xyz = lat.sc_block(6, 6, 6, R)
    bonds, hub = _hub_bonds(xyz)
    g = lpmg.System(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    o = Oracle(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    bptr, _, _ = o.bonds()
    cptr, _, _ = o.conns()
    assert np.diff(bptr).max() > 64 and np.diff(cptr).max() > 64
    for a, b in zip(g.bonds(), o.bonds()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(g.conns(), o.conns()):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(4)
    x = xyz + 5e-4 * R * rng.standard_normal(xyz.shape)
    z = xyz[:, 2]
    cons = np.zeros((o.N, 3))
    cons[z < z.min() + 0.5 * R] = 1
    for s in (g, o):
        s.set_material(*MAT)
        s.set("XYZ", x)
        s.set("CONSTRAINT", cons)
        s.set("PEX", np.tile([0.3, -0.2, 0.5], o.N) * (1 - cons.ravel()))
        s.forward()
    g.assemble(precision=1)
    o.assemble()
    for a, b in zip(g.csr()[:3], o.csr()[:3]):
        np.testing.assert_array_equal(a, b)
    Kg, Ko = g.csr()[3], o.csr()[3]
    assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()
    v = rng.standard_normal(o.dim * o.N)
    bound = abs(_csr(o)) @ np.abs(v)
    y_ref = o.spmv(v)
    assert np.all(np.abs(g.spmv(v, 0) - y_ref) <= 64 * 2.2e-16 * bound)
    assert np.all(np.abs(g.spmv(v, 1) - y_ref) <= 256 * 1.2e-7 * bound)
    for s in (g, o):
        s.update_rr()
    x0 = o.get("XYZ")
    g.solve_linear(**GPU_OPTS)
    o.solve_linear(**ORC_OPTS)
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-10


def test_row_of_exactly_32_bonds_matches_oracle():
    """A hub of 18 + 14 = 32 bonds: the largest row the lane-per-bond forward
    pass takes (32 lanes per particle, every lane holding a bond)."""
    xyz = lat.sc_block(6, 6, 6, R)
    bonds, hub = _hub_bonds(xyz, n_extra=14)
    g = lpmg.System(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    o = Oracle(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    bptr, _, _ = o.bonds()
    assert np.diff(bptr).max() == 32
    rng = np.random.default_rng(6)
    x = xyz + 5e-4 * R * rng.standard_normal(xyz.shape)
    for s in (g, o):
        s.set_material(*MAT)
        s.set("XYZ", x)
        s.forward()
    for f in ("PIN", "STRESS"):
        a, b_ = g.get(f), o.get(f)
        assert np.abs(a - b_).max() <= 1e-13 * np.abs(b_).max(), f
    g.assemble()
    o.assemble()
    Kg, Ko = g.csr()[3], o.csr()[3]
    assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()


def test_repeated_bond_in_a_bond_list_matches_oracle():
    """A bond file may list a bond twice (Assembly::readBond keeps both): two of
    a particle's bonds then end at the same particle, and the tangent walks
    those rows' paths one by one instead of through its per-connection lists."""
    xyz = lat.sc_block(4, 4, 4, R)
    o = Oracle(lat.SC3D, R, lat.ELASTIC, xyz)
    bptr, nb0, p2 = o.bonds()
    n = len(xyz)
    p1 = np.repeat(np.arange(n), np.diff(bptr))
    layer = np.concatenate([np.r_[np.zeros(nb0[i]), np.ones(bptr[i + 1] - bptr[i] - nb0[i])]
                            for i in range(n)])
    c = int(np.argmin(np.linalg.norm(xyz - xyz.mean(0), axis=1)))
    b = bptr[c] + nb0[c]  # c's first layer-1 bond, listed again in both directions
    k = int(p2[b])
    bonds = (np.r_[p1, c, k].astype(np.int32), np.r_[p2, k, c].astype(np.int32),
             np.r_[layer, 1, 1].astype(np.int32))
    g = lpmg.System(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    o = Oracle(lat.SC3D, R, lat.ELASTIC, xyz, bonds=bonds)
    for a, b_ in zip(g.bonds(), o.bonds()):
        np.testing.assert_array_equal(a, b_)
    rng = np.random.default_rng(5)
    x = xyz + 5e-4 * R * rng.standard_normal(xyz.shape)
    for s in (g, o):
        s.set_material(*MAT)
        s.set("XYZ", x)
        s.forward()
    g.assemble()
    o.assemble()
    for a, b_ in zip(g.csr()[:3], o.csr()[:3]):
        np.testing.assert_array_equal(a, b_)
    Kg, Ko = g.csr()[3], o.csr()[3]
    assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()


def test_particle_without_bonds_is_an_input_error():
    xyz = np.r_[lat.sc_block(3, 3, 3, R), [[5.0, 5.0, 5.0]]]
    with pytest.raises(Exception, match="bond"):
        lpmg.System(lat.SC3D, R, lat.ELASTIC, xyz)
    with pytest.raises(Exception):
        lpmg.System(lat.SC3D, R, lat.ELASTIC, np.zeros((1, 3)))


def test_two_particles_one_bond_match_oracle():
    xyz = np.array([[0.0, 0.0, 0.0], [2 * R, 0.0, 0.0]])
    prob = P.Problem("pair", lat.SC3D, R, lat.ELASTIC, xyz, MAT, dict(a=np.array([0]), b=np.array([1])),
                     [[("disp", "a", P.RELATIVE, "x", 0.0), ("disp", "a", P.RELATIVE, "y", 0.0),
                       ("disp", "a", P.RELATIVE, "z", 0.0), ("disp", "b", P.RELATIVE, "y", 0.0),
                       ("disp", "b", P.RELATIVE, "z", 0.0), ("force", "b", P.RELATIVE, 5.0, 0.0, 0.0)]] * 2,
                     tol=1e-11)
    out = []
    for backend, opts in ((lpmg.System, GPU_OPTS), (Oracle, ORC_OPTS)):
        s, gids = P.build(backend, prob)
        s.forward()
        P.solve_problem(s, gids, prob, opts)
        out.append(s.get("XYZ").reshape(-1, 3))
    ug, uo = out[0] - xyz, out[1] - xyz
    assert uo[1, 0] > 0 and np.all(uo[0] == 0)
    assert _relerr(ug, uo) < 1e-10


def test_empty_group_and_zero_step():
    prob = P.hex_plate(nx=10, ny=8, steps=2)
    prob.groups["none"] = np.array([], dtype=np.int64)
    zero = [("disp", "none", P.RELATIVE, "y", 1.0), ("force", "none", P.RELATIVE, 1.0, 1.0, 0.0),
            ("disp", "bottom", P.RELATIVE, "y", 0.0), ("disp", "top", P.RELATIVE, "y", 0.0)]
    res = []
    for backend, opts in ((lpmg.System, GPU_OPTS), (Oracle, ORC_OPTS)):
        s, gids = P.build(backend, prob)
        s.forward()
        P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}), opts)
        before = s.get("XYZ").copy()
        P.apply_step(s, gids, zero)
        ok, st = s.solve_step(prob.undamaged, prob.max_iter, 1e-11, **opts)
        assert ok
        res.append((before, s.get("XYZ"), st))
    (bg, ag, stg), (bo, ao, sto) = res
    np.testing.assert_array_equal(ag, bg)   # nothing moves on the GPU
    assert _relerr(ag - prob.xyz.ravel(), ao - prob.xyz.ravel()) < 1e-8


def test_solver_iteration_timing_runs_exactly_the_asked_iterations():
    # lpmg_time_kernel kinds 5-8, 10, 11 time fixed-length solves of the current
    # residual; with a zero residual there is nothing to time (an error), and
    # the solve leaves the positions where they were
    prob = P.hex_plate(nx=40, ny=30, steps=1)
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    s.assemble(precision=1)
    s.update_rr()
    with pytest.raises(lpmg.LpmgError, match="residual is zero"):
        s.time_kernel(5, reps=8)
    P.apply_step(s, gids, prob.steps[0])
    s.forward()
    s.update_rr()
    x0 = s.get("XYZ").copy()
    for which in (5, 6, 7, 8, 10, 11):
        s.time_kernel(which, reps=16)  # the multigrid hierarchy is built on its first use
        ms, nbytes = s.time_kernel(which, reps=64)
        assert 0 < ms < 50, (which, ms)
        if which <= 6:
            w = 8 if which == 5 else 4
            assert nbytes > 11 * w * s.dim * s.N
        else:
            assert nbytes == 0
    np.testing.assert_array_equal(s.get("XYZ"), x0)
