"""Pins the CPU oracle (oracle/lpm_oracle.cpp) before it is trusted as the checker.

LPM-CPP ships no tests, fixtures or golden vectors, and it cannot be built in
this image (its headers need Intel MKL's mkl.h and Boost, both absent), so every
check here comes from mathematics the reference's algorithm must satisfy:
closed forms, invariants, brute force on small inputs and finite differences.
Each test names the plausible mistake it would catch.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import Oracle, unit_cell
from lpmg import lattice as lat
from lpmg import problems as P

R = 0.1


def _sc(n=6):
    return lat.sc_block(n, n, n, R)


def _jitter(xyz, amp, seed=0):
    rng = np.random.default_rng(seed)
    out = xyz + amp * rng.standard_normal(xyz.shape)
    if np.all(xyz[:, 2] == 0):
        out[:, 2] = 0
    return out


CASES = [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(10, 10, R)),
         ("square", lat.SQUARE2D, lambda: lat.sc_block(9, 9, 1, R)),
         ("sc", lat.SC3D, lambda: _sc(6)),
         ("bcc", lat.BCC3D, lambda: lat.bcc_block(4, 4, 4, R)),
         ("fcc", lat.FCC3D, lambda: lat.fcc_block(3, 3, 3, R))]


def test_unit_cell_closed_forms():
    # volume per particle of each lattice (a wrong exponent or constant shows here)
    r = 0.37
    want = {lat.SQUARE2D: 4 * r * r, lat.HEXAGON2D: math.sqrt(3) / 2 * (2 * r) ** 2,
            lat.SC3D: (2 * r) ** 3, lat.FCC3D: (2 * math.sqrt(2) * r) ** 3 / 4,
            lat.BCC3D: (4 * r / math.sqrt(3)) ** 3 / 2}
    cut1 = {lat.SQUARE2D: 2 * math.sqrt(2) * r, lat.HEXAGON2D: 2 * math.sqrt(3) * r,
            lat.SC3D: 2 * math.sqrt(2) * r, lat.FCC3D: 2 * math.sqrt(2) * r,
            lat.BCC3D: 4 * r / math.sqrt(3)}
    for L, v in want.items():
        u = unit_cell(L, r)
        assert u[2] == pytest.approx(v, rel=1e-14)
        assert u[3] == pytest.approx(2 * r, rel=1e-15)
        assert u[4] == pytest.approx(cut1[L], rel=1e-15)
    # second-shell distance of the generated lattices equals the unit cell's cutoff
    assert np.linalg.norm(lat.bcc_block(2, 2, 2, r)[2] - lat.bcc_block(2, 2, 2, r)[0]) == pytest.approx(
        4 * r / math.sqrt(3))


@pytest.mark.parametrize("name,L,gen", CASES)
def test_bonds_match_brute_force(name, L, gen):
    xyz = _jitter(gen(), 0.002 * R)
    a = Oracle(L, R, lat.ELASTIC, xyz)
    b = Oracle(L, R, lat.ELASTIC, xyz, brute=True)
    for x, y in zip(a.bonds(), b.bonds()):
        np.testing.assert_array_equal(x, y)
    # independent numpy search: partner sets per layer, ascending ids
    u = unit_cell(L, R)
    rc0, rc1 = 1.01 * u[3], 1.01 * u[4]
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    bptr, nb0, p2 = a.bonds()
    for i in range(len(xyz)):
        l0 = np.nonzero((d[i] < rc0) & (np.arange(len(xyz)) != i))[0]
        l1 = np.nonzero((d[i] >= rc0) & (d[i] < rc1))[0]
        np.testing.assert_array_equal(p2[bptr[i]:bptr[i] + nb0[i]], l0)
        np.testing.assert_array_equal(p2[bptr[i] + nb0[i]:bptr[i + 1]], l1)


@pytest.mark.parametrize("name,L,gen,nb_full,nb0_full,nconn_full", [
    ("hex", lat.HEXAGON2D, lambda: lat.hex_plate(12, 12, R), 12, 6, 31),
    ("sc", lat.SC3D, lambda: _sc(8), 18, 6, 61),
    ("bcc", lat.BCC3D, lambda: lat.bcc_block(6, 6, 6, R), 14, 8, 41),
    ("fcc", lat.FCC3D, lambda: lat.fcc_block(5, 5, 5, R), 18, 12, 61)])
def test_shell_and_conn_counts(name, L, gen, nb_full, nb0_full, nconn_full):
    # unit_cell.h: nneighbors, nneighbors1 and nneighbors_AFEM + 1 (self)
    o = Oracle(L, R, lat.ELASTIC, gen())
    bptr, nb0, p2 = o.bonds()
    nb = np.diff(bptr)
    assert nb.max() == nb_full and nb0.max() == nb0_full
    cptr, conns, ncl = o.conns()
    assert np.diff(cptr).max() == nconn_full
    # conns = sorted unique 1-/2-hop same-layer partners (independent python sets)
    for i in range(o.N):
        s = set()
        for Lr in (0, 1):
            lo, hi = (bptr[i], bptr[i] + nb0[i]) if Lr == 0 else (bptr[i] + nb0[i], bptr[i + 1])
            for k in p2[lo:hi]:
                s.add(int(k))
                klo, khi = (bptr[k], bptr[k] + nb0[k]) if Lr == 0 else (bptr[k] + nb0[k], bptr[k + 1])
                s.update(int(m) for m in p2[klo:khi])
        want = np.array(sorted(s))
        np.testing.assert_array_equal(conns[cptr[i]:cptr[i + 1]], want)
        assert ncl[i] == np.count_nonzero(want >= i)


@pytest.mark.parametrize("name,L,gen", CASES)
def test_rigid_motion_gives_no_force(name, L, gen):
    # a sign error in the bond stretch or a missing dis0 would load a rigid body
    xyz = gen()
    o = Oracle(L, R, lat.ELASTIC, xyz)
    o.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    th = 0.3
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    o.set("XYZ", xyz @ Rz.T + np.array([1.3, -0.7, 0.0]))
    o.forward()
    pin_rigid = np.abs(o.get("PIN")).max()
    o.set("XYZ", _jitter(xyz, 1e-4 * R, seed=3))
    o.forward()
    pin_def = np.abs(o.get("PIN")).max()
    assert pin_rigid < 1e-7 * pin_def


@pytest.mark.parametrize("name,L,gen", CASES)
def test_force_balance_and_pair_antisymmetry(name, L, gen):
    # internal forces of a free body sum to zero: a wrong opposite-bond index or a
    # one-sided 0.5*(f_ij + f_ji) average breaks it
    xyz = gen()
    o = Oracle(L, R, lat.ELASTIC, xyz)
    o.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    o.set("XYZ", _jitter(xyz, 3e-4 * R, seed=1))
    o.forward()
    pin = o.get("PIN").reshape(-1, 3)
    assert np.abs(pin.sum(0)).max() < 1e-9 * np.abs(pin).sum()


@pytest.mark.parametrize("name,L,gen,ps", [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(18, 18, R), 1),
                                           ("sc", lat.SC3D, lambda: _sc(11), 0)])
def test_affine_strain_interior_stress_is_hooke(name, L, gen, ps):
    # LPM calibration (unit_cell.h el_mapping) reproduces C:eps in the interior of
    # the hexagonal and simple-cubic lattices; a wrong KnTv mapping, Tv layer or
    # stress volume V_m fails it
    E, nu, eps = 200e3, 0.3, 1e-6
    xyz = gen()
    o = Oracle(L, R, lat.ELASTIC, xyz)
    o.set_material(0, [0, ps, E, nu, 1.0])
    if ps:
        C11, C12 = E / (1 - nu * nu), E * nu / (1 - nu * nu)
    else:
        C11, C12 = E * (1 - nu) / (1 + nu) / (1 - 2 * nu), E * nu / (1 + nu) / (1 - 2 * nu)
    C44 = E / 2 / (1 + nu)
    cptr, conns, _ = o.conns()
    bptr, _, _ = o.bonds()
    full = np.diff(bptr) == np.diff(bptr).max()
    deep = np.array([full[conns[cptr[i]:cptr[i + 1]]].all() for i in range(o.N)])
    assert deep.sum() > 4
    for F, want in [(np.diag([1 + eps, 1, 1]), (C11, C12, 0)),
                    (np.array([[1, eps / 2, 0], [eps / 2, 1, 0], [0, 0, 1]]), (0, 0, C44))]:
        o.set("XYZ", xyz @ F.T)
        o.forward()
        s = o.get("STRESS").reshape(-1, 6)[deep] / eps
        assert np.allclose(s[:, 0], want[0], rtol=2e-3, atol=1e-3 * C11)
        assert np.allclose(s[:, 1], want[1], rtol=2e-3, atol=1e-3 * C11)
        assert np.allclose(s[:, 5], want[2], rtol=2e-3, atol=1e-3 * C11)


@pytest.mark.parametrize("name,L,gen", CASES)
def test_tangent_equals_finite_difference_jacobian_at_rest(name, L, gen):
    # At a state with zero bond forces the reference's analytic tangent
    # (derivative.h) is the exact Jacobian of Pin; central differences pin every
    # term: direct bond, common-neighbour and two-hop dilatation couplings.
    xyz = gen()
    if len(xyz) > 130:
        xyz = xyz[:130]
    o = Oracle(L, R, lat.ELASTIC, xyz)
    o.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    o.forward()
    o.assemble()
    d, n = o.dim, o.dim * o.N
    Kp, IK, JK, K = o.csr()
    rows = np.repeat(np.arange(n), np.diff(IK))
    A = sp.csr_matrix((K, (rows, JK - 1)), shape=(n, n)).toarray()
    A = A + np.triu(A, 1).T
    h = 1e-7
    J = np.zeros((n, n))
    for col in range(n):
        i, k = divmod(col, d)
        for sgn in (1, -1):
            x = xyz.copy()
            x[i, k] += sgn * h
            o.set("XYZ", x)
            o.forward()
            J[:, col] += sgn * o.get("PIN").reshape(-1, 3)[:, :d].ravel() / (2 * h)
    assert np.abs(J - A).max() < 1e-7 * np.abs(A).max()
    assert np.abs(A - A.T).max() == 0.0


@pytest.mark.parametrize("name,L,gen", CASES)
def test_csr_layout_closed_form(name, L, gen):
    # IK/JK/K_pointer rebuilt from conns by the reference's layout rules
    # (stiffness.h:54-71, 220-288): row r of particle i holds the upper entries of
    # the diagonal block, then d columns per greater conn, ascending
    o = Oracle(L, R, lat.ELASTIC, gen())
    o.set_material(0, [0, 1, 200e3, 0.3, 1.0])
    o.forward()
    o.assemble()
    Kp, IK, JK, K = o.csr()
    cptr, conns, ncl = o.conns()
    d = o.dim
    ik, jk, kp = [1], [], [0]
    for i in range(o.N):
        greater = conns[cptr[i]:cptr[i + 1]]
        greater = greater[greater > i]
        kp.append(kp[-1] + d * d * ncl[i] - (1 if d == 2 else 3))
        for r in range(d):
            cols = [d * i + s + 1 for s in range(r, d)] + [d * j + s + 1 for j in greater for s in range(d)]
            jk.extend(cols)
            ik.append(ik[-1] + len(cols))
    np.testing.assert_array_equal(Kp, kp)
    np.testing.assert_array_equal(IK, ik)
    np.testing.assert_array_equal(JK, jk)


def _dense_K(o):
    d, n = o.dim, o.dim * o.N
    Kp, IK, JK, K = o.csr()
    rows = np.repeat(np.arange(n), np.diff(IK))
    A = sp.csr_matrix((K, (rows, JK - 1)), shape=(n, n))
    return (A + sp.triu(A, 1).T).tocsc()


def test_dirichlet_elimination_and_cg_against_direct_solve():
    # constrained rows/columns zeroed with ||diag||_2 on the diagonal
    # (stiffness.h:363-438); CG restating MKL's dcg against SuperLU
    prob = P.sc_cube(n=6, steps=1)
    o, g = P.build(Oracle, prob)
    o.forward()
    P.apply_step(o, g, prob.steps[0])
    o.assemble()
    A = _dense_K(o)
    cons = o.get("CONSTRAINT").reshape(-1, 3)[:, :o.dim].ravel() == 1
    Ad = A.toarray()
    diag_norm = np.linalg.norm(np.diag(Ad)[~cons]) if False else None
    assert np.all(Ad[cons][:, ~cons] == 0) and np.all(Ad[~cons][:, cons] == 0)
    dvals = np.diag(Ad)[cons]
    assert np.all(dvals == dvals[0])
    o.forward()
    o.update_rr()
    b = o.get("RESIDUAL")
    x_direct = spla.spsolve(A, b)
    u0 = o.get("XYZ").reshape(-1, 3)[:, :o.dim].ravel().copy()
    st = o.solve_linear(precond=0, rtol=1e-26, atol=0)
    du = o.get("XYZ").reshape(-1, 3)[:, :o.dim].ravel() - u0
    assert np.linalg.norm(du - x_direct) < 1e-9 * np.linalg.norm(x_direct)
    assert st[0] > 0
    assert np.linalg.norm(o.spmv(x_direct) - A @ x_direct) < 1e-13 * np.linalg.norm(A @ x_direct)
    del diag_norm


def test_newton_step_reaches_equilibrium():
    prob = P.sc_cube(n=6, steps=2)
    o, g = P.build(Oracle, prob)
    o.forward()
    for step in prob.steps:
        P.apply_step(o, g, step)
        ok, st = o.solve_step(-1, 30, 1e-11, precond=0, rtol=1e-24, atol=0)
        assert ok
    o.forward()
    nr = o.update_rr()
    pin = o.get("PIN")
    assert nr[0] < 1e-9 * np.abs(pin).max() * math.sqrt(o.N)
    # the top face moved by exactly the prescribed displacement
    u = (o.get("XYZ") - o.get("XYZ_INITIAL")).reshape(-1, 3)
    assert np.allclose(u[prob.groups["top"], 2], 2 * prob.steps[0][3][4], rtol=0, atol=1e-15)


def test_j2_return_mapping_textbook_properties():
    # one J2 update after a converged tension step: the plastic strain increment
    # recovered from the bonds' dLp changes must be deviatoric, parallel to the
    # deviatoric trial stress, of equivalent magnitude dlambda (= the change of
    # the accumulated plastic strain), with dlambda = f/(3G+H) (radial return with
    # linear isotropic hardening)
    prob = P.bcc_j2_bar(nx=3, ny=3, nz=6, steps=8, strain_step=6e-4)
    o, g = P.build(Oracle, prob)
    o.forward()
    for step in prob.steps[:-1]:
        P.apply_step(o, g, step)
        assert o.solve_step(-1, 30, 1e-11, precond=0, rtol=1e-24, atol=0)[0]
    P.apply_step(o, g, prob.steps[-1])
    o.assemble()
    o.newton(30, 1e-11, precond=0, rtol=1e-24, atol=0)
    sv0 = o.get("STATE_VAR").reshape(-1, 7).copy()
    dlp0 = o.get("DLP").copy()
    stress = o.get("STRESS").reshape(-1, 6).copy()
    o.update_state_var()
    sv1 = o.get("STATE_VAR").reshape(-1, 7)
    ddlp = o.get("DLP") - dlp0
    cs = o.get("CS").reshape(-1, 3)
    dis = o.get("DIS")
    bptr, _, _ = o.bonds()
    E, nu, sy, H = 200e3, 0.3, 250.0, 2000.0
    G = E / 2 / (1 + nu)
    checked = 0
    for i in range(o.N):
        dl = sv1[i, 0] - sv0[i, 0]
        s = stress[i].copy()
        m = s[:3].mean()
        s[:3] -= m
        seq = math.sqrt(1.5 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)))
        f = seq - (sy + H * sv0[i, 0])
        assert dl == pytest.approx(max(f, 0) / (3 * G + H), rel=1e-12, abs=1e-300)
        if dl <= 0:
            continue
        b = np.arange(bptr[i], bptr[i + 1])
        n = cs[b]
        A = np.stack([n[:, 0] ** 2, n[:, 1] ** 2, n[:, 2] ** 2, 2 * n[:, 1] * n[:, 2], 2 * n[:, 0] * n[:, 2],
                      2 * n[:, 0] * n[:, 1]], axis=1) * dis[b, None]
        e, _, rank, _ = np.linalg.lstsq(A, ddlp[b], rcond=None)
        if rank < 6:  # too few bond directions (corner particles) to recover the strain
            continue
        assert abs(e[:3].sum()) < 1e-9 * np.abs(e).max()
        assert math.sqrt(2 / 3 * (e[0] ** 2 + e[1] ** 2 + e[2] ** 2 + 2 * (e[3] ** 2 + e[4] ** 2 + e[5] ** 2))) == \
            pytest.approx(dl, rel=1e-9)
        assert np.allclose(e, 1.5 * dl * s / seq, rtol=1e-8, atol=1e-12 * dl)
        checked += 1
    assert checked > 0


def test_bond_breaking_follows_geometric_stretch():
    # ParticleElastic::updateParticleStaticDamage breaks exactly the bonds whose
    # stretch (dis - dis0)/dis reaches the critical value (computed here from the
    # positions alone)
    xyz = lat.hex_plate(10, 10, R)
    o = Oracle(lat.HEXAGON2D, R, lat.ELASTIC, xyz)
    crit = 2e-3
    o.set_material(0, [0, 1, 200e3, 0.3, crit])
    x = xyz.copy()
    x[:, 1] *= 1 + 2.5e-3 * (x[:, 0] > x[:, 0].mean())
    o.set("XYZ", x)
    o.forward()
    assert o.update_static_damage(-1)
    bptr, nb0, p2 = o.bonds()
    i_of = np.repeat(np.arange(o.N), np.diff(bptr))
    d0 = np.linalg.norm(xyz[i_of] - xyz[p2], axis=1)
    d1 = np.linalg.norm(x[i_of] - x[p2], axis=1)
    np.testing.assert_array_equal(o.get("BDAMAGE") == 1, (d1 - d0) / d1 >= crit)
