"""Parity of the HIP path (liblpmg.so through its C ABI) with the CPU oracle.

Integer work (bonds, layers, conns, K_pointer, IK, JK, damage and plastic
flags) must be bit-exact.  Floating point: the forward pass and the tangent are
computed in the reference's summation order, so they are compared at 1e-13;
solutions of full quasi-static steps are compared at the north-star bar,
relative L2 error <= 1e-8 for displacements and forces.  At the benchmark size
(SC 40^3) size-independent properties are checked: bit-exact topology, the
matrix against the oracle's, SpMV linearity and the CG residual.
"""
import numpy as np
import pytest

import lpmg
from lpmg import geometry
from lpmg import lattice as lat
from lpmg import problems as P
from oracle import Oracle

pytestmark = pytest.mark.gpu

R = 0.1
GPU_OPTS = dict(precond=1, rtol=1e-24, atol=0.0)
ORC_OPTS = dict(precond=0, rtol=1e-24, atol=0.0)


def _jitter(xyz, amp, seed=0):
    rng = np.random.default_rng(seed)
    out = xyz + amp * rng.standard_normal(xyz.shape)
    if np.all(xyz[:, 2] == 0):
        out[:, 2] = 0
    return out


LATTICES = [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(11, 9, R)),
            ("square", lat.SQUARE2D, lambda: lat.sc_block(8, 7, 1, R)),
            ("sc", lat.SC3D, lambda: lat.sc_block(7, 6, 5, R)),
            ("bcc", lat.BCC3D, lambda: lat.bcc_block(4, 4, 5, R)),
            ("fcc", lat.FCC3D, lambda: lat.fcc_block(3, 3, 4, R)),
            ("bcc_poly", lat.BCC3D, lambda: geometry.bcc_polycrystal(6, 5, radius=R, seed=11)["xyz"])]


def _pair(L, xyz, ptype=lat.ELASTIC, model=0, params=(0, 1, 200e3, 0.3, 1.0)):
    g = lpmg.System(L, R, ptype, xyz)
    o = Oracle(L, R, ptype, xyz)
    for s in (g, o):
        s.set_material(model, list(params))
    return g, o


def _relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_topology_is_bit_exact(name, L, gen):
    xyz = _jitter(gen(), 0.002 * R)
    g, o = _pair(L, xyz)
    assert (g.N, g.dim, g.nbond, g.nnz, g.nnzb) == (o.N, o.dim, o.nbond, o.nnz, o.nnzb)
    for a, b in zip(g.bonds(), o.bonds()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(g.conns(), o.conns()):
        np.testing.assert_array_equal(a, b)
    for s in (g, o):
        s.forward()
        s.assemble()
    for a, b in zip(g.csr()[:3], o.csr()[:3]):
        np.testing.assert_array_equal(a, b)


def test_bond_file_order_is_kept():
    # Assembly::readBond keeps the file order inside each (particle, layer)
    xyz = lat.hex_plate(8, 8, R)
    o = Oracle(lat.HEXAGON2D, R, lat.ELASTIC, xyz)
    bptr, nb0, p2 = o.bonds()
    p1 = np.repeat(np.arange(o.N), np.diff(bptr)).astype(np.int32)
    layer = np.concatenate([np.r_[np.zeros(nb0[i]), np.ones(bptr[i + 1] - bptr[i] - nb0[i])]
                            for i in range(o.N)]).astype(np.int32)
    perm = np.random.default_rng(5).permutation(len(p1))
    bonds = (p1[perm], p2[perm], layer[perm])
    g = lpmg.System(lat.HEXAGON2D, R, lat.ELASTIC, xyz, bonds=bonds)
    o2 = Oracle(lat.HEXAGON2D, R, lat.ELASTIC, xyz, bonds=bonds)
    for a, b in zip(g.bonds(), o2.bonds()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(g.conns(), o2.conns()):
        np.testing.assert_array_equal(a, b)


def _perturbed_state(g, o, xyz, ptype, seed):
    rng = np.random.default_rng(seed)
    x = _jitter(xyz, 5e-4 * R, seed)
    for s in (g, o):
        s.set("XYZ", x)
    nb = g.nbond
    bd = np.zeros(nb)
    bptr, nb0, p2 = o.bonds()
    i_of = np.repeat(np.arange(o.N), np.diff(bptr))
    broken = rng.random(nb) < 0.03
    # break both directions of a bond
    key = set(zip(i_of[broken], p2[broken]))
    bd[[k for k in range(nb) if (i_of[k], p2[k]) in key or (p2[k], i_of[k]) in key]] = 1.0
    for s in (g, o):
        s.set("BDAMAGE", bd)
    if ptype == lat.J2PLASTICITY:
        dlp = 1e-5 * rng.standard_normal(nb)
        for s in (g, o):
            s.set("DLP", dlp)
    if ptype != lat.ELASTIC:
        dm = np.where(rng.random(o.N) < 0.05, rng.random(o.N) * 0.5, 0.0)
        for s in (g, o):
            s.set("DAMAGE", dm)


FWD_FIELDS = ("DIS", "CS", "DLE", "BSTRAIN", "BFORCE", "BDAMAGE", "PIN", "STRESS", "DAMAGE_VISUAL",
              "CS_SUM", "DLE_TOTAL", "TDLE_TOTAL")


@pytest.mark.parametrize("name,L,gen", LATTICES)
@pytest.mark.parametrize("ptype,model,params", [
    (lat.ELASTIC, 0, (0, 1, 200e3, 0.3, 1.0)),
    (lat.ELASTIC_DAMAGE, 2, (0, 1, 200e3, 0.3, 1e-3, 5.0, 10.0, 0.95)),
    (lat.J2PLASTICITY, 4, (0, 1, 200e3, 0.3, 250.0, 0.0, 2000.0, 0.0, 1.0))])
def test_forward_pass_matches_oracle(name, L, gen, ptype, model, params):
    xyz = gen()
    g, o = _pair(L, xyz, ptype, model, params)
    _perturbed_state(g, o, xyz, ptype, seed=11)
    for s in (g, o):
        s.forward()
    for f in FWD_FIELDS:
        a, b = g.get(f), o.get(f)
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-13 * scale, f
    ng, no = g.update_rr(), o.update_rr()
    np.testing.assert_allclose(ng, no, rtol=1e-12)
    np.testing.assert_allclose(g.get("RESIDUAL"), o.get("RESIDUAL"), rtol=0,
                               atol=1e-13 * np.abs(o.get("RESIDUAL")).max() + 1e-300)


def _csr_close(g, o):
    """Entries within 1e-13 of their row's largest magnitude.  The diagonal of an
    eliminated DOF is ||diag(K)||_2 (stiffness.h:378), a sum over all DOFs whose
    rounding depends on the summation order (the oracle sums serially, the GPU
    by a tree): those entries are compared at 1e-12."""
    Kg, Ko = g.csr()[3], o.csr()[3]
    Kp, IK, JK, _ = o.csr()
    rows = np.repeat(np.arange(len(IK) - 1), np.diff(IK))
    rowmax = np.zeros(len(IK) - 1)
    np.maximum.at(rowmax, rows, np.abs(Ko))
    ok = np.abs(Kg - Ko) <= 1e-13 * rowmax[rows] + 1e-300
    cons = o.get("CONSTRAINT").reshape(-1, 3)[:, :o.dim].ravel() == 1
    diag_pos = IK[:-1] - 1
    elim = diag_pos[cons]
    if elim.size:
        nd_o, nd_g = Ko[elim], Kg[elim]
        assert np.all(nd_o == nd_o[0]) and np.all(nd_g == nd_g[0])
        gap = abs(nd_g[0] - nd_o[0]) / nd_o[0]
        print(f"||diag||_2 relative gap GPU vs oracle: {gap:.2e}")
        ok[elim] = gap <= 1e-12
    return ok


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_assembly_matches_oracle(name, L, gen):
    xyz = gen()
    g, o = _pair(L, xyz)
    _perturbed_state(g, o, xyz, lat.ELASTIC, seed=7)
    rng = np.random.default_rng(2)
    cons = (rng.random((o.N, 3)) < 0.1).astype(float)
    for s in (g, o):
        s.set("CONSTRAINT", cons)
        s.forward()
        s.assemble()
    assert _csr_close(g, o).all()


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_stiffness_exactly_symmetric_with_one_sided_damage(name, L, gen):
    # Bond damage set per direction (a bond broken on one side only, partial
    # damage differing between a bond and its reverse), constrained DOFs: the
    # assembled blocks match the oracle at 1e-13, and every block below the
    # diagonal is bit for bit the transpose of the one above.  Column a of K is
    # read exactly as K e_a (every other product is an exact zero).
    xyz = _jitter(gen(), 4e-4 * R, seed=5)
    g, o = _pair(L, xyz)
    bptr, _, p2 = o.bonds()
    rng = np.random.default_rng(17)
    nb = len(p2)
    bd = np.where(rng.random(nb) < 0.25, rng.random(nb) * 0.8, 0.0)
    bd[rng.random(nb) < 0.04] = 1.0
    cons = (rng.random((o.N, 3)) < 0.05).astype(float)
    for s in (g, o):
        s.set("BDAMAGE", bd)
        s.set("CONSTRAINT", cons)
        s.forward()
        s.assemble()
    assert _csr_close(g, o).all()
    Kp, IK, JK, K = g.csr()
    import scipy.sparse as sp
    n = g.dim * g.N
    U = sp.csr_matrix((K, (np.repeat(np.arange(n), np.diff(IK)), JK - 1)), shape=(n, n)).toarray()
    for a in rng.choice(n, size=min(n, 24), replace=False):
        x = np.zeros(n)
        x[a] = 1.0
        col = g.spmv(x, 0)
        np.testing.assert_array_equal(col[a + 1:], U[a, a + 1:])
        np.testing.assert_array_equal(col[:a + 1], U[:a + 1, a])


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_spmv_matches_oracle(name, L, gen):
    xyz = gen()
    g, o = _pair(L, xyz)
    _perturbed_state(g, o, xyz, lat.ELASTIC, seed=3)
    for s in (g, o):
        s.forward()
    g.assemble(precision=1)
    o.assemble()
    x = np.random.default_rng(1).standard_normal(o.dim * o.N)
    # elementwise bound: rows of K nearly cancel (translation invariance), so
    # the error is measured against |K||x|
    Kp, IK, JK, K = o.csr()
    import scipy.sparse as sp
    n = o.dim * o.N
    rows = np.repeat(np.arange(n), np.diff(IK))
    A = sp.csr_matrix((K, (rows, JK - 1)), shape=(n, n))
    A = A + sp.triu(A, 1).T
    bound = abs(A) @ np.abs(x)
    y_ref = o.spmv(x)
    assert np.all(np.abs(g.spmv(x, 0) - y_ref) <= 64 * 2.2e-16 * bound)
    assert np.all(np.abs(g.spmv(x, 1) - y_ref) <= 256 * 1.2e-7 * bound)


@pytest.mark.parametrize("name,L,gen", LATTICES)
def test_linear_solve_matches_oracle(name, L, gen):
    xyz = gen()
    g, o = _pair(L, xyz)
    z = xyz[:, o.dim - 1]
    cons = np.zeros((o.N, 3))
    cons[z < z.min() + 0.5 * R] = 1
    for s in (g, o):
        s.set("CONSTRAINT", cons)
        s.set("PEX", np.tile([0.3, -0.2, 0.5], o.N) * (1 - cons.ravel()))
        s.forward()
        s.assemble()
        s.update_rr()
    x0 = o.get("XYZ")
    g.solve_linear(**GPU_OPTS)
    o.solve_linear(**ORC_OPTS)
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-10


def _run_problem(prob, backend, opts, tol=1e-11):
    s, gids = P.build(backend, prob)
    s.forward()
    stats = P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": tol}), opts)
    return s, stats


@pytest.mark.parametrize("prob", [P.sc_cube(n=8, steps=3), P.hex_plate(nx=14, ny=12, steps=3),
                                  P.bcc_j2_bar(nx=3, ny=3, nz=6, steps=10, strain_step=5e-4),
                                  P.bcc_polycrystal(n_cells=7, n_grains=6, steps=3, strain_step=2e-4)],
                         ids=["sc_cube", "hex_plate", "bcc_j2_bar", "bcc_polycrystal_cubic"])
def test_quasi_static_steps_match_oracle(prob):
    g, _ = _run_problem(prob, lpmg.System, GPU_OPTS)
    o, _ = _run_problem(prob, Oracle, ORC_OPTS)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8
    if prob.ptype == lat.J2PLASTICITY:
        ag, ao = g.get("STATE_VAR")[0::7], o.get("STATE_VAR")[0::7]
        np.testing.assert_array_equal(ag > 0, ao > 0)  # plastic flags
        assert (ao > 0).any()
        assert _relerr(ag, ao) < 1e-8
        assert _relerr(g.get("DLP"), o.get("DLP")) < 1e-8


def test_mixed_precision_newton_matches_fp64_oracle():
    prob = P.sc_cube(n=8, steps=3)
    g, _ = _run_problem(prob, lpmg.System, dict(precond=1, rtol=1e-10, atol=0.0, precision=1))
    o, _ = _run_problem(prob, Oracle, ORC_OPTS)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8


def _notched_hex(crit):
    xyz = lat.hex_plate(16, 14, R)
    y = xyz[:, 1]
    bottom = np.nonzero(y < y.min() + 0.5 * R)[0]
    top = np.nonzero(y > y.max() - 0.5 * R)[0]
    step = [("disp", "bottom", P.RELATIVE, "x", 0.0), ("disp", "bottom", P.RELATIVE, "y", 0.0),
            ("disp", "top", P.RELATIVE, "x", 0.0), ("disp", "top", P.RELATIVE, "y", 0.004)]
    prob = P.Problem("notched", lat.HEXAGON2D, R, lat.ELASTIC, xyz, (0, [0.0, 1, 200e3, 0.3, crit]),
                     dict(bottom=bottom, top=top), [list(step) for _ in range(4)])
    return prob


def test_bond_breaking_flags_match_oracle():
    prob = _notched_hex(crit=1.2e-3)
    # pre-crack: break the bonds crossing x = mid along the lower third
    o0 = Oracle(prob.lattice, R, prob.ptype, prob.xyz)
    bptr, nb0, p2 = o0.bonds()
    i_of = np.repeat(np.arange(o0.N), np.diff(bptr))
    xm = prob.xyz[:, 0].mean()
    ymax = prob.xyz[:, 1].min() + 0.4 * np.ptp(prob.xyz[:, 1])
    xi, xj = prob.xyz[i_of], prob.xyz[p2]
    cut = ((xi[:, 0] - xm) * (xj[:, 0] - xm) < 0) & (np.minimum(xi[:, 1], xj[:, 1]) < ymax)
    bd = cut.astype(float)
    res = []
    for backend, opts in ((lpmg.System, GPU_OPTS), (Oracle, ORC_OPTS)):
        s, gids = P.build(backend, prob)
        s.set("BDAMAGE", bd)
        s.store_state()
        s.forward()
        P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}), opts)
        res.append(s)
    g, o = res
    bg, bo = g.get("BDAMAGE"), o.get("BDAMAGE")
    np.testing.assert_array_equal(bg, bo)
    assert bo.sum() > bd.sum()  # new bonds broke
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8


def test_elastic_damage_with_nonlocal_averaging_matches_oracle():
    xyz = lat.hex_plate(12, 12, R)
    y = xyz[:, 1]
    bottom = np.nonzero(y < y.min() + 0.5 * R)[0]
    top = np.nonzero(y > y.max() - 0.5 * R)[0]
    step = [("disp", "bottom", P.RELATIVE, "x", 0.0), ("disp", "bottom", P.RELATIVE, "y", 0.0),
            ("disp", "top", P.RELATIVE, "y", 0.003)]
    prob = P.Problem("damage", lat.HEXAGON2D, R, lat.ELASTIC_DAMAGE, xyz,
                     (2, [3 * R, 1, 200e3, 0.3, 2e-3, 20.0, 10.0, 0.9]),
                     dict(bottom=bottom, top=top), [list(step) for _ in range(3)], tol=1e-11)
    res = []
    for backend, opts in ((lpmg.System, GPU_OPTS), (Oracle, ORC_OPTS)):
        s, gids = P.build(backend, prob)
        s.search_nonlocal(1.5)
        s.forward()
        P.solve_problem(s, gids, prob, opts)
        res.append(s)
    g, o = res
    dg, do = g.get("DAMAGE"), o.get("DAMAGE")
    assert (do > 0).any()
    np.testing.assert_array_equal(dg >= 1, do >= 1)
    assert np.abs(dg - do).max() < 1e-8
    assert _relerr(g.get("XYZ") - o.get("XYZ_INITIAL"), o.get("XYZ") - o.get("XYZ_INITIAL")) < 1e-8


def test_restart_and_rerun_is_bit_identical():
    # reductions are ordered (last-workgroup sums), so a re-run on the same device
    # reproduces every bit; lpmg_restart returns to the initial configuration
    prob = P.sc_cube(n=8, steps=3)
    s, gids = P.build(lpmg.System, prob)
    outs = []
    for _ in range(2):
        s.restart()
        P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}), GPU_OPTS)
        outs.append((s.get("XYZ"), s.get("PIN")))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_graph_replays_match_eager_launches():
    # The CG iterations run as a captured HIP graph, replayed chunk after chunk
    # with eager kernels (forward pass, assembly) and host syncs in between; the
    # profiling mode launches the same kernels eagerly.  Every replay must give
    # the eager result bit for bit (the kernels are deterministic).
    prob = P.sc_cube(n=8, steps=3)
    s, gids = P.build(lpmg.System, prob)
    res = []
    for prof in (False, True, False):
        s.restart()
        s.profile(prof)
        P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": 1e-11}), GPU_OPTS)
        s.profile(False)
        res.append((s.get("XYZ"), s.get("PIN")))
    for a, b in res[1:]:
        np.testing.assert_array_equal(res[0][0], a)
        np.testing.assert_array_equal(res[0][1], b)


def test_benchmark_size_properties():
    # SC 40^3 (the FP64 drop-in benchmark size): bit-exact topology and CSR
    # indices, the stiffness against the oracle's, SpMV linearity and the CG residual
    prob = P.sc_cube(n=40, steps=20)
    g, gg = P.build(lpmg.System, prob)
    o, og = P.build(Oracle, prob)
    for a, b in zip(g.bonds(), o.bonds()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(g.conns(), o.conns()):
        np.testing.assert_array_equal(a, b)
    for s, gid in ((g, gg), (o, og)):
        s.forward()
        P.apply_step(s, gid, prob.steps[0])
        s.assemble()
    cg, co = g.csr(), o.csr()
    for a, b in zip(cg[:3], co[:3]):
        np.testing.assert_array_equal(a, b)
    assert _csr_close(g, o).all()
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(g.dim * g.N), rng.standard_normal(g.dim * g.N)
    Ax, Ay, Axy = g.spmv(x), g.spmv(y), g.spmv(2.0 * x - 3.0 * y)
    assert np.linalg.norm(Axy - (2 * Ax - 3 * Ay)) < 1e-13 * np.linalg.norm(Axy)
    Kp, IK, JK, K = co
    import scipy.sparse as sp
    n = g.dim * g.N
    A = sp.csr_matrix((K, (np.repeat(np.arange(n), np.diff(IK)), JK - 1)), shape=(n, n))
    A = A + sp.triu(A, 1).T
    assert np.all(np.abs(Ax - o.spmv(x)) <= 64 * 2.2e-16 * (abs(A) @ np.abs(x)))
    g.forward()
    g.update_rr()
    r = g.get("RESIDUAL")
    x0 = g.get("XYZ")
    st = g.solve_linear(precond=1, rtol=1e-20, atol=0.0)
    du = (g.get("XYZ") - x0).reshape(-1, 3).ravel()
    res = np.linalg.norm(o.spmv(du) - r) / np.linalg.norm(r)
    assert res < 1e-9, (res, st)
