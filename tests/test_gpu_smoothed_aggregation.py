"""Smoothed-aggregation multigrid PCG (solver option precond=4) on the GPU.

No reference counterpart: LPM-CPP solves with MKL's CG or PARDISO
(solver.h:190-354).  The preconditioner changes the CG's path, not the system
it solves, so quasi-static steps must meet the Jacobi CG's bar against the CPU
oracle: displacements and forces within 1e-8 relative L2 at a tight Newton
tolerance, damage flags exactly equal.  The hierarchy itself is pinned by
`sa_pcg_np`, a plain numpy/scipy restatement of the method on the oracle's
matrix: grid-cell aggregates (cell edge 3 lattice spacings, x3.5 per level,
doubled while a level keeps more than 80% of its nodes), each cell split into
the connected components of the intact bonds inside it (coarser levels: of
the bonds between aggregates; cells first, a cell's components by smallest
node; no split that leaves more than 80% of the nodes), the prolongation
P = (I - om_p D^-1 A) T of the piecewise-constant T, Galerkin coarse
operators P^T A P, a dense coarsest inverse, one V-cycle of damped Jacobi
before and after the coarse correction, z = 0 on the eliminated DOFs.  Per
level, lambda is the largest eigenvalue of the 12-step Lanczos tridiagonal of
D^-1 A from a fixed hashed start vector; the smoother's damping is
min(0.9, 1.65 / lambda) and om_p = min(1, 1.5 / lambda).  The level sizes
must agree exactly and the PCG iteration counts to rounding.  A dropped
smoothing term, a transposed P or a wrong sign in it changes the count by far
more than that (the tentative hierarchy takes 2-3x as many iterations on
CT plates).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import lpmg
from lpmg import lattice as lat
from lpmg import problems as P
from lpmg import solver_opts
from lpmg.serve import Request, Server
from oracle import Oracle

pytestmark = pytest.mark.gpu

R = 0.1
SA64 = dict(precond=4, rtol=1e-24, atol=0.0, precision=0)
ORC = dict(precond=0, rtol=1e-24, atol=0.0)
OMEGA = 2.0 / 3.0        # the damping when no estimate is usable
LZ_STEPS, OMEGA_MAX, OMEGA_TARGET, P_TARGET = 12, 0.9, 1.65, 1.5


def _relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _aggregate(pos, h):
    """Nonempty cells of the grid of edge h over pos (n x 3), numbered in
    (z, y, x) cell order; returns (aggregate of each node, count)."""
    cc = np.floor((pos - pos.min(axis=0)) / h).astype(np.int64)
    dims = cc.max(axis=0) + 1
    key = (cc[:, 2] * dims[1] + cc[:, 1]) * dims[0] + cc[:, 0]
    _, agg = np.unique(key, return_inverse=True)
    return agg.ravel(), int(agg.max()) + 1


def intact_edges(o):
    """Node pairs (min, max) of every intact bond slot (|bdamage - 1| >= 1e-6)."""
    bptr, _, p2 = o.bonds()
    bd = o.get("BDAMAGE")[: len(p2)]
    src = np.repeat(np.arange(o.N), np.diff(bptr))
    ok = np.abs(bd - 1.0) >= 1e-6
    return np.stack([np.minimum(src[ok], p2[ok]), np.maximum(src[ok], p2[ok])], 1)


def split_aggregates(agg, nc, e):
    """Each aggregate split into the connected components of the edges inside
    it, in cell order and a cell's components by smallest node; unchanged when
    more than 80% of the nodes would be aggregates."""
    n = len(agg)
    inside = agg[e[:, 0]] == agg[e[:, 1]]
    G = sp.csr_matrix((np.ones(int(inside.sum())), (e[inside, 0], e[inside, 1])), shape=(n, n))
    ncc, lab = connected_components(G, directed=False)
    low = np.full(ncc, n)
    np.minimum.at(low, lab, np.arange(n))
    _, agg2 = np.unique(agg.astype(np.int64) * (n + 1) + low[lab], return_inverse=True)
    nc2 = int(agg2.max()) + 1
    return (agg, nc) if nc2 > 0.8 * n else (agg2.ravel(), nc2)


def coarse_edges(agg, e):
    """The next level's edges: distinct aggregate pairs joined by an edge."""
    a, b = agg[e[:, 0]], agg[e[:, 1]]
    k = a != b
    if not k.any():
        return np.zeros((0, 2), np.int64)
    return np.unique(np.stack([np.minimum(a[k], b[k]), np.maximum(a[k], b[k])], 1), axis=0)


def lanczos_start(n):
    k = np.arange(n, dtype=np.uint64)
    h = ((k * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) >> np.uint64(16)
    return (h & np.uint64(1023)).astype(np.float64) / 1024.0 - 0.5


def dampings(A):
    """(smoother damping min(OMEGA_MAX, OMEGA_TARGET / lambda), prolongator
    smoothing min(1, P_TARGET / lambda)): lambda the largest eigenvalue of the
    LZ_STEPS-step Lanczos tridiagonal of D^-1/2 A D^-1/2 from the hashed start
    vector (a lower estimate of rho(D^-1 A))."""
    s = 1.0 / np.sqrt(A.diagonal())
    v = lanczos_start(A.shape[0])
    v /= np.linalg.norm(v)
    vp, beta, al, be = np.zeros_like(v), 0.0, [], []
    for _ in range(LZ_STEPS):
        w = s * (A @ (s * v))
        a = v @ w
        w = w - a * v - beta * vp
        al.append(a)
        be.append(beta)
        beta = np.linalg.norm(w)
        if not beta > 1e-12 * max(abs(x) for x in al):
            break
        vp, v = v, w / beta
    m = len(al)
    T = np.diag(al) + np.diag(be[1:m], 1) + np.diag(be[1:m], -1)
    lam = np.linalg.eigvalsh(T)[-1]
    if not lam > 0:
        return OMEGA, OMEGA
    return min(OMEGA_MAX, OMEGA_TARGET / lam), min(1.0, P_TARGET / lam)


def sa_hierarchy(A, pos, radius, dim, coarsest=128, edges=None):
    """Levels [(A_l, P_l, dinv_l, smoother damping_l)] and the coarsest matrix;
    with edges (intact bonds as node pairs) the cells are split along them."""
    levels = []
    h = 3.0 * 2.0 * radius
    while A.shape[0] > coarsest:
        n = len(pos)
        agg, nc = _aggregate(pos, h)
        if nc > 0.8 * n:
            h *= 2.0
            continue
        if edges is not None:
            agg, nc = split_aggregates(agg, nc, edges)
            edges = coarse_edges(agg, edges)
        rows = np.arange(n * dim)
        T = sp.csr_matrix((np.ones(n * dim), (rows, np.repeat(agg, dim) * dim + np.tile(np.arange(dim), n))),
                          shape=(n * dim, nc * dim))
        dinv = 1.0 / A.diagonal()
        om, om_p = dampings(A)
        Pm = (T - om_p * (sp.diags(dinv) @ (A @ T))).tocsr()
        levels.append((A, Pm, dinv, om))
        A = (Pm.T @ (A @ Pm)).tocsr()
        cen = np.zeros((nc, 3))
        np.add.at(cen, agg, pos)
        pos = cen / np.bincount(agg)[:, None]
        h *= 3.5
    return levels, A.toarray()


def sa_pcg_np(A, b, mask, pos, radius, dim, rtol, maxit, edges=None):
    """PCG with one smoothed-aggregation V-cycle as M; stops when
    ||r||^2 <= rtol ||r0||^2.  Returns (x, iterations, level sizes)."""
    levels, Ac = sa_hierarchy(A, pos, radius, dim, edges=edges)
    inv = np.linalg.inv(Ac)

    def vcycle(l, rhs):
        if l == len(levels):
            return inv @ rhs
        Al, Pm, dinv, om = levels[l]
        x1 = om * dinv * rhs
        xc = vcycle(l + 1, Pm.T @ (rhs - Al @ x1))
        x = x1 + Pm @ xc
        return x + om * dinv * (rhs - Al @ x)

    def M(r):
        return vcycle(0, r) * mask

    x = np.zeros_like(b)
    r = b.copy()
    thresh = rtol * (r @ r)
    z = M(r)
    p = z.copy()
    rz = r @ z
    sizes = [lv[0].shape[0] for lv in levels] + [Ac.shape[0]]
    for it in range(1, maxit + 1):
        w = A @ p
        alpha = rz / (p @ w)
        x += alpha * p
        r -= alpha * w
        if r @ r <= thresh:
            return x, it, sizes
        z = M(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, maxit, sizes


def _full_matrix(s):
    _, IK, JK, K = s.csr()
    n = s.dim * s.N
    U = sp.csr_matrix((K, (np.repeat(np.arange(n), np.diff(IK)), JK - 1)), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


def _loaded_pair(L, xyz):
    g = lpmg.System(L, R, lat.ELASTIC, xyz)
    o = Oracle(L, R, lat.ELASTIC, xyz)
    dim = o.dim
    z = xyz[:, dim - 1]
    cons = np.zeros((o.N, 3))
    cons[z < z.min() + 0.5 * R] = 1
    for s in (g, o):
        s.set_material(0, [0, 1, 200e3, 0.3, 1.0])
        s.set("CONSTRAINT", cons)
        s.set("PEX", np.tile([0.3, -0.2, 0.5], o.N) * (1 - cons.ravel()))
        s.forward()
        s.assemble()
        s.update_rr()
    return g, o, cons[:, :dim].ravel()


LATTICES = [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(60, 52, R)),
            ("square", lat.SQUARE2D, lambda: lat.sc_block(40, 36, 1, R)),
            ("sc", lat.SC3D, lambda: lat.sc_block(14, 12, 11, R)),
            ("bcc", lat.BCC3D, lambda: lat.bcc_block(12, 12, 13, R))]


@pytest.mark.parametrize("name,L,gen", LATTICES, ids=[x[0] for x in LATTICES])
def test_sa_levels_and_iterations_match_numpy_restatement(name, L, gen):
    xyz = gen()
    g, o, fixed = _loaded_pair(L, xyz)
    A = _full_matrix(o)
    b = o.get("RESIDUAL")
    x0 = g.get("XYZ")
    st = g.solve_linear(precond=4, rtol=1e-16, atol=0.0)
    x_np, it_np, sizes = sa_pcg_np(A, b, 1.0 - fixed, xyz, R, o.dim, 1e-16, 2000, edges=intact_edges(o))
    assert g.multigrid_levels() == sizes
    assert len(sizes) >= 3
    assert abs(int(st[0]) - it_np) <= max(2, it_np // 20), (int(st[0]), it_np)
    du_g = (g.get("XYZ") - x0).reshape(-1, 3)[:, : o.dim].ravel()
    assert _relerr(du_g, x_np) < 1e-6
    # the tentative hierarchy on the same system takes more iterations
    g.set("XYZ", x0)
    st2 = g.solve_linear(precond=2, rtol=1e-16, atol=0.0)
    assert st2[0] > st[0], (st2[0], st[0])


def _crack(o, xyz):
    """Bond damage of a straight crack: every bond crossing the line midway
    between two middle rows, over the left half of the plate, broken."""
    bptr, _, p2 = o.bonds()
    src = np.repeat(np.arange(o.N), np.diff(bptr))
    x, y = xyz[:, 0], xyz[:, 1]
    rows = np.unique(np.round(y, 9))
    ym = 0.5 * (rows[len(rows) // 2] + rows[len(rows) // 2 + 1])
    xm = 0.5 * (x.min() + x.max())
    cut = ((y[src] - ym) * (y[p2] - ym) < 0) & (np.minimum(x[src], x[p2]) < xm)
    return cut.astype(np.float64)


def _level_sizes(xyz, dim, edges, coarsest=128):
    """The hierarchy's level sizes from its aggregation alone."""
    sizes, pos, h, n = [len(xyz) * dim], xyz, 6.0 * R, len(xyz)
    while n * dim > coarsest:
        agg, nc = _aggregate(pos, h)
        if nc > 0.8 * n:
            h *= 2.0
            continue
        if edges is not None:
            agg, nc = split_aggregates(agg, nc, edges)
            edges = coarse_edges(agg, edges)
        cen = np.zeros((nc, 3))
        np.add.at(cen, agg, pos)
        pos, h, n = cen / np.bincount(agg)[:, None], h * 3.5, nc
        sizes.append(nc * dim)
    return sizes


def test_sa_aggregates_split_along_a_crack():
    # cells cut by a crack become one aggregate per side: the level sizes and
    # the PCG path follow the restatement with the split, and the split cells
    # make more aggregates than the cells alone
    xyz = lat.hex_plate(60, 52, R)
    g, o, fixed = _loaded_pair(lat.HEXAGON2D, xyz)
    bd = _crack(o, xyz)
    assert bd.sum() > 50
    for s in (g, o):
        s.set("BDAMAGE", bd)
        s.forward()
        s.assemble()
        s.update_rr()
    A = _full_matrix(o)
    b = o.get("RESIDUAL")
    x0 = g.get("XYZ")
    e = intact_edges(o)
    st = g.solve_linear(precond=4, rtol=1e-16, atol=0.0)
    x_np, it_np, sizes = sa_pcg_np(A, b, 1.0 - fixed, xyz, R, 2, 1e-16, 2000, edges=e)
    assert g.multigrid_levels() == sizes
    assert sizes == _level_sizes(xyz, 2, e)
    unsplit = _level_sizes(xyz, 2, None)
    assert sizes[1] > unsplit[1], (sizes, unsplit)
    assert abs(int(st[0]) - it_np) <= max(2, it_np // 20), (int(st[0]), it_np)
    du_g = (g.get("XYZ") - x0).reshape(-1, 3)[:, :2].ravel()
    assert _relerr(du_g, x_np) < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_sa_hierarchy_follows_bonds_broken_before_a_step(precision):
    # a hierarchy built on the intact plate is kept by plain solves and rebuilt
    # at the first solve of the next load step once bonds broke across its
    # cells; the rebuilt one solves bit for bit as one built on the cracked
    # plate from the start
    xyz = lat.hex_plate(60, 52, R)
    opts = dict(precond=4, rtol=1e-20 if precision == 0 else 1e-5, atol=0.0, precision=precision)
    g, o, _ = _loaded_pair(lat.HEXAGON2D, xyz)
    x0 = g.get("XYZ")
    g.assemble(precision=precision)
    g.solve_linear(**opts)
    assert g.multigrid_levels() == _level_sizes(xyz, 2, None)
    bd = _crack(o, xyz)
    o.set("BDAMAGE", bd)

    def cracked(s):
        s.set("BDAMAGE", bd)
        s.set("XYZ", x0)
        s.forward()
        s.assemble(precision=precision)
        s.update_rr()

    cracked(g)
    g.solve_linear(**opts)  # not a load step: the hierarchy stays
    assert g.multigrid_levels() == _level_sizes(xyz, 2, None)
    g.set("XYZ", x0)
    g.solve_step(-1, 1, 1e-11, **opts)  # one Newton iteration of a load step
    assert g.multigrid_levels() == _level_sizes(xyz, 2, intact_edges(o))
    fresh, _, _ = _loaded_pair(lat.HEXAGON2D, xyz)
    out = []
    for s in (g, fresh):
        cracked(s)
        st = s.solve_linear(**opts)
        out.append((int(st[0]), s.get("XYZ") - x0))
    assert fresh.multigrid_levels() == g.multigrid_levels()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def _run(prob, backend, opts, tol=1e-11):
    s, gids = P.build(backend, prob)
    s.forward()
    stats = P.solve_problem(s, gids, P.Problem(**{**prob.__dict__, "tol": tol}), opts)
    return s, stats


@pytest.mark.parametrize("prob", [P.sc_cube(n=10, steps=3), P.hex_plate(nx=40, ny=36, steps=3),
                                  P.bcc_j2_bar(nx=3, ny=3, nz=6, steps=10, strain_step=5e-4)],
                         ids=["sc_cube", "hex_plate", "bcc_j2_bar"])
def test_sa_quasi_static_matches_oracle(prob):
    g, _ = _run(prob, lpmg.System, SA64)
    o, _ = _run(prob, Oracle, ORC)
    assert len(g.multigrid_levels()) >= 2
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8
    assert _relerr(g.get("PIN"), o.get("PIN")) < 1e-8
    if prob.ptype == lat.J2PLASTICITY:
        ag, ao = g.get("STATE_VAR")[0::7], o.get("STATE_VAR")[0::7]
        np.testing.assert_array_equal(ag > 0, ao > 0)
        assert (ao > 0).any()


@pytest.mark.parametrize("precision", [0, 1])
def test_sa_ct_with_bond_breaking_matches_oracle(precision):
    prob = P.ct_plate(3000, 0.30, steps=12, opening=3e-5)
    opts = dict(SA64) if precision == 0 else dict(precond=4, rtol=1e-4, atol=0.0, precision=1)
    g, _ = _run(prob, lpmg.System, opts, tol=1e-10)
    o, _ = _run(prob, Oracle, ORC, tol=1e-10)
    x0 = o.get("XYZ_INITIAL")
    bg, bo = g.get("BDAMAGE") > 0.5, o.get("BDAMAGE") > 0.5
    assert bo.sum() > 0
    np.testing.assert_array_equal(bg, bo)
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8


def test_sa_fp32_inner_solve_matches_fp64_oracle():
    prob = P.hex_plate(nx=40, ny=36, steps=3)
    g, _ = _run(prob, lpmg.System, dict(precond=4, rtol=1e-4, atol=0.0, precision=1))
    o, _ = _run(prob, Oracle, ORC)
    x0 = o.get("XYZ_INITIAL")
    assert _relerr(g.get("XYZ") - x0, o.get("XYZ") - x0) < 1e-8


def test_sa_single_level_dense_coarsest():
    prob = P.hex_plate(nx=8, ny=8, steps=2)  # 128 unknowns: the coarsest level is the system
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.apply_step(s, gids, prob.steps[0])
    s.assemble()
    s.forward()
    s.update_rr()
    x0 = s.get("XYZ")
    st = s.solve_linear(precond=4, rtol=1e-20, atol=0.0)
    assert s.multigrid_levels() == [128]
    assert st[0] <= 2
    du = s.get("XYZ") - x0
    s.set("XYZ", x0)
    s.solve_linear(precond=1, rtol=1e-24, atol=0.0)
    assert _relerr(du, s.get("XYZ") - x0) < 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_sa_reruns_are_bit_identical(precision):
    prob = P.hex_plate(nx=60, ny=50, steps=3)
    opts = SA64 if precision == 0 else dict(precond=4, rtol=1e-4, atol=0.0, precision=1)
    a, _ = _run(prob, lpmg.System, opts)
    b, _ = _run(prob, lpmg.System, opts)
    np.testing.assert_array_equal(a.get("XYZ"), b.get("XYZ"))


def test_switching_between_hierarchies_on_one_handle():
    # the handle rebuilds its hierarchy when the kind asked for changes
    prob = P.hex_plate(nx=50, ny=44, steps=1)
    s, gids = P.build(lpmg.System, prob)
    s.forward()
    P.apply_step(s, gids, prob.steps[0])
    s.assemble()
    s.forward()
    s.update_rr()
    x0 = s.get("XYZ")
    out = {}
    for pc in (4, 2, 4):
        s.set("XYZ", x0)
        st = s.solve_linear(precond=pc, rtol=1e-20, atol=0.0)
        out.setdefault(pc, []).append((int(st[0]), s.get("XYZ") - x0))
    assert out[4][0][0] == out[4][1][0]
    np.testing.assert_array_equal(out[4][0][1], out[4][1][1])
    assert out[2][0][0] > out[4][0][0]
    assert _relerr(out[2][0][1], out[4][0][1]) < 1e-7


def test_served_sa_request_within_its_memory_estimate():
    # explicit option 4 in 3D and the server's own choice for a 2D request
    # above the fusion limit: each stays within the admission estimate and
    # gives the same bits as the same request solved on its own handle
    reqs = [(P.sc_cube(n=16, steps=2), dict(precond=4, rtol=1e-4, atol=0.0, precision=1)),
            (P.ct_plate(12000, 0.35, steps=4, opening=7.5e-6, crit=1e-3, tol=1e-9),
             dict(precond=-1, rtol=1e-4, atol=0.0, precision=1, predict=1.0))]
    for prob, o in reqs:
        out = []
        for slots in (1, 3):
            rq = Request(prob, solver_opts(**o), out_fields=("XYZ",))
            with Server(max_active=slots, batching=10000) as srv:
                extra = [srv.submit(Request(P.hex_plate(nx=30, ny=30, steps=2),
                                            solver_opts(**dict(o, precond=-1)), out_fields=("XYZ",)))
                         for _ in range(slots - 1)]
                status, st = srv.wait(srv.submit(rq))
                for t in extra:
                    assert srv.wait(t)[0] == 0
            assert status == 0, st.get("error")
            assert st["bytes"] <= st["bytes_est"], (st["bytes"], st["bytes_est"])
            out.append(rq.out["XYZ"].copy())
        np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("prob", [P.ct_plate(12000, 0.35, steps=3, opening=7.5e-6, crit=1e-3, tol=1e-9),
                                  P.bcc_j2_bar(nx=9, ny=9, nz=18, steps=3, strain_step=5e-4)],
                         ids=["ct", "bcc"])
def test_device_and_host_symbolic_lists_agree(prob, precision, monkeypatch):
    # the Y and coarse symbolic lists are built on the device (the host builds
    # them for rows the device kernel cannot hold, or with
    # LPMG_SA_HOST_SYMBOLIC set): both must give the same hierarchy, so the
    # same bits in a full quasi-static run
    opts = dict(precond=4, rtol=1e-8 if precision == 0 else 1e-4, atol=0.0, precision=precision)
    runs = []
    for host in (False, True):
        if host:
            monkeypatch.setenv("LPMG_SA_HOST_SYMBOLIC", "1")
        else:
            monkeypatch.delenv("LPMG_SA_HOST_SYMBOLIC", raising=False)
        s, stats = _run(prob, lpmg.System, opts, tol=1e-9)
        runs.append((s.get("XYZ"), s.multigrid_levels(), [tuple(x) for x in stats]))
        s.close()
    assert runs[0][1] == runs[1][1] and len(runs[0][1]) >= 3
    assert runs[0][2] == runs[1][2]
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
