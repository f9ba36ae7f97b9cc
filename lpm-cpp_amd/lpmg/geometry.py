"""Specimen pre-processors and the reference's particle/bond file formats.

CT specimen (geometry/geo1_CT_configuration.cpp:44-171): particles of a
lattice are kept when `isValid` (outside both pin holes and the notch); a
candidate bond (distance < 1.01 * second-shell cutoff) is kept when none of the
midpoints of its 10 equal sub-segments lies in an excluded region; its layer is
0 below 1.01 * first-shell cutoff, else 1.  Bonds are listed by first particle,
partners ascending (`searchNeighbor`, geo1_CT_configuration.cpp:78-132).  The
reference scans all O(N^2) pairs; here candidates come from a uniform grid of
cell size 1.01 * cutoff, then each bond's test uses the reference's arithmetic
(the sub-segment points are xi + k*dx and xi + (k+1)*dx with dx = (xj - xi)/10,
the midpoint their half-sum; distances are sqrt of summed squares, as g++
compiles pow(d, 2.0) to d*d; `abs` of a double is fabs under libstdc++).

`inNotch` (geo1_CT_configuration.cpp:52-62) excludes the wedge opening towards
+x from the tip: |y - yt| <= hw and |y - yt| <= (x - xt)/sqrt(3).  The CT here
therefore has its pin holes and load line on the right and the ligament on the
left, as the reference's Ti64 CT (`isValid`, :137-171).

Dump and bond files: writeConfigurationDump / readDump and writeBond / readBond
(include/assembly.h:268-424).  The dump keeps 5 significant digits ("%.4e"),
so coordinates read back from a dump are the rounded ones, as in the reference.
"""
import math

import numpy as np

from . import lattice as lat

SQRT3_INV = 1.0 / math.sqrt(3.0)


def in_circle(x, y, cx, cy, R):
    """inCircle (geo1_CT_configuration.cpp:44-49), z ignored."""
    dx = x - cx
    dy = y - cy
    return dx * dx + dy * dy <= R * R


def in_notch(x, y, tx, ty, hw):
    """inNotch (geo1_CT_configuration.cpp:52-62): the wedge towards +x from the tip."""
    d = y - ty
    return (np.abs(d) <= hw) & (d >= SQRT3_INV * (tx - x)) & (d <= SQRT3_INV * (-tx + x))


class Region:
    """Excluded regions of a specimen: circles (cx, cy, R) and notches (tx, ty, hw)."""

    def __init__(self, circles=(), notches=()):
        self.circles = [tuple(map(float, c)) for c in circles]
        self.notches = [tuple(map(float, n)) for n in notches]

    def invalid(self, x, y):
        bad = np.zeros(np.broadcast(x, y).shape, dtype=bool)
        for cx, cy, R in self.circles:
            bad |= in_circle(x, y, cx, cy, R)
        for tx, ty, hw in self.notches:
            bad |= in_notch(x, y, tx, ty, hw)
        return bad

    def valid(self, xyz):
        """isValid per particle."""
        return ~self.invalid(xyz[:, 0], xyz[:, 1])


def _candidate_pairs(xyz, rc):
    """All ordered pairs (i, j), i != j, with |xj - xi| < rc (grid of cell size rc)."""
    n = xyz.shape[0]
    lo = xyz.min(axis=0)
    cell = np.floor((xyz - lo) / rc).astype(np.int64)
    dims = cell.max(axis=0) + 1
    key = (cell[:, 2] * dims[1] + cell[:, 1]) * dims[0] + cell[:, 0]
    order = np.argsort(key, kind="stable")
    skey = key[order]
    ncell = int(dims[0] * dims[1] * dims[2])
    start = np.searchsorted(skey, np.arange(ncell + 1))
    I, J = [], []
    offs = [(a, b, c) for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)]
    for ox, oy, oz in offs:
        nc = cell + np.array([ox, oy, oz])
        ok = np.all((nc >= 0) & (nc < dims), axis=1)
        src = np.nonzero(ok)[0]
        nkey = (nc[src, 2] * dims[1] + nc[src, 1]) * dims[0] + nc[src, 0]
        s0, s1 = start[nkey], start[nkey + 1]
        cnt = s1 - s0
        if cnt.sum() == 0:
            continue
        ii = np.repeat(src, cnt)
        pos = np.repeat(s0 - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt) + np.arange(cnt.sum())
        jj = order[pos]
        I.append(ii)
        J.append(jj)
    I = np.concatenate(I)
    J = np.concatenate(J)
    keep = I != J
    I, J = I[keep], J[keep]
    d = xyz[J] - xyz[I]
    dis = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    keep = dis < rc
    return I[keep], J[keep], dis[keep]


def search_neighbor(xyz, cell, region=None, chunk=1 << 20):
    """searchNeighbor (geo1_CT_configuration.cpp:78-132): bonds (layer, p1, p2)
    of the particle set `xyz`, excluding those crossing `region`."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    c0, c1 = 1.01 * cell["cut"][0], 1.01 * cell["cut"][1]
    I, J, dis = _candidate_pairs(xyz, c1)
    if region is not None and (region.circles or region.notches):
        valid = np.ones(I.shape[0], dtype=bool)
        for s in range(0, I.shape[0], chunk):
            i, j = I[s:s + chunk], J[s:s + chunk]
            xi, yi = xyz[i, 0], xyz[i, 1]
            dx = (xyz[j, 0] - xi) / 10
            dy = (xyz[j, 1] - yi) / 10
            v = np.ones(i.shape[0], dtype=bool)
            for k in range(10):
                x1 = xi + k * dx
                x2 = xi + (k + 1) * dx
                y1 = yi + k * dy
                y2 = yi + (k + 1) * dy
                v &= ~region.invalid((x1 + x2) / 2.0, (y1 + y2) / 2.0)
            valid[s:s + chunk] = v
        I, J, dis = I[valid], J[valid], dis[valid]
    order = np.lexsort((J, I))
    I, J, dis = I[order], J[order], dis[order]
    layer = np.where(dis < c0, 0, 1).astype(np.int32)
    return layer, I.astype(np.int32), J.astype(np.int32)


def ct_specimen(n_target, a_over_w, W=50.0, hw_factor=0.5):
    """Hexagonal-lattice CT specimen of about `n_target` particles.

    Width W (load line to back edge), overall 1.25 W x 1.2 W; pin holes of
    radius 0.125 W centred on the load line x = W at 0.275 W above and below
    the crack line; a slit of half-width hw_factor * r from the right edge to
    the tip at x = W - a.  The crack line sits half-way between two particle
    rows, so the slit removes no particle and cuts every bond that crosses it.
    Returns a dict with xyz, bonds (p1, p2, layer), radius, geometry."""
    area = 1.25 * 1.2 - 2.0 * math.pi * 0.125 ** 2
    r = W * math.sqrt(area / (2.0 * math.sqrt(3.0) * n_target))
    box = [0.0, 1.25 * W, 0.0, 1.2 * W, 0.0, 1.0]
    xyz = lat.create_plate_hex2d(box, r)
    rows = np.unique(xyz[:, 1])
    j = int(np.searchsorted(rows, 0.6 * W, side="right")) - 1
    j = min(max(j, 0), len(rows) - 2)
    yc = 0.5 * (rows[j] + rows[j + 1])
    R = 0.125 * W
    a = a_over_w * W
    region = Region(circles=[(W, yc + 0.275 * W, R), (W, yc - 0.275 * W, R)],
                    notches=[(W - a, yc, hw_factor * r)])
    xyz = xyz[region.valid(xyz)]
    cell = lat.unit_cell_info(lat.HEXAGON2D, r)
    layer, p1, p2 = search_neighbor(xyz, cell, region)
    # `bonds` in the argument order of lpmg.System / lpmg_create_with_bonds
    return dict(xyz=xyz, bonds=(p1, p2, layer), radius=r, box=box, W=W, a=a, yc=yc, R=R, region=region,
                pin_top=(W, yc + 0.275 * W), pin_bottom=(W, yc - 0.275 * W))


# ---- files -----------------------------------------------------------------

def write_dump(path, xyz, box, types=None, ids=None):
    """writeConfigurationDump (assembly.h:284-306): id type x y z."""
    n = xyz.shape[0]
    types = np.zeros(n, dtype=np.int64) if types is None else types
    ids = np.arange(n) if ids is None else ids
    with open(path, "w") as f:
        f.write("ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % n)
        for k in range(3):
            f.write("%8.8f %8.8f\n" % (box[2 * k], box[2 * k + 1]))
        f.write("ITEM: ATOMS id type x y z\n")
        for i in range(n):
            f.write("%d %d %.4e %.4e %.4e \n" % (ids[i], types[i], xyz[i, 0], xyz[i, 1], xyz[i, 2]))


def read_dump(path):
    """readDump (assembly.h:339-392): (ids, types, xyz, box), particles in file order."""
    with open(path) as f:
        lines = f.read().split("\n")
    n = int(lines[3].split()[0])
    box = [float(v) for k in range(3) for v in lines[5 + k].split()[:2]]
    rows = [ln.split() for ln in lines[9:] if ln.strip()]
    rows = rows[:n] if len(rows) >= n else rows
    ids = np.array([int(r[0]) for r in rows], dtype=np.int64)
    types = np.array([int(r[1]) for r in rows], dtype=np.int32)
    xyz = np.array([[float(r[2]), float(r[3]), float(r[4])] for r in rows], dtype=np.float64)
    return ids, types, xyz, box


def write_bond(path, layer, p1, p2):
    """The bond file of geo1's writeBond (geo1_CT_configuration.cpp:33-41): id layer p1 p2."""
    with open(path, "w") as f:
        for i in range(len(p1)):
            f.write("%d %d %d %d \n" % (i, layer[i], p1[i], p2[i]))


def read_bond(path):
    """readBond (assembly.h:394-424): (layer, p1, p2) in file order."""
    a = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if a.size == 0:
        z = np.zeros(0, dtype=np.int32)
        return z, z, z
    return a[:, 1].astype(np.int32), a[:, 2].astype(np.int32), a[:, 3].astype(np.int32)


# ---- BCC polycrystal ---------------------------------------------------------

def random_rotations(n, rng):
    """Uniformly distributed rotations (unit quaternions, Shoemake's method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def bcc_polycrystal(n_cells=80, n_grains=200, radius=0.1, seed=7, offset=1.3, threshold=1.3, min_bonds=6):
    """Voronoi BCC polycrystal in a cube of n_cells BCC cells per edge.

    The algorithm of benchmark/BCC_Hailong/create_BCC_configuration.F90:169-333:
    per grain, a BCC lattice (lattice constant 4r/sqrt(3), a corner and a body
    centre per cell) is rotated by the grain's orientation and centred on its
    seed; a point is kept when it is closer to its own seed than to every other
    seed by less than `offset` r (the cell widened a little) and lies in the
    box; grains are added in order, and a point closer than `threshold` r to a
    point of an earlier grain is dropped (that earlier point is marked a
    grain-boundary particle).  Seeds are uniform in the box, orientations
    uniform over SO(3), both from `seed`.  The nearest-seed test uses a k-d tree
    instead of the reference's neighbour-seed list.  Last, particles with fewer
    than `min_bonds` neighbours inside the second-shell cutoff are removed,
    repeatedly until none is left: a grain-boundary particle held by two or
    three central bonds is a mechanism that makes the stiffness singular.
    Returns xyz, the grain of each particle (n_grains for grain-boundary
    particles), seeds, rotations."""
    from scipy.spatial import cKDTree
    a = 4.0 * radius / np.sqrt(3.0)
    L = n_cells * a
    rng = np.random.default_rng(seed)
    seeds = rng.random((n_grains, 3)) * L
    rots = random_rotations(n_grains, rng)
    tree = cKDTree(seeds)
    # the lattice block around each seed: a ball that covers its Voronoi cell
    # (cells farther than 3 mean grain radii from their seed are not expected)
    rg = 3.0 * (3.0 * L ** 3 / (4.0 * np.pi * n_grains)) ** (1.0 / 3.0)
    m = int(np.ceil(rg / a)) + 1
    g = np.arange(-m, m + 1, dtype=np.float64) * a
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    corner = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    block = np.concatenate([corner, corner + 0.5 * a], axis=0)
    block = block[np.einsum("ij,ij->i", block, block) <= rg * rg]
    pts, grain = [], []
    for k in range(n_grains):
        p = block @ rots[k].T + seeds[k]
        inside = np.all((p >= 0.0) & (p <= L), axis=1)
        p = p[inside]
        d, idx = tree.query(p, k=2)
        d_own = np.linalg.norm(p - seeds[k], axis=1)
        other = np.where(idx[:, 0] == k, d[:, 1], d[:, 0])
        keep = other > d_own - offset * radius
        pts.append(p[keep])
        grain.append(np.full(int(keep.sum()), k, dtype=np.int32))
    xyz = np.concatenate(pts, axis=0)
    grain = np.concatenate(grain)
    # overlaps: grains in order; a point within threshold*r of a kept point of
    # an earlier grain is dropped
    pairs = cKDTree(xyz).query_pairs(threshold * radius, output_type="ndarray")
    gi, gj = grain[pairs[:, 0]], grain[pairs[:, 1]]
    cross = gi != gj
    pairs = pairs[cross]
    early = np.where(grain[pairs[:, 0]] < grain[pairs[:, 1]], pairs[:, 0], pairs[:, 1])
    late = np.where(grain[pairs[:, 0]] < grain[pairs[:, 1]], pairs[:, 1], pairs[:, 0])
    order = np.argsort(grain[late], kind="stable")
    early, late = early[order], late[order]
    kept = np.ones(len(xyz), dtype=bool)
    boundary = np.zeros(len(xyz), dtype=bool)
    lg = grain[late]
    starts = np.searchsorted(lg, np.arange(n_grains + 1))
    for k in range(n_grains):
        e, l = early[starts[k]:starts[k + 1]], late[starts[k]:starts[k + 1]]
        hit = kept[e]
        kept[l[hit]] = False
        boundary[e[hit]] = True
    grain = np.where(boundary, n_grains, grain)
    xyz, grain = xyz[kept], grain[kept]
    if min_bonds > 0:
        rc = 1.01 * lat.unit_cell_info(lat.BCC3D, radius)["cut"][1]
        pr = cKDTree(xyz).query_pairs(rc, output_type="ndarray")
        nb = np.bincount(pr.ravel(), minlength=len(xyz))
        alive = np.ones(len(xyz), dtype=bool)
        while True:
            weak = alive & (nb < min_bonds)
            if not weak.any():
                break
            alive &= ~weak
            for a_, b_ in ((0, 1), (1, 0)):  # partners of the removed lose a bond
                hit = weak[pr[:, a_]]
                np.subtract.at(nb, pr[hit, b_], 1)
            pr = pr[~(weak[pr[:, 0]] | weak[pr[:, 1]])]
        xyz, grain = xyz[alive], grain[alive]
    return dict(xyz=xyz, grain=grain, seeds=seeds, rotations=rots, L=L, radius=radius)
