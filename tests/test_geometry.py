"""CT specimen pre-processor and the reference's dump / bond file formats
(geometry/geo1_CT_configuration.cpp, include/assembly.h:268-424).

The grid-based bond search of lpmg.geometry must give the bond list of the
reference's O(N^2) searchNeighbor + isValid (restated by brute force in the
oracle) bit for bit, and the exclusion must do what the geometry says: no kept
bond crosses the slit or enters a pin hole, while bonds across the ligament
ahead of the tip stay.
"""
import os

import numpy as np
import pytest

from lpmg import geometry as G
from lpmg import lattice as lat
import oracle as O


def _segments_cross(p, q, y0, x_lo, x_hi):
    """Does segment p->q cross the horizontal line y = y0 at x in [x_lo, x_hi]?"""
    s = (p[:, 1] - y0) * (q[:, 1] - y0) < 0
    t = (y0 - p[:, 1]) / np.where(q[:, 1] != p[:, 1], q[:, 1] - p[:, 1], 1.0)
    x = p[:, 0] + t * (q[:, 0] - p[:, 0])
    return s & (x >= x_lo) & (x <= x_hi)


@pytest.mark.parametrize("n,aw", [(2000, 0.2), (6000, 0.5), (15000, 0.37)])
def test_ct_bonds_match_brute_force(n, aw):
    ct = G.ct_specimen(n, aw)
    cell = lat.unit_cell_info(lat.HEXAGON2D, ct["radius"])
    layer, p1, p2 = O.ct_search_neighbor(ct["xyz"], cell["cut"], cell["nneighbors"], ct["region"])
    g1, g2, gl = ct["bonds"]
    assert np.array_equal(p1, g1) and np.array_equal(p2, g2) and np.array_equal(layer, gl)
    full = lat.create_plate_hex2d(ct["box"], ct["radius"])
    assert np.array_equal(O.ct_valid(full, ct["region"]), ct["region"].valid(full))


def test_ct_exclusion_geometry():
    ct = G.ct_specimen(8000, 0.4)
    xyz, (p1, p2, layer) = ct["xyz"], ct["bonds"]
    W, a, yc, R = ct["W"], ct["a"], ct["yc"], ct["R"]
    tip = W - a
    # no particle removed by the slit: it lies between two rows
    assert np.all(np.abs(xyz[:, 1] - yc) > 0.5 * ct["radius"])
    cross_slit = _segments_cross(xyz[p1], xyz[p2], yc, tip + 1e-9, 2 * W)
    assert not cross_slit.any()
    cross_lig = _segments_cross(xyz[p1], xyz[p2], yc, 0.0, tip - 2 * ct["radius"])
    assert cross_lig.sum() > 0
    # no particle inside a hole; every bond is symmetric (both directions kept)
    for cx, cy in (ct["pin_top"], ct["pin_bottom"]):
        assert np.all(np.hypot(xyz[:, 0] - cx, xyz[:, 1] - cy) > R)
    fwd = set(zip(p1.tolist(), p2.tolist()))
    assert all((j, i) in fwd for i, j in fwd)
    # interior particles keep the full 6 + 6 shells
    nb = np.bincount(p1, minlength=len(xyz))
    assert nb.max() == 12 and (nb == 12).mean() > 0.8
    assert np.array_equal(np.bincount(p1[layer == 0], minlength=len(xyz)) <= 6, np.ones(len(xyz), bool))


def test_particle_count_tracks_target():
    for n in (5000, 50000):
        ct = G.ct_specimen(n, 0.3)
        assert abs(len(ct["xyz"]) - n) / n < 0.03


def test_dump_and_bond_files_round_trip(tmp_path):
    ct = G.ct_specimen(3000, 0.3)
    xyz = ct["xyz"]
    types = np.arange(len(xyz)) % 5
    G.write_dump(os.path.join(tmp_path, "s.dump"), xyz, ct["box"], types=types)
    p1, p2, layer = ct["bonds"]
    G.write_bond(os.path.join(tmp_path, "s.bond"), layer, p1, p2)
    ids, t2, x2, box = G.read_dump(os.path.join(tmp_path, "s.dump"))
    assert np.array_equal(ids, np.arange(len(xyz))) and np.array_equal(t2, types)
    assert box == ct["box"]
    # "%.4e": 5 significant digits, as the reference's writer
    assert np.allclose(x2, xyz, rtol=5.01e-5, atol=0)
    assert np.array_equal(x2, np.array([[float("%.4e" % v) for v in row] for row in xyz]))
    l2, q1, q2 = G.read_bond(os.path.join(tmp_path, "s.bond"))
    assert np.array_equal(l2, layer) and np.array_equal(q1, p1) and np.array_equal(q2, p2)
    with open(os.path.join(tmp_path, "s.bond")) as f:
        first = f.readline().split()
    assert first == ["0", str(layer[0]), str(p1[0]), str(p2[0])]


def test_bcc_polycrystal_structure():
    """Voronoi BCC polycrystal (create_BCC_configuration.F90's algorithm): the
    rotations are proper, points keep the overlap threshold, the particle count
    is the lattice's, grain interiors are rotated BCC lattices (nearest
    neighbour at 2r), each point sits in (or within the offset of) its grain's
    Voronoi cell, and the result is reproducible from the seed."""
    from scipy.spatial import cKDTree
    r = 0.1
    pc = G.bcc_polycrystal(12, 10, radius=r, seed=3)
    R = pc["rotations"]
    assert np.allclose(np.einsum("kij,kil->kjl", R, R), np.eye(3)[None], atol=1e-12)
    assert np.allclose(np.linalg.det(R), 1.0)
    xyz, grain = pc["xyz"], pc["grain"]
    n_ideal = 2 * 12 ** 3
    assert 0.9 * n_ideal < len(xyz) < 1.05 * n_ideal
    L = pc["L"]
    assert xyz.min() >= 0.0 and xyz.max() <= L
    d, _ = cKDTree(xyz).query(xyz, k=2)
    assert d[:, 1].min() >= 1.3 * r - 1e-12
    interior = grain < 10
    seeds = pc["seeds"]
    dist = np.linalg.norm(xyz[:, None, :] - seeds[None, :, :], axis=2)
    own = dist[np.arange(len(xyz))[interior], grain[interior]]
    assert np.all(own <= dist[interior].min(axis=1) + 1.3 * r + 1e-12)
    # far from any other seed, a grain is a perfect rotated lattice
    second = np.sort(dist, axis=1)[:, 1]
    deep = interior & (second - dist.min(axis=1) > 6 * r)
    assert deep.sum() > 100
    np.testing.assert_allclose(d[deep, 1], 2 * r, rtol=1e-12)
    pc2 = G.bcc_polycrystal(12, 10, radius=r, seed=3)
    assert np.array_equal(pc2["xyz"], xyz)


def _elemental(axis, a):
    """Active rotation by angle a about a fixed axis (right-handed)."""
    c, s = np.cos(a), np.sin(a)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


@pytest.mark.parametrize("seed", range(4))
def test_rotation_matrix_is_the_product_of_elemental_rotations(seed):
    # createRMatrix (include/utilities.h:15-91) expands a product of elemental
    # rotations: flag 0 = X(a1) Y(a2) Z(a3), flag 2 (Bunge) = Z(a1) X(a2) Z(a3),
    # flag 1 (Kocks) = Z(a1) Y(a2) Z(pi - a3).  The restatement's expanded
    # entries must equal the matrix products; a swapped sign, entry or angle fails.
    rng = np.random.default_rng(seed)
    a1, a2, a3 = rng.uniform(-np.pi, np.pi, 3)
    want = {0: _elemental("x", a1) @ _elemental("y", a2) @ _elemental("z", a3),
            1: _elemental("z", a1) @ _elemental("y", a2) @ _elemental("z", np.pi - a3),
            2: _elemental("z", a1) @ _elemental("x", a2) @ _elemental("z", a3)}
    for flag, M in want.items():
        R = lat.create_rmatrix(flag, [a1, a2, a3]).reshape(3, 3)
        np.testing.assert_allclose(R, M, atol=2e-15, rtol=0)
    # 2D lattices use only the third angle: a rotation in the plane
    R = lat.create_rmatrix(0, [0.0, 0.0, a3]).reshape(3, 3)
    np.testing.assert_allclose(R[:2, :2] @ [1.0, 0.0], [np.cos(a3), np.sin(a3)], atol=1e-15)
    with pytest.raises(ValueError):
        lat.create_rmatrix(3, [a1, a2, a3])


def test_rotated_lattice_is_the_identity_lattice_rotated():
    # createCuboidSC3D rotates a block of lattice points about the origin and
    # keeps those inside the box: every kept point, rotated back, lies on the
    # unrotated lattice x0 + 2r (i, j, k) with x0 the identity generator's origin
    r = 0.1
    box = [0.0, 1.0, 0.0, 1.0, 0.0, 1.0]
    R = lat.create_rmatrix(2, [np.radians(30.0), np.radians(45.0), np.radians(60.0)])
    pts = lat.create_cuboid_sc3d(box, r, R)
    back = pts @ R.reshape(3, 3)  # R^T p for each row p
    ident = lat.create_cuboid_sc3d(box, r)
    origin = ident.min(axis=0)
    k = (back - origin) / (2 * r)
    np.testing.assert_allclose(k, np.round(k), atol=1e-9)
    lo, hi = np.array(box[0::2]), np.array(box[1::2])
    assert np.all(pts >= lo - 1e-12) and np.all(pts <= hi + 1e-12)
