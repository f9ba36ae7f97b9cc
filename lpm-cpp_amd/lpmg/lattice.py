"""Lattice generators for LPM particle systems (host-side input formats).

`create_plate_hex2d` and `create_cuboid_sc3d` restate LPM-CPP's
createPlateHEX2D (include/utilities.h:127-177) and createCuboidSC3D
(include/utilities.h:179-220) for the identity rotation, where R*p reduces to p
exactly whatever order the reference's cblas_dgemv sums in.  The reference has
no BCC generator; `bcc_block` places a corner and a body-centre particle per
cubic cell of edge a = 4r/sqrt(3), which makes the first shell 2r and the
second shell a, matching the BCC unit cell's cutoffs (include/unit_cell.h:108-122).
"""
import math

import numpy as np

# LatticeType (include/lpm.h:99-108); the three BCC slip variants share one unit cell
SQUARE2D, HEXAGON2D, SC3D, FCC3D, BCC3D = 0, 1, 2, 3, 4
# ParticleType (include/lpm.h:91-97)
ELASTIC, ELASTIC_DAMAGE, J2PLASTICITY, FATIGUE_HCF = 0, 1, 2, 3


def create_plate_hex2d(box, radius):
    """createPlateHEX2D with R = identity (include/utilities.h:127-177)."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2)
    box_t = [-a, a, -a, a, -a, a]
    hx = 2 * radius
    hy = hx * math.sqrt(3.0) / 2.0
    pfr = 1 + math.floor((box_t[1] - box_t[0]) / hx)
    rows = 1 + math.floor((box_t[3] - box_t[2]) / hy)
    pts = []
    i_odd = np.arange(1, pfr + 1, dtype=np.float64)
    i_even = np.arange(1, pfr, dtype=np.float64)
    x_odd = box_t[0] + hx * (i_odd - 1)
    x_even = box_t[0] + hx * (2 * i_even - 1) / 2.0
    for j in range(1, rows + 1):
        y = box_t[2] + hy * (j - 1)
        if y < box[2] or y > box[3]:
            continue
        xs = x_odd if j % 2 == 1 else x_even
        xs = xs[(xs >= box[0]) & (xs <= box[1])]
        if xs.size:
            pts.append(np.stack([xs, np.full_like(xs, y), np.zeros_like(xs)], axis=1))
    return np.concatenate(pts, axis=0) if pts else np.zeros((0, 3))


def create_cuboid_sc3d(box, radius):
    """createCuboidSC3D with R = identity (include/utilities.h:179-220); note the
    x grid starts at box[0] while y and z start at -a, as in the reference."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2 + (box[5] - box[4]) ** 2)
    box_t = [-a, a, -a, a, -a, a]
    hx = 2 * radius
    n_i = 1 + int(math.floor((box_t[1] - box_t[0]) / hx))
    n_j = 1 + int(math.floor((box_t[3] - box_t[2]) / hx))
    n_k = 1 + int(math.floor((box_t[5] - box_t[4]) / hx))
    xs = box[0] + hx * (np.arange(1, n_i + 1, dtype=np.float64) - 1)
    ys = box_t[2] + hx * (np.arange(1, n_j + 1, dtype=np.float64) - 1)
    zs = box_t[4] + hx * (np.arange(1, n_k + 1, dtype=np.float64) - 1)
    xs = xs[(xs >= box[0]) & (xs <= box[1])]
    ys = ys[(ys >= box[2]) & (ys <= box[3])]
    zs = zs[(zs >= box[4]) & (zs <= box[5])]
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def hex_plate(nx, ny, radius):
    """Hexagonal plate of ny rows; even rows hold nx particles at x = 2r*i, odd
    rows are shifted by r (and hold nx-1 when `short_rows`)."""
    pts = []
    hy = math.sqrt(3.0) * radius
    for j in range(ny):
        xs = 2 * radius * np.arange(nx, dtype=np.float64) + (radius if j % 2 else 0.0)
        pts.append(np.stack([xs, np.full(nx, hy * j), np.zeros(nx)], axis=1))
    return np.concatenate(pts, axis=0)


def sc_block(nx, ny, nz, radius):
    """Simple-cubic block, spacing 2r, ids x-fastest then y then z."""
    h = 2 * radius
    Z, Y, X = np.meshgrid(h * np.arange(nz, dtype=np.float64), h * np.arange(ny, dtype=np.float64),
                          h * np.arange(nx, dtype=np.float64), indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def bcc_block(nx, ny, nz, radius):
    """BCC block of nx*ny*nz cubic cells (2 particles per cell), cell edge 4r/sqrt(3);
    per cell the corner particle precedes the body-centre one, cells x-fastest."""
    a = 4.0 * radius / math.sqrt(3.0)
    Z, Y, X = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    corner = np.stack([X.ravel() * a, Y.ravel() * a, Z.ravel() * a], axis=1)
    centre = corner + 0.5 * a
    out = np.empty((2 * corner.shape[0], 3))
    out[0::2] = corner
    out[1::2] = centre
    return out


def fcc_block(nx, ny, nz, radius):
    """FCC block of nx*ny*nz cubic cells (4 particles per cell), cell edge
    2*sqrt(2)*r, which makes the first shell 2r and the second shell 2*sqrt(2)*r
    as the FCC unit cell's cutoffs (include/unit_cell.h:92-106); per cell the
    corner particle, then the xy-, xz- and yz-face centres, cells x-fastest.
    The reference reads FCC systems from dump files and has no generator."""
    a = 2.0 * math.sqrt(2.0) * radius
    Z, Y, X = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    corner = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    return ((corner[:, None, :] + basis[None, :, :]) * a).reshape(-1, 3)


def unit_cell_info(lattice_type, radius):
    """Neighbour counts and shell cutoffs of UnitCell (include/unit_cell.h:29-122)."""
    r = float(radius)
    if lattice_type == SQUARE2D:
        return dict(dim=2, nneighbors=8, cut=(2.0 * r, 2.0 * math.sqrt(2.0) * r))
    if lattice_type == HEXAGON2D:
        return dict(dim=2, nneighbors=12, cut=(2.0 * r, 2.0 * math.sqrt(3.0) * r))
    if lattice_type == SC3D:
        return dict(dim=3, nneighbors=18, cut=(2.0 * r, 2.0 * math.sqrt(2.0) * r))
    if lattice_type == FCC3D:
        return dict(dim=3, nneighbors=18, cut=(2.0 * r, 2.0 * math.sqrt(2.0) * r))
    if lattice_type == BCC3D:
        return dict(dim=3, nneighbors=14, cut=(2.0 * r, 4.0 / math.sqrt(3.0) * r))
    raise ValueError(f"unknown lattice {lattice_type}")
