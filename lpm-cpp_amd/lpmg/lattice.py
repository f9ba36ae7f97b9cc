"""Lattice generators for LPM particle systems (host-side input formats).

`create_rmatrix`, `create_plate_sq2d`, `create_plate_hex2d`,
`create_cuboid_sc3d` and `create_cuboid_fcc3d` restate LPM-CPP's createRMatrix
and lattice generators (include/utilities.h:15-330), rotation included: each
rotated coordinate is the row sum R[0]*p0 + R[1]*p1 + R[2]*p2 taken left to
right (the reference calls cblas_dgemv; for the identity rotation R*p is p
exactly whatever order it sums in).  include/lpmcpp/utilities.h holds the same
generators for C++ drivers.  The reference has
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


def create_rmatrix(eulerflag, angles):
    """createRMatrix (include/utilities.h:15-91): row-major 3x3 rotation for
    Euler angles (radians); eulerflag 0 = X1-Y2-Z3, 1 = Kocks Z1-Y2-Z3,
    2 = Bunge Z1-X2-Z3.  Each entry is formed by the reference's expression."""
    c1, s1 = math.cos(angles[0]), math.sin(angles[0])
    c2, s2 = math.cos(angles[1]), math.sin(angles[1])
    c3, s3 = math.cos(angles[2]), math.sin(angles[2])
    if eulerflag == 0:
        m = [c2 * c3, -c2 * s3, s2,
             c1 * s3 + s1 * s2 * c3, c1 * c3 - s1 * s2 * s3, -s1 * c2,
             s1 * s3 - c1 * s2 * c3, s1 * c3 + c1 * s2 * s3, c1 * c2]
    elif eulerflag == 1:
        m = [-c1 * c2 * c3 - s1 * s3, c3 * s1 - s3 * c2 * c1, s2 * c1,
             s3 * c1 - c3 * c2 * s1, -c1 * c3 - s1 * c2 * s3, s2 * s1,
             c3 * s2, s3 * s2, c2]
    elif eulerflag == 2:
        m = [-s1 * c2 * s3 + c1 * c3, -s3 * c1 - c3 * c2 * s1, s1 * s2,
             c3 * s1 + s3 * c2 * c1, -s1 * s3 + c1 * c2 * c3, -c1 * s2,
             s2 * s3, s2 * c3, c2]
    else:
        raise ValueError("eulerflag must be 0, 1 or 2")
    return np.array(m, dtype=np.float64)


IDENTITY = np.array([1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0])


def _rotate_keep(pts, R, box, dims):
    """R p for every point (each row summed left to right, as C evaluates
    R[0]*p0 + R[1]*p1 + R[2]*p2), then the points inside the box's first
    `dims` ranges, in order."""
    R = np.asarray(R, dtype=np.float64).reshape(3, 3)
    q = np.empty_like(pts)
    for r in range(3):
        q[:, r] = R[r, 0] * pts[:, 0] + R[r, 1] * pts[:, 1] + R[r, 2] * pts[:, 2]
    keep = np.ones(len(q), dtype=bool)
    for k in range(dims):
        keep &= (q[:, k] >= box[2 * k]) & (q[:, k] <= box[2 * k + 1])
    return q[keep]


def create_plate_sq2d(box, radius, R=IDENTITY):
    """createPlateSQ2D (include/utilities.h:93-125): square grid of cell
    centres over the padded block, z = 0 after rotation."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2)
    hx = 2 * radius
    nx, ny = int(math.floor((a - -a) / hx)), int(math.floor((a - -a) / hx))
    xs = -a + hx * (np.arange(1, nx + 1, dtype=np.float64) - 1) + radius
    ys = -a + hx * (np.arange(1, ny + 1, dtype=np.float64) - 1) + radius
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    q = _rotate_keep(pts, R, box, 2)
    q[:, 2] = 0.0
    return q


def create_plate_hex2d(box, radius, R=IDENTITY):
    """createPlateHEX2D (include/utilities.h:127-177): odd rows one particle
    longer than the even rows, which are shifted by r; rotated by R and kept
    inside the box."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2)
    box_t = [-a, a, -a, a, -a, a]
    hx = 2 * radius
    hy = hx * math.sqrt(3.0) / 2.0
    pfr = 1 + math.floor((box_t[1] - box_t[0]) / hx)
    rows = 1 + math.floor((box_t[3] - box_t[2]) / hy)
    i_odd = np.arange(1, pfr + 1, dtype=np.float64)
    i_even = np.arange(1, pfr, dtype=np.float64)
    x_odd = box_t[0] + hx * (i_odd - 1)
    x_even = box_t[0] + hx * (2 * i_even - 1) / 2.0
    pts = []
    for j in range(1, rows + 1):
        y = box_t[2] + hy * (j - 1)
        xs = x_odd if j % 2 == 1 else x_even
        pts.append(np.stack([xs, np.full_like(xs, y), np.zeros_like(xs)], axis=1))
    return _rotate_keep(np.concatenate(pts, axis=0), R, box, 2)


def create_cuboid_sc3d(box, radius, R=IDENTITY):
    """createCuboidSC3D (include/utilities.h:179-220); the x grid starts at
    box[0] while y and z start at -a, as in the reference."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2 + (box[5] - box[4]) ** 2)
    box_t = [-a, a, -a, a, -a, a]
    hx = 2 * radius
    n_i = 1 + int(math.floor((box_t[1] - box_t[0]) / hx))
    n_j = 1 + int(math.floor((box_t[3] - box_t[2]) / hx))
    n_k = 1 + int(math.floor((box_t[5] - box_t[4]) / hx))
    xs = box[0] + hx * (np.arange(1, n_i + 1, dtype=np.float64) - 1)
    ys = box_t[2] + hx * (np.arange(1, n_j + 1, dtype=np.float64) - 1)
    zs = box_t[4] + hx * (np.arange(1, n_k + 1, dtype=np.float64) - 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    return _rotate_keep(pts, R, box, 3)


def create_cuboid_fcc3d(box, radius, R=IDENTITY):
    """createCuboidFCC3D (include/utilities.h:222-330): layers half a cell
    apart; in odd layers the odd rows start at the block's edge and the even
    rows half a cell in, the other way round in even layers."""
    a = math.sqrt((box[1] - box[0]) ** 2 + (box[3] - box[2]) ** 2 + (box[5] - box[4]) ** 2)
    h = 4 * radius / math.sqrt(2.0)
    nx = 1 + int(math.floor((a - -a) / h))
    ny = 1 + int(math.floor((a - -a) / (h / 2.0)))
    full_x = -a + h * (np.arange(1, nx + 1, dtype=np.float64) - 1)
    half_x = -a + h * (np.arange(1, nx, dtype=np.float64) - 1) + h / 2.0
    pts = []
    for k in range(1, ny + 1):
        z = -a + h / 2.0 * (k - 1)
        for j in range(1, ny + 1):
            y = -a + h / 2.0 * (j - 1)
            xs = full_x if (k % 2 == 1) == (j % 2 == 1) else half_x
            pts.append(np.stack([xs, np.full_like(xs, y), np.full_like(xs, z)], axis=1))
    return _rotate_keep(np.concatenate(pts, axis=0), R, box, 3)


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
