"""Quasi-static LPM problems (particles, material, boundary groups, load steps)
and the load-stepping driver.

`solve_problem` restates SolverStatic::solveProblem (solver_static.h:64-95):
each load step applies its force BCs then its displacement BCs
(solver_static.h:101-102) and runs solveProblemStep; a step that does not
converge resets the state to the last converged one (positions included),
halves the step and runs it twice -- the force BCs of the failed attempt stay
applied, as in the reference (resetStateVar does not restore Pex).

The workloads (sizes, material, loading) are the benchmark configurations:
  hex_plate   2D hexagonal plate 100x100, r=0.1, E=200 GPa, nu=0.3 plane
              stress, bottom row fixed, top row pulled 0.01 mm in 10 steps
  sc_cube     3D simple-cubic 40^3 cube, r=0.1, E=70 GPa, nu=0.33, bottom face
              fixed, top face pulled 0.05 mm in 20 steps
  bcc_j2_bar  3D BCC J2 bar, 30x30x60 cells, E=200 GPa, nu=0.3, sigma_y=250 MPa,
              H=2 GPa, tension 2.5e-4 strain per step
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import lattice as lat

MAT_ELASTIC, MAT_ELASTIC_CUBIC, MAT_DAMAGE, MAT_DAMAGE_CUBIC, MAT_J2 = 0, 1, 2, 3, 4
RELATIVE, ABSOLUTE = 0, 1


@dataclass
class Problem:
    name: str
    lattice: int
    radius: float
    ptype: int
    xyz: np.ndarray
    material: tuple                       # (model, params)
    groups: dict                          # name -> particle ids
    steps: list                           # list of load steps; a step is a list of BC tuples
    types: np.ndarray = None
    undamaged: int = -1
    max_iter: int = 30
    tol: float = 1e-5
    note: str = ""
    extra: dict = field(default_factory=dict)


def _bc_steps(n, per_step):
    return [list(per_step) for _ in range(n)]


def hex_plate(nx=100, ny=100, radius=0.1, steps=10, total=0.01, E=200e3, nu=0.3):
    xyz = lat.hex_plate(nx, ny, radius)
    y = xyz[:, 1]
    bottom = np.nonzero(y < y.min() + 0.5 * radius)[0]
    top = np.nonzero(y > y.max() - 0.5 * radius)[0]
    d = total / steps
    step = [("disp", "bottom", RELATIVE, "x", 0.0), ("disp", "bottom", RELATIVE, "y", 0.0),
            ("disp", "top", RELATIVE, "x", 0.0), ("disp", "top", RELATIVE, "y", d)]
    return Problem("hex_plate", lat.HEXAGON2D, radius, lat.ELASTIC, xyz,
                   (MAT_ELASTIC, [0.0, 1, E, nu, 1.0]), dict(bottom=bottom, top=top),
                   _bc_steps(steps, step),
                   note=f"hex {nx}x{ny}, r={radius}, E={E}, nu={nu} plane stress, top uy {total} in {steps} steps")


def sc_cube(n=40, radius=0.1, steps=20, total=0.05, E=70e3, nu=0.33):
    xyz = lat.sc_block(n, n, n, radius)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 0.5 * radius)[0]
    top = np.nonzero(z > z.max() - 0.5 * radius)[0]
    d = total / steps
    step = [("disp", "bottom", RELATIVE, "x", 0.0), ("disp", "bottom", RELATIVE, "y", 0.0),
            ("disp", "bottom", RELATIVE, "z", 0.0), ("disp", "top", RELATIVE, "z", d)]
    return Problem("sc_cube", lat.SC3D, radius, lat.ELASTIC, xyz,
                   (MAT_ELASTIC, [0.0, 0, E, nu, 1.0]), dict(bottom=bottom, top=top),
                   _bc_steps(steps, step),
                   note=f"SC {n}^3, r={radius}, E={E}, nu={nu}, top uz {total} in {steps} steps")


def bcc_j2_bar(nx=30, ny=30, nz=60, radius=0.1, steps=45, strain_step=2.5e-4, E=200e3, nu=0.3,
               sigmay=250.0, H=2000.0, xi=0.0, A=0.0):
    xyz = lat.bcc_block(nx, ny, nz, radius)
    z = xyz[:, 2]
    a = 4.0 * radius / math.sqrt(3.0)
    bottom = np.nonzero(z < z.min() + 0.25 * a)[0]
    top = np.nonzero(z > z.max() - 0.25 * a)[0]
    d = strain_step * (z.max() - z.min())
    step = [("disp", "bottom", RELATIVE, "x", 0.0), ("disp", "bottom", RELATIVE, "y", 0.0),
            ("disp", "bottom", RELATIVE, "z", 0.0), ("disp", "top", RELATIVE, "z", d)]
    return Problem("bcc_j2_bar", lat.BCC3D, radius, lat.J2PLASTICITY, xyz,
                   (MAT_J2, [0.0, 0, E, nu, sigmay, xi, H, A, 1.0]), dict(bottom=bottom, top=top),
                   _bc_steps(steps, step),
                   note=f"BCC {nx}x{ny}x{nz} cells J2, r={radius}, strain {strain_step}/step x {steps}")


def build(backend, prob):
    """Create the system with `backend(lattice, radius, ptype, xyz)` and set it up."""
    s = backend(prob.lattice, prob.radius, prob.ptype, prob.xyz)
    s.set_material(*prob.material)
    if prob.types is not None:
        s.set_types(prob.types)
    gids = {name: s.define_group(idx) for name, idx in prob.groups.items()}
    return s, gids


def apply_step(s, gids, step, scale=1.0):
    """solveProblemStep's updateForceBC then updateDisplacementBC (solver_static.h:101-102)."""
    for bc in step:
        if bc[0] == "force":
            _, g, mode, fx, fy, fz = bc
            s.force_bc(gids[g], mode, fx * scale, fy * scale, fz * scale)
    for bc in step:
        if bc[0] == "disp":
            _, g, mode, flag, d = bc
            s.disp_bc(gids[g], mode, flag, d * scale)


def solve_problem(s, gids, prob, opts, on_step=None, max_halvings=8):
    """SolverStatic::solveProblem (solver_static.h:64-95) with step halving."""
    stats = []
    queue = [(st, 1.0) for st in prob.steps]
    i = 0
    halvings = 0
    while i < len(queue):
        step, scale = queue[i]
        apply_step(s, gids, step, scale)
        ok, st = s.solve_step(prob.undamaged, prob.max_iter, prob.tol, **opts)
        if not ok:
            if halvings >= max_halvings:
                raise RuntimeError(f"step {i} did not converge after {halvings} halvings")
            halvings += 1
            s.reset_state(True)
            queue[i] = (step, scale / 2)
            queue.insert(i, (step, scale / 2))
            continue
        stats.append(st)
        if on_step:
            on_step(i, st)
        i += 1
    return stats


def prepare(s):
    """The examples' pre-solve pass: updateGeometry + updateForceState
    (e.g. examples/ex1_isoelasticity_3d_sc.cpp:104-105)."""
    s.forward()
