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

MAT_ELASTIC, MAT_ELASTIC_CUBIC, MAT_DAMAGE, MAT_DAMAGE_CUBIC, MAT_J2, MAT_FATIGUE = 0, 1, 2, 3, 4, 5
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
    bonds = prob.extra.get("bonds")
    if bonds is not None:
        s = backend(prob.lattice, prob.radius, prob.ptype, prob.xyz, bonds=bonds)
    else:
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
    """SolverStatic::solveProblem (solver_static.h:64-95) with step halving.
    With opts["predict"] > 0 (GPU path only) each step's predictor scale is
    that value times the step's size over the last converged step's size."""
    stats = []
    queue = [(st, 1.0) for st in prob.steps]
    i = 0
    halvings = 0
    predict = float(opts.get("predict", 0.0))
    last_scale = 1.0
    while i < len(queue):
        step, scale = queue[i]
        apply_step(s, gids, step, scale)
        kw = dict(opts, predict=predict * scale / last_scale) if predict > 0 else opts
        ok, st = s.solve_step(prob.undamaged, prob.max_iter, prob.tol, **kw)
        if not ok:
            if halvings >= max_halvings:
                raise RuntimeError(f"step {i} did not converge after {halvings} halvings")
            halvings += 1
            s.reset_state(True)
            queue[i] = (step, scale / 2)
            queue.insert(i, (step, scale / 2))
            continue
        stats.append(st)
        last_scale = scale
        if on_step:
            on_step(i, st)
        i += 1
    return stats


def prepare(s):
    """The examples' pre-solve pass: updateGeometry + updateForceState
    (e.g. examples/ex1_isoelasticity_3d_sc.cpp:104-105)."""
    s.forward()


def ct_plate(n_target=20000, a_over_w=0.35, steps=30, opening=2.5e-4, W=50.0, E=200e3, nu=0.3,
             crit=1e-3, max_iter=30, tol=1e-5):
    """2D hexagonal CT specimen (lpmg.geometry.ct_specimen) of brittle elastic
    particles: ParticleElastic bonds break once their stretch reaches `crit`
    (particle_elastic.h:52-68).  The pin particles (first particle ring of the
    upper half of the top hole and the lower half of the bottom hole) move
    +-opening*W in y per step and are held in x; the back-edge particle on the
    crack line is fixed.  Particles within 0.2 W of a pin centre are type 4, the
    undamaged type, as the pin regions of
    examples/ex4_isoelasticity_2d_crack.cpp:47-69.  Holding the pins in x keeps
    each half fully constrained once the crack has cut the ligament (the
    example's pin particles are held in y only, which leaves a separated half
    a free rigid motion and the stiffness singular)."""
    from . import geometry as geo
    ct = geo.ct_specimen(n_target, a_over_w, W=W)
    xyz, r = ct["xyz"], ct["radius"]
    x, y = xyz[:, 0], xyz[:, 1]
    (tx, ty), (bx, by) = ct["pin_top"], ct["pin_bottom"]
    dt = np.hypot(x - tx, y - ty)
    db = np.hypot(x - bx, y - by)
    R = ct["R"]
    top = np.nonzero((dt < R + 2.0 * r) & (y >= ty))[0]
    bottom = np.nonzero((db < R + 2.0 * r) & (y <= by))[0]
    fix = np.array([int(np.argmin(np.hypot(x - 0.0, y - ct["yc"])))])
    types = np.zeros(len(xyz), dtype=np.int32)
    types[(dt < 0.2 * W) | (db < 0.2 * W)] = 4
    d = opening * W
    step = [("disp", "fix", RELATIVE, "x", 0.0), ("disp", "fix", RELATIVE, "y", 0.0),
            ("disp", "top", RELATIVE, "x", 0.0), ("disp", "top", RELATIVE, "y", d),
            ("disp", "bottom", RELATIVE, "x", 0.0), ("disp", "bottom", RELATIVE, "y", -d)]
    p = Problem("ct_plate", lat.HEXAGON2D, r, lat.ELASTIC, xyz, (MAT_ELASTIC, [0.0, 1, E, nu, crit]),
                dict(fix=fix, top=top, bottom=bottom), _bc_steps(steps, step), types=types, undamaged=4,
                max_iter=max_iter, tol=tol,
                note=f"CT hex N={len(xyz)}, a/W={a_over_w}, W={W}, r={r:.4g}, E={E}, nu={nu} plane stress, "
                     f"critical bond stretch {crit}, pin opening {d:.4g} per step x {steps}")
    p.extra["bonds"] = ct["bonds"]
    p.extra["ct"] = dict(W=W, a=ct["a"], yc=ct["yc"], radius=r)
    return p


def bcc_polycrystal(n_cells=80, n_grains=200, radius=0.1, steps=10, strain_step=1e-4, C11=230e3, C12=135e3,
                    C44=117e3, seed=7):
    """3D BCC Voronoi polycrystal (lpmg.geometry.bcc_polycrystal) with cubic
    elasticity C11/C12/C44 (MPa) in every grain -- the grains differ only in
    the orientation of their lattice -- through setParticleProperty(L, C11,
    C12, C44, critical_bstrain) (particle_elastic.h:110-140).  Bottom face
    fixed, top face pulled `strain_step` of the height per step."""
    from . import geometry as geo
    pc = geo.bcc_polycrystal(n_cells, n_grains, radius, seed=seed)
    xyz = pc["xyz"]
    z = xyz[:, 2]
    a = 4.0 * radius / math.sqrt(3.0)
    bottom = np.nonzero(z < z.min() + 0.5 * a)[0]
    top = np.nonzero(z > z.max() - 0.5 * a)[0]
    d = strain_step * (z.max() - z.min())
    step = [("disp", "bottom", RELATIVE, "x", 0.0), ("disp", "bottom", RELATIVE, "y", 0.0),
            ("disp", "bottom", RELATIVE, "z", 0.0), ("disp", "top", RELATIVE, "z", d)]
    p = Problem("bcc_polycrystal", lat.BCC3D, radius, lat.ELASTIC, xyz,
                (MAT_ELASTIC_CUBIC, [0.0, C11, C12, C44, 1.0]), dict(bottom=bottom, top=top),
                _bc_steps(steps, step),
                note=f"BCC polycrystal {n_cells}^3 cells, {n_grains} grains (seed {seed}), N={len(xyz)}, "
                     f"cubic C11/C12/C44 = {C11}/{C12}/{C44} MPa, tension {strain_step}/step x {steps}")
    p.extra["grain"] = pc["grain"]
    return p


def fatigue_gage(radius=0.12, d_max=0.043927, R=0.0, n_cycles=1_000_000, nonlocal_L=8.0, cutoff_ratio=1.5,
                 tau=1e-4, hole=0.0, E=71.7e3, nu=0.306, A=2.003e-5, B=2.833, d=0.5, d_thres=1.0,
                 f_lmt_ratio=1.108, max_iter=30, tol=1e-5):
    """2D hexagonal gage section of FatigueHCF particles under cyclic uniaxial
    displacement, as examples/ex5_isoelasticity_2d_fatigue_gage.cpp: the
    5.08 x 14.7 mm plate (r = 0.12), Al 7075-T651 plane stress (E = 71.7 GPa,
    nu = 0.306) with the fatigue law A = 2.003e-5, B = 2.833, d = 0.5; the
    left edge (x < 0.05) is held in x, the grip particles (y > 13.7, y < 1) are
    of the undamaged type 4, the bottom grip is held in y and the top grip
    moves through the spectrum low, high, low, ... (R * d_max, d_max) one step
    per half cycle.  `hole` > 0 cuts a central circular hole of that radius,
    which concentrates the strain energy so that damage localises.

    Problem.steps holds the static step to the spectrum's first value
    (solveProblemStatic); extra["fatigue"] holds what solve_cyclic needs."""
    from . import fatigue as F
    box = [0.0, 5.08, 0.0, 12.7 + 2, 0.0, 4.8]
    xyz = lat.create_plate_hex2d(box, radius)
    if hole > 0:
        keep = np.hypot(xyz[:, 0] - 2.54, xyz[:, 1] - 7.35) >= hole
        xyz = xyz[keep]
    x, y = xyz[:, 0], xyz[:, 1]
    types = np.zeros(len(xyz), dtype=np.int32)
    left = np.nonzero(x < 0.05)[0]
    types[left] = 1
    top = np.nonzero(y > 14.7 / 2 + 12.7 / 2)[0]
    types[top] = 2
    bottom = np.nonzero(y < 14.7 / 2 - 12.7 / 2)[0]
    types[bottom] = 3
    undamaged = 4
    types[(types == 2) | (types == 3)] = undamaged
    d_min = R * d_max
    static = [[("disp", "left", RELATIVE, "x", 0.0), ("disp", "top", RELATIVE, "y", d_min),
               ("disp", "bottom", RELATIVE, "y", 0.0)]]
    p = Problem("fatigue_gage", lat.HEXAGON2D, radius, lat.FATIGUE_HCF, xyz,
                (MAT_FATIGUE, [nonlocal_L, 1, E, nu, A, B, d, d_thres, f_lmt_ratio]),
                dict(left=left, top=top, bottom=bottom), static, types=types, undamaged=undamaged,
                max_iter=max_iter, tol=tol,
                note=f"hex gage 5.08x14.7 r={radius} N={len(xyz)}" + (f" hole r={hole}" if hole else "") +
                     f", fatigue A={A} B={B} d={d}, d_max={d_max} R={R}, L={nonlocal_L}, tau={tau}")
    p.extra["nonlocal_ratio"] = cutoff_ratio
    p.extra["fatigue"] = dict(ltype=F.UNIAXIAL_DISP, groups=("left", "top", "bottom"),
                              spectrum=F.constant_amplitude(d_min, d_max, n_cycles), n_interval=1,
                              tau=tau, t_mode=F.LINEAR)
    return p
