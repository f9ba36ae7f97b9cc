"""Cyclic (high-cycle fatigue) loading of FatigueHCF particles.

`SolverFatigue` restates LPM-CPP's SolverFatigue (include/solver_fatigue.h) and
`load_cycle` the fatigue load steps of include/load_step_fatigue.h, over any
backend with the System interface (lpmg.System on the GPU; the oracle's Oracle
has the same methods).  One simulated cycle applies the load and unload
increments of the spectrum, each as a quasi-static step (force BCs, then
displacement BCs, stiffness assembly, Newton) followed by the particle
state-variable update; then the damage of every damageable particle grows by
its nonlocal damage rate times dN/dt, the number of real cycles the simulated
one stands for, and the converged state is stored.
"""
import math

import numpy as np

from .problems import RELATIVE, apply_step

# TimeMapMode (include/lpm.h:85-89)
LINEAR, EXPONENTIAL = 0, 1
# FatigueLoadType (include/load_step_fatigue.h:13-19)
UNIAXIAL_DISP, UNIAXIAL_FORCE, DOGBONE_DISP, CT_FORCE = 0, 1, 2, 3


def constant_amplitude(low, high, n_cycles):
    """A constant-amplitude spectrum as loading/generate_constant_loading.ipynb
    writes them: `low`, then `high, low` once per cycle."""
    out = np.empty(2 * n_cycles + 1)
    out[0::2] = low
    out[1::2] = high
    return out


def read_load(path):
    """SolverFatigue::readLoad (solver_fatigue.h:41-63): the numbers of the file
    in order, up to the first token that is not one."""
    vals = []
    with open(path) as f:
        for tok in f.read().split():
            try:
                vals.append(float(tok))
            except ValueError:
                break
    return np.array(vals)


def dNdt_linear(tau):
    """func_dNdt_linear (solver_fatigue.h:65-68): (int)(1 / tau)."""
    return int(1 / tau)


def dNdt_exponential(tau, t):
    """func_dNdt_exponental (solver_fatigue.h:70-73): (int)(exp(t / tau) / tau)."""
    return int(1 / tau * math.exp(t / tau))


def _fatigue_step(ltype, v, groups):
    """One load step of load_step_fatigue.h:22-85 with increment v (group names)."""
    if ltype == UNIAXIAL_DISP:    # LoadUniaxialDisp(dy, left, top, bottom)
        left, top, bottom = groups
        return [("disp", left, RELATIVE, "x", 0.0), ("disp", top, RELATIVE, "y", v),
                ("disp", bottom, RELATIVE, "y", 0.0)]
    if ltype == UNIAXIAL_FORCE:   # LoadUniaxialForce(fy, left, top, bottom)
        left, top, bottom = groups
        return [("disp", left, RELATIVE, "x", 0.0), ("force", top, RELATIVE, 0.0, v, 0.0),
                ("disp", bottom, RELATIVE, "y", 0.0)]
    if ltype == DOGBONE_DISP:     # LoadDogBoneDisp(0, dy, 0, group1, group2)
        g1, g2 = groups
        return [("disp", g1, RELATIVE, "x", 0.0), ("disp", g1, RELATIVE, "y", v),
                ("disp", g1, RELATIVE, "z", 0.0), ("disp", g2, RELATIVE, "x", 0.0),
                ("disp", g2, RELATIVE, "y", 0.0), ("disp", g2, RELATIVE, "z", 0.0)]
    if ltype == CT_FORCE:         # LoadCTForce(fy, mid, top, bottom)
        mid, top, bottom = groups
        return [("disp", mid, RELATIVE, "y", 0.0), ("disp", top, RELATIVE, "x", 0.0),
                ("disp", bottom, RELATIVE, "x", 0.0), ("force", top, RELATIVE, 0.0, v, 0.0),
                ("force", bottom, RELATIVE, 0.0, -v, 0.0)]
    raise ValueError(f"unknown fatigue load type {ltype}")


def load_cycle(ltype, spectrum, N, n_interval, groups):
    """SolverFatigue::generateLoadCycle (solver_fatigue.h:128-180): cycle N goes
    from spectrum[2N] to spectrum[2N+1] in n_interval equal steps and back to
    spectrum[2N+2] in n_interval more."""
    f0, f1, f2 = spectrum[2 * N], spectrum[2 * N + 1], spectrum[2 * N + 2]
    up = [_fatigue_step(ltype, (f1 - f0) / n_interval, groups) for _ in range(n_interval)]
    down = [_fatigue_step(ltype, (f2 - f1) / n_interval, groups) for _ in range(n_interval)]
    return up + down


class SolverFatigue:
    """SolverFatigue (solver_fatigue.h:18-38) over system `s` with named groups `gids`.

    opts are the linear-solver options of System.newton (precond, rtol, atol,
    maxit, precision); precision 1 assembles the FP32 copy of the stiffness
    for the mixed-precision inner solve."""

    def __init__(self, s, gids, undamaged, t_mode=LINEAR, tau=1e-4, max_iter=30, tol=1e-5, opts=None):
        self.s, self.gids = s, gids
        self.undamaged, self.t_mode, self.tau = undamaged, t_mode, tau
        self.max_iter, self.tol = max_iter, tol
        self.opts = dict(opts or {})

    def _assemble(self):
        p = int(self.opts.get("precision", 0))
        if p:
            self.s.assemble(p)
        else:
            self.s.assemble()

    def solve_step(self, step, scale=1.0):
        """solveProblemStep (solver_fatigue.h:277-301): BCs, stiffness, Newton;
        converged when Newton returns before max_iter."""
        apply_step(self.s, self.gids, step, scale)
        self._assemble()
        n, st = self.s.newton(self.max_iter, self.tol, **self.opts)
        return n < self.max_iter, st

    def solve_static(self, steps, max_halvings=8):
        """solveProblemStatic (solver_fatigue.h:243-275): each step to
        convergence, halving a step that fails (state reset to the last
        converged one, positions included); after each converged step the state
        variables are updated and stored.  Returns the Newton stats per step."""
        queue = [(st, 1.0) for st in steps]
        out, i, halvings = [], 0, 0
        while i < len(queue):
            step, scale = queue[i]
            ok, st = self.solve_step(step, scale)
            if not ok:
                if halvings >= max_halvings:
                    raise RuntimeError(f"static step {i} did not converge after {halvings} halvings")
                halvings += 1
                self.s.reset_state(True)
                queue[i] = (step, scale / 2)
                queue.insert(i, (step, scale / 2))
                continue
            self.s.update_state_var()
            self.s.store_state()
            out.append(st)
            i += 1
        return out

    def solve_one_cycle(self, steps, dNdt, on_substep=None):
        """solveProblemOneCycle (solver_fatigue.h:220-241).  A substep that does
        not converge is kept, as in the reference.  on_substep(i, stats) runs
        after each substep's state-variable update.  Returns (any noticeable
        new damage, Newton stats per substep)."""
        stats = []
        for i, step in enumerate(steps):
            _, st = self.solve_step(step)
            self.s.update_state_var()
            stats.append(st)
            if on_substep:
                on_substep(i, st)
        new_damaged = self.s.update_fatigue_damage(self.undamaged, float(dNdt))
        self.s.store_state()
        self.s.forward()
        return new_damaged, stats

    def solve_cyclic(self, ltype, spectrum, n_interval, groups, on_cycle=None, on_substep=None):
        """solveProblemCyclic (solver_fatigue.h:182-218): simulated cycles until
        the real cycle count N reaches the spectrum's (len - 1) / 2.  Returns one
        record per simulated cycle: N after it, dN/dt, the mean of state_var[0]
        (the greatest strain energy density so far) over damageable particles,
        whether damage grew noticeably, and the Newton stats of its substeps."""
        n_cycles = 0.5 * (len(spectrum) - 1)
        t, N = 0, 0
        history = []
        while True:
            steps = load_cycle(ltype, spectrum, N, n_interval, groups)
            if self.t_mode == LINEAR:
                dNdt = dNdt_linear(self.tau)
            else:
                dNdt = dNdt_exponential(self.tau, t)
                t += 1
            new_damaged, stats = self.solve_one_cycle(steps, dNdt, on_substep)
            sv = self.s.get("STATE_VAR").reshape(-1, 7)
            damageable = self.s.get("TYPE").astype(np.int64) != self.undamaged
            # the reference's running sum in particle order (solver_fatigue.h:202-212)
            ave_u = float(np.cumsum(sv[damageable, 0])[-1] / damageable.sum()) if damageable.any() \
                else float("nan")
            N += int(dNdt)
            rec = dict(N=N, dNdt=dNdt, ave_energy=ave_u, new_damaged=new_damaged, stats=stats)
            history.append(rec)
            if on_cycle:
                on_cycle(rec)
            if not N < n_cycles:
                break
        return history
