"""Parity of cyclic fatigue loading (SolverFatigue over FatigueHCF particles,
solver_fatigue.h) on the HIP path with the CPU oracle, through the same host
driver (lpmg.fatigue).

Every simulated cycle is compared: each substep's displacements at the
north-star bar (L2 error <= 1e-8 of the peak-load displacement norm, Newton
run to 1e-11), damage and the cycle energy
extremes to 1e-8 relative, the set of fully damaged particles and the
"noticeable new damage" return exactly.
"""
import numpy as np
import pytest

import lpmg
from lpmg import fatigue as F
from lpmg import problems as P
from oracle import Oracle

pytestmark = pytest.mark.gpu

GPU_OPTS = dict(precond=1, rtol=1e-24, atol=0.0)
ORC_OPTS = dict(precond=0, rtol=1e-24, atol=0.0)


def _run(backend, prob, opts, n_sim):
    fz = prob.extra["fatigue"]
    s, g = P.build(backend, prob)
    s.search_nonlocal(prob.extra["nonlocal_ratio"])
    s.forward()
    sf = F.SolverFatigue(s, g, prob.undamaged, t_mode=fz["t_mode"], tau=fz["tau"], max_iter=prob.max_iter,
                         tol=prob.tol, opts=opts)
    sf.solve_static(prob.steps)
    dNdt = F.dNdt_linear(fz["tau"])
    spec = fz["spectrum"][: 2 * n_sim * dNdt + 1]
    snaps, sub = [], []

    def on_substep(i, st):
        sub.append(s.get("XYZ").reshape(-1, 3))

    def on_cycle(rec):
        snaps.append(dict(N=rec["N"], new=rec["new_damaged"], xyz=list(sub), damage=s.get("DAMAGE"),
                          sv=s.get("STATE_VAR").reshape(-1, 7)[:, :3]))
        sub.clear()

    sf.solve_cyclic(fz["ltype"], spec, fz["n_interval"], fz["groups"], on_cycle=on_cycle, on_substep=on_substep)
    x0 = s.get("XYZ_INITIAL").reshape(-1, 3)
    return snaps, x0


def _compare(prob, n_sim, gpu_opts=GPU_OPTS, gpu_tol=1e-11):
    import copy
    pg, po = copy.copy(prob), copy.copy(prob)
    pg.tol, po.tol = gpu_tol, 1e-11
    print(f"fatigue parity: {prob.note}, {n_sim} cycles, opts {gpu_opts}", flush=True)
    a, x0 = _run(lpmg.System, pg, gpu_opts, n_sim)
    b, _ = _run(Oracle, po, ORC_OPTS, n_sim)
    assert len(a) == len(b) == n_sim
    for k, (ga, ob) in enumerate(zip(a, b)):
        assert ga["N"] == ob["N"] and ga["new"] == ob["new"], k
        # every substep's displacements against the peak-load displacement norm
        # (the unloaded states carry almost no displacement to be relative to)
        scale = max(np.linalg.norm(x - x0) for x in ob["xyz"])
        for xa, xb in zip(ga["xyz"], ob["xyz"]):
            assert np.linalg.norm(xa - xb) <= 1e-8 * scale, k
        np.testing.assert_array_equal(ga["damage"] == 1, ob["damage"] == 1)
        np.testing.assert_allclose(ga["damage"], ob["damage"], rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(ga["sv"], ob["sv"], rtol=1e-8, atol=1e-14)
    return a, b


def test_holed_gage_cycles_match_oracle():
    # a central hole localises the damage: particles reach the threshold at
    # different cycles and the crack runs across the gage
    prob = P.fatigue_gage(hole=1.0, nonlocal_L=0.6, tau=1e-4, n_cycles=40000)
    a, _ = _compare(prob, 4)
    n_full = [int((s["damage"] == 1).sum()) for s in a]
    assert n_full[0] > 0 and n_full[-1] > n_full[0]


def test_reference_gage_settings_match_oracle():
    # examples/ex5_isoelasticity_2d_fatigue_gage.cpp as written (L = 8, R = 0):
    # near-uniform damage growth, about 0.012 per simulated cycle of 10^4
    prob = P.fatigue_gage(n_cycles=30000)
    a, _ = _compare(prob, 3)
    d = a[-1]["damage"]
    assert 0.02 < d.max() < 0.05


def test_mixed_precision_cycles_match_oracle():
    # FP32 inner solves inside the FP64 Newton iteration reach the same
    # converged states (Newton to 1e-9, above the mixed iteration's floor), so
    # the fatigue history is unchanged
    prob = P.fatigue_gage(hole=1.0, nonlocal_L=0.6, tau=1e-4, n_cycles=20000)
    _compare(prob, 2, gpu_opts=dict(precond=1, rtol=1e-4, atol=0.0, precision=1), gpu_tol=1e-9)
