"""Pins the oracle's fatigue path (ParticleFatigueHCF, SolverFatigue) and the
host-side cyclic driver (lpmg.fatigue) before they are trusted as the checker.

The reference ships no fatigue fixtures (its load spectra are generated by a
notebook and absent), so each check comes from mathematics: the strain energy
density of an affine strain against the continuum 1/2 eps:C:eps, the cycle
extremes of a known load history, the damage-rate law evaluated on those
closed-form energies with B != d and a nonzero cycle minimum, the nonlocal
average as a normalised weighting, and the damage growth of a uniformly
strained gage section.  Each test names the mistake it would catch.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from lpmg import lattice as lat
from lpmg import problems as P
from lpmg import fatigue as F

R = 0.1
E, NU = 200e3, 0.3
INT_MAX = 2147483647.0


def _deep(o):
    """Particles whose connections all have full bond shells (bulk interior)."""
    cptr, conns, _ = o.conns()
    bptr, _, _ = o.bonds()
    full = np.diff(bptr) == np.diff(bptr).max()
    return np.array([full[conns[cptr[i]:cptr[i + 1]]].all() for i in range(o.N)])


def _fatigue_oracle(L, xyz, A=1e-3, B=2.0, d=0.5, thr=1.0, ps=1, nonlocal_L=0.0):
    o = Oracle(L, R, lat.FATIGUE_HCF, xyz)
    o.set_material(P.MAT_FATIGUE, [nonlocal_L, ps, E, NU, A, B, d, thr, 1.108])
    return o


def _moduli(ps):
    if ps:
        return E / (1 - NU * NU), E * NU / (1 - NU * NU)
    return E * (1 - NU) / (1 + NU) / (1 - 2 * NU), E * NU / (1 + NU) / (1 - 2 * NU)


def test_state_starts_at_cycle_sentinels():
    # ParticleFatigueHCF starts with state_var = {0, INT_MAX, 0}: a zero cycle
    # minimum would make the first cycle's range the peak energy itself
    o = _fatigue_oracle(lat.HEXAGON2D, lat.hex_plate(6, 6, R))
    sv = o.get("STATE_VAR").reshape(-1, 7)
    assert np.all(sv[:, 1] == INT_MAX) and np.all(sv[:, [0, 2, 3, 4, 5, 6]] == 0)
    o2 = Oracle(lat.HEXAGON2D, R, lat.ELASTIC_DAMAGE, lat.hex_plate(6, 6, R))
    assert np.all(o2.get("STATE_VAR") == 0)


def test_material_and_particle_type_must_agree():
    o = Oracle(lat.HEXAGON2D, R, lat.ELASTIC, lat.hex_plate(4, 4, R))
    with pytest.raises(RuntimeError):
        o.set_material(P.MAT_FATIGUE, [0, 1, E, NU, 1e-3, 2.0, 0.5, 1.0, 1.108])
    f = Oracle(lat.HEXAGON2D, R, lat.FATIGUE_HCF, lat.hex_plate(4, 4, R))
    with pytest.raises(RuntimeError):
        f.set_material(P.MAT_ELASTIC, [0, 1, E, NU, 1.0])


@pytest.mark.parametrize("name,L,gen,ps", [("hex", lat.HEXAGON2D, lambda: lat.hex_plate(18, 18, R), 1),
                                           ("sc", lat.SC3D, lambda: lat.sc_block(9, 9, 9, R), 0)])
def test_energy_of_affine_strain_is_continuum_energy(name, L, gen, ps):
    # calcEqEnergy's energy_total in the bulk equals 1/2 eps:C:eps for uniaxial
    # and equibiaxial strain: a dropped volumetric (Tv) term fails the biaxial
    # case, a wrong modified volume V_m or a missing 1/2 fails both
    xyz = gen()
    o = _fatigue_oracle(L, xyz, ps=ps)
    C11, C12 = _moduli(ps)
    deep = _deep(o)
    assert deep.sum() > 4
    eps = 1e-5
    for Fm, want in [(np.diag([1 + eps, 1, 1]), 0.5 * C11 * eps ** 2),
                     (np.diag([1 + eps, 1 + eps, 1]), (C11 + C12) * eps ** 2)]:
        o.set("XYZ", xyz @ Fm.T)
        o.forward()
        o.update_state_var()
        e = o.get("STATE_VAR").reshape(-1, 7)[deep, 0]
        np.testing.assert_allclose(e, want, rtol=1e-4)
        o.reset_state(True)


def test_cycle_extremes_rate_law_and_reset():
    # A cycle through stretches a1 < a2 > a3 in the bulk: sv[1] is the least
    # and sv[2] the greatest energy of the cycle, sv[0] the greatest ever, and
    # the local rate is A * sv0^B * (sv2 - sv1)^d on the closed-form energies;
    # B != d and a nonzero minimum catch swapped exponents or a range taken
    # from zero.  updateFatigueDamage then restores the cycle sentinels and
    # keeps the history maximum.
    xyz = lat.hex_plate(18, 18, R)
    A, B, d = 3e-2, 2.5, 0.7
    o = _fatigue_oracle(lat.HEXAGON2D, xyz, A=A, B=B, d=d)
    C11, _ = _moduli(1)
    deep = _deep(o)
    amps = [4e-6, 1e-5, 6e-6]
    for a in amps:
        o.set("XYZ", xyz @ np.diag([1 + a, 1, 1]).T)
        o.forward()
        o.update_state_var()
    sv = o.get("STATE_VAR").reshape(-1, 7)[deep]
    U = [0.5 * C11 * a * a for a in amps]
    np.testing.assert_allclose(sv[:, 1], min(U), rtol=1e-4)
    np.testing.assert_allclose(sv[:, 2], max(U), rtol=1e-4)
    np.testing.assert_allclose(sv[:, 0], max(U), rtol=1e-4)
    rate = A * max(U) ** B * (max(U) - min(U)) ** d
    np.testing.assert_allclose(o.get("DDOT_LOCAL")[deep], rate, rtol=1e-3)
    o.search_nonlocal(1.5)  # nonlocal_L = 0: each particle averages only itself
    o.update_fatigue_damage(-1, 1.0)
    sv2 = o.get("STATE_VAR").reshape(-1, 7)
    assert np.all(sv2[:, 1] == INT_MAX) and np.all(sv2[:, 2] == 0)
    np.testing.assert_array_equal(sv2[deep, 0], sv[:, 0])
    # a second cycle at lower amplitude keeps the history maximum
    o.set("XYZ", xyz @ np.diag([1 + 2e-6, 1, 1]).T)
    o.forward()
    o.update_state_var()
    sv3 = o.get("STATE_VAR").reshape(-1, 7)[deep]
    np.testing.assert_allclose(sv3[:, 0], max(U), rtol=1e-4)
    np.testing.assert_allclose(sv3[:, 1], sv3[:, 2])
    assert np.all(o.get("DDOT_LOCAL")[deep] == 0)  # no range yet


def _nonlocal_system(nonlocal_L=0.35):
    xyz = lat.hex_plate(20, 20, R)
    o = _fatigue_oracle(lat.HEXAGON2D, xyz, nonlocal_L=nonlocal_L)
    o.search_nonlocal(1.5)
    return o, xyz


def test_nonlocal_average_reproduces_constant_and_linear_fields():
    # The phi-weighted average is normalised (a constant rate stays exactly that
    # rate everywhere) and symmetric (a linear field is reproduced where the
    # neighbourhood is complete); weighting by the wrong particle's volume or
    # distance breaks one of them
    o, xyz = _nonlocal_system()
    c = 2.5e-6
    o.set("DDOT_LOCAL", np.full(o.N, c))
    assert not o.update_fatigue_damage(-1, 100.0)   # 2.5e-4 < 1e-3: not noticeable
    np.testing.assert_allclose(o.get("DDOT_NONLOCAL"), c, rtol=1e-14)
    np.testing.assert_allclose(o.get("DAMAGE"), 100.0 * c, rtol=1e-14)
    o2, _ = _nonlocal_system()
    x = xyz[:, 0]
    o2.set("DDOT_LOCAL", 1e-6 * (1 + x))
    o2.update_fatigue_damage(-1, 1.0)
    cut = 1.5 * 0.35 + 2 * R
    inner = ((x > x.min() + cut) & (x < x.max() - cut) & (xyz[:, 1] > xyz[:, 1].min() + cut) &
             (xyz[:, 1] < xyz[:, 1].max() - cut))
    assert inner.sum() > 20
    np.testing.assert_allclose(o2.get("DDOT_NONLOCAL")[inner], 1e-6 * (1 + x[inner]), rtol=2e-3)


def test_damage_threshold_undamaged_type_and_return_value():
    # damage += rate * dN/dt; reaching d_thres sets 1; the undamaged type is
    # neither updated nor averaged over; the return says whether any damage
    # moved by more than 1e-3 since the last stored state
    o, xyz = _nonlocal_system()
    types = np.zeros(o.N, dtype=np.int32)
    types[xyz[:, 1] > xyz[:, 1].max() - 0.5] = 4
    o.set_types(types)
    o.set("DDOT_LOCAL", np.where(xyz[:, 0] < 1.5, 1e-4, 1e-7))
    assert o.update_fatigue_damage(4, 1e4)
    dmg = o.get("DAMAGE")
    assert np.all(dmg[types == 4] == 0)
    nl = o.get("DDOT_NONLOCAL")
    want = np.where(nl * 1e4 >= 1.0, 1.0, nl * 1e4)
    np.testing.assert_array_equal(dmg[types != 4], want[types != 4])
    assert (dmg == 1).sum() > 0 and ((dmg > 0) & (dmg < 1)).sum() > 0


def test_static_damage_leaves_fatigue_particles_alone():
    # the fatigue particle keeps Particle::updateParticleStaticDamage's no-op
    o, _ = _nonlocal_system()
    o.set("DDOT_LOCAL", np.full(o.N, 0.5))
    assert not o.update_static_damage(-1)
    assert np.all(o.get("DAMAGE") == 0)


def test_time_maps_spectrum_and_load_cycle():
    assert F.dNdt_linear(1e-4) == 10000
    assert F.dNdt_linear(1e-5) == 99999        # (int)(1/1e-5) truncates 99999.99999999999
    assert F.dNdt_exponential(0.1, 0) == 10
    assert F.dNdt_exponential(0.5, 1) == int(2 * math.e ** 2) == 14
    spec = F.constant_amplitude(0.1, 1.0, 3)
    np.testing.assert_array_equal(spec, [0.1, 1.0, 0.1, 1.0, 0.1, 1.0, 0.1])
    steps = F.load_cycle(F.UNIAXIAL_DISP, np.array([0.0, 1.0, 0.25, 2.0, 0.0]), 1, 2,
                         ("l", "t", "b"))
    assert len(steps) == 4
    assert [s[1][4] for s in steps] == [0.875, 0.875, -1.0, -1.0]
    assert steps[0][0] == ("disp", "l", P.RELATIVE, "x", 0.0) and steps[0][2][1:4] == ("b", P.RELATIVE, "y")
    ct = F.load_cycle(F.CT_FORCE, np.array([1.0, 3.0, 1.0]), 0, 1, ("m", "t", "b"))
    assert ct[0][3] == ("force", "t", P.RELATIVE, 0.0, 2.0, 0.0)
    assert ct[0][4] == ("force", "b", P.RELATIVE, 0.0, -2.0, 0.0)
    assert ct[1][3][4] == -2.0
    db = F.load_cycle(F.DOGBONE_DISP, np.array([0.0, 0.5, 0.0]), 0, 1, ("g1", "g2"))
    assert [b[3] for b in db[0]] == ["x", "y", "z", "x", "y", "z"] and db[0][1][4] == 0.5


def test_read_load(tmp_path):
    p = tmp_path / "spec.txt"
    p.write_text("0\n0.043927\n0\n0.043927\n0\nend\n7\n")
    np.testing.assert_array_equal(F.read_load(str(p)), [0, 0.043927, 0, 0.043927, 0])


def test_gage_damage_rate_matches_closed_form():
    # examples/ex5_isoelasticity_2d_fatigue_gage.cpp at R = 0.5: the gage
    # strains uniformly by d/12.7 between the grips, so after one simulated
    # cycle the bulk damage is A*Umax^B*(Umax - Umin)^d * dN/dt with
    # U = 1/2 E eps^2 (uniaxial plane stress) and Umin = R^2 Umax.  Swapping B
    # and d changes it by 0.75^(B-d) ~ 0.5; a rate taken from zero by 1/0.75^d.
    prob = P.fatigue_gage(R=0.5, n_cycles=1, nonlocal_L=8.0)
    o, g = P.build(Oracle, prob)
    o.search_nonlocal(prob.extra["nonlocal_ratio"])
    o.forward()
    fz = prob.extra["fatigue"]
    sf = F.SolverFatigue(o, g, prob.undamaged, tau=fz["tau"], max_iter=prob.max_iter, tol=prob.tol,
                         opts=dict(precond=1, rtol=1e-12, atol=1e-12))
    sf.solve_static(prob.steps)
    hist = sf.solve_cyclic(fz["ltype"], fz["spectrum"], fz["n_interval"], fz["groups"])
    assert [h["N"] for h in hist] == [10000]
    xyz = prob.xyz
    mid = (np.abs(xyz[:, 1] - 7.35) < 3.0) & (np.abs(xyz[:, 0] - 2.54) < 1.5)
    Umax = 0.5 * 71.7e3 * (0.043927 / 12.7) ** 2
    Umin = 0.25 * Umax
    want = 2.003e-5 * Umax ** 2.833 * (Umax - Umin) ** 0.5 * 10000
    got = np.median(o.get("DAMAGE")[mid])
    assert abs(got / want - 1) < 0.05, (got, want)
    sv = o.get("STATE_VAR").reshape(-1, 7)
    assert abs(np.median(sv[mid, 0]) / Umax - 1) < 0.05
