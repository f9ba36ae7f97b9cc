"""Pins the oracle's damage branches and the unit-cell data before they are
trusted as the checker.

* UnitCell data (include/unit_cell.h:44-122): el_mapping, cutoffs, volume,
  nneighbors and dim of the five lattices against tests/golden/unit_cells.json,
  which tests/golden/make_unit_cells.py evaluates from the reference header's
  own expressions -- bit for bit, for the oracle and for the product library.
* ElasticDamage rate law (particle_elastic_damage.h:150-161): Ddot_local =
  (U - U_last) / k1 when f = U - k1 D - k0 > 0 and U - U_last > 0, else 0,
  with U the closed-form strain energy density of an affine strain; every gate
  is driven both ways.
* The static-damage increment (particle_elastic_damage.h:163-173,
  particle_j2plasticity.h:142-152): damage += Ddot_nonlocal, 1 at the
  threshold, the undamaged type skipped, and the |D - D_last| > 1e-3 return,
  on hand-set damages whose sums are exact in binary.
* The bond-damage update of the damage models (particle_elastic_damage.h:80-84,
  particle_j2plasticity.h:71-73): bdamage = max(bdamage, D_i, D_j) and the
  bond force scaled by (1 - bdamage), on hand-set particle and bond damages.

Each test names the mistake it would catch.
"""
import json
import os

import numpy as np
import pytest

from oracle import Oracle, unit_cell
from lpmg import lattice as lat
from lpmg import problems as P

R = 0.1
E, NU = 200e3, 0.3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "unit_cells.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)["lattices"]


def _check_cell(out, rec, r):
    rad = rec["radii"][repr(r)]
    assert int(out[0]) == rec["dim"] and int(out[1]) == rec["nneighbors"]
    assert out[2] == float.fromhex(rad["volume"])
    assert [out[3], out[4]] == [float.fromhex(v) for v in rad["cutoff"]]
    assert list(out[5:14]) == [float.fromhex(v) for v in rec["el_mapping"]]


@pytest.mark.parametrize("lattice", range(5))
@pytest.mark.parametrize("r", [0.05, 0.1, 0.12, 0.13, 0.5, 1.0])
def test_unit_cell_data_equals_reference_header(lattice, r):
    # a mistyped el_mapping entry (the BCC/FCC tables are pinned by nothing
    # else: the finite-difference tests use the oracle's own Kn/Tv), a wrong
    # cutoff or volume expression fails here, in the oracle and in the product
    import lpmg

    rec = _golden()[str(lattice)]
    _check_cell(unit_cell(lattice, r), rec, r)
    _check_cell(lpmg.unit_cell(lattice, r), rec, r)


def test_unit_cell_fixture_covers_every_lattice():
    g = _golden()
    assert sorted(g) == ["0", "1", "2", "3", "4"]
    assert [g[str(k)]["function"] for k in range(5)] == ["setSquare", "setHexagon", "setSimpleCubic",
                                                          "setFCC", "setBCC"]


def _deep(o):
    """Particles whose connections all have full bond shells (bulk interior)."""
    cptr, conns, _ = o.conns()
    bptr, _, _ = o.bonds()
    full = np.diff(bptr) == np.diff(bptr).max()
    return np.array([full[conns[cptr[i]:cptr[i + 1]]].all() for i in range(o.N)])


# --------------------------------------------------------- ElasticDamage rate law

C11_PS = E / (1 - NU * NU)  # plane-stress uniaxial modulus of the hexagonal plate


def _damage_plate(k0, k1, nonlocal_L=0.0, thr=1.0, n=18):
    xyz = lat.hex_plate(n, n, R)
    o = Oracle(lat.HEXAGON2D, R, lat.ELASTIC_DAMAGE, xyz)
    o.set_material(P.MAT_DAMAGE, [nonlocal_L, 1, E, NU, k0, k1, 1.0, thr])
    return o, xyz


def _strain(o, xyz, eps):
    o.set("XYZ", xyz @ np.diag([1 + eps, 1, 1]).T)
    o.forward()
    o.update_state_var()
    return o.get("STATE_VAR").reshape(-1, 7)[:, 0], o.get("DDOT_LOCAL")


def test_elastic_damage_rate_law_both_gates():
    eps1 = 1e-5
    U1 = 0.5 * C11_PS * eps1 ** 2
    k1 = 0.01
    # f > 0 and delta > 0 (U_last = 0 at the start): Ddot = U / k1.  A rate of
    # U * k1, of U / k0, or measured from the damage surface fails here.
    o, xyz = _damage_plate(k0=0.5 * U1, k1=k1)
    deep = _deep(o)
    assert deep.sum() > 10
    U, dd = _strain(o, xyz, eps1)
    np.testing.assert_allclose(U[deep], U1, rtol=1e-4)
    np.testing.assert_allclose(dd[deep], U1 / k1, rtol=1e-4)
    # f <= 0 through k0 alone, delta > 0: no rate (a gate of "f > 0 or delta > 0" fails)
    o, xyz = _damage_plate(k0=2.0 * U1, k1=k1)
    _, dd = _strain(o, xyz, eps1)
    assert np.all(dd[deep] == 0)


def test_elastic_damage_surface_includes_the_damage_term():
    eps1 = 1e-5
    U1 = 0.5 * C11_PS * eps1 ** 2
    k1 = 1e-4
    # f = U - k1 D - k0 with k1 D = 0.75 U, k0 = 0.5 U: f < 0, no rate.  A
    # surface without the k1 D term, or with +k1 D, gives a rate here.
    o, xyz = _damage_plate(k0=0.5 * U1, k1=k1)
    deep = _deep(o)
    o.set("DAMAGE", np.full(o.N, 0.75 * U1 / k1))
    _, dd = _strain(o, xyz, eps1)
    assert np.all(dd[deep] == 0)
    # k1 D = 0.25 U: f > 0 and the rate is still delta / k1 (the damage moves
    # the surface, not the rate)
    o, xyz = _damage_plate(k0=0.5 * U1, k1=k1)
    o.set("DAMAGE", np.full(o.N, 0.25 * U1 / k1))
    _, dd = _strain(o, xyz, eps1)
    np.testing.assert_allclose(dd[deep], U1 / k1, rtol=1e-4)


def test_elastic_damage_rate_is_measured_from_the_stored_state():
    eps1 = 1e-5
    U1 = 0.5 * C11_PS * eps1 ** 2
    k1 = 0.02
    o, xyz = _damage_plate(k0=0.1 * U1, k1=k1)
    deep = _deep(o)
    U2, _ = _strain(o, xyz, 2 * eps1)   # 4 U1
    o.store_state()                     # state_var_last[0] = 4 U1
    # unloading to eps1 while f > 0: delta < 0, no rate (a delta measured from
    # zero, or a gate on f alone, gives one)
    _, dd = _strain(o, xyz, eps1)
    assert np.all(dd[deep] == 0)
    # reloading to 3 eps1: delta = 9 U1 - 4 U1 = 5 U1
    _, dd = _strain(o, xyz, 3 * eps1)
    np.testing.assert_allclose(dd[deep], 5 * U1 / k1, rtol=1e-4)
    np.testing.assert_allclose(U2[deep], 4 * U1, rtol=1e-4)


# --------------------------------------------------------- static damage increment

def _j2_plate(nonlocal_L, crit):
    xyz = lat.hex_plate(14, 14, R)
    o = Oracle(lat.HEXAGON2D, R, lat.J2PLASTICITY, xyz)
    # nonlocal_L, plane_stress, E, mu, sigmay, xi, H, A, critical_bstrain
    o.set_material(P.MAT_J2, [nonlocal_L, 1, E, NU, 250.0, 0.0, 2000.0, 0.0, crit])
    return o, xyz


def _ed_plate(nonlocal_L, thr):
    return _damage_plate(k0=1.0, k1=1.0, nonlocal_L=nonlocal_L, thr=thr, n=14)


def test_static_damage_increment_threshold_and_undamaged_type():
    # Rates of 0.25 (a power of two) make every phi-weighted average exactly
    # 0.25 and every sum below exact, so the threshold test is met with
    # equality: damage 0.5 + 0.25 reaches d_thres = 0.75 and becomes 1 (a ">"
    # test keeps 0.75); 0.25 + 0.25 stays 0.5.  Undamaged-type particles (type
    # 4) carry a different rate and keep their damage: averaging over them
    # would move the others' average off 0.25, updating them would change theirs.
    o, xyz = _ed_plate(nonlocal_L=0.3, thr=0.75)
    o.search_nonlocal(1.5)
    n = o.N
    types = np.zeros(n, dtype=np.int32)
    types[xyz[:, 1] > xyz[:, 1].max() - 0.25] = 4
    o.set_types(types)
    d0 = np.where(np.arange(n) % 3 == 0, 0.5, np.where(np.arange(n) % 3 == 1, 0.25, 0.0))
    o.set("DAMAGE", d0)
    o.store_state()
    o.set("DDOT_LOCAL", np.where(types == 4, 0.75, 0.25))
    assert o.update_static_damage(4)
    dn = o.get("DDOT_NONLOCAL")
    assert np.all(dn[types != 4] == 0.25)
    want = np.where(types == 4, d0, np.where(d0 + 0.25 >= 0.75, 1.0, d0 + 0.25))
    np.testing.assert_array_equal(o.get("DAMAGE"), want)
    assert (want == 1.0).sum() > 0 and ((want > 0) & (want < 1)).sum() > 0


def test_static_damage_return_value():
    # The return says whether any damage moved by more than 1e-3 from the
    # stored state, after the threshold: 2^-11 (< 1e-3) is not noticeable,
    # 2^-9 is, and 2^-11 that lifts one particle onto the threshold (so it
    # jumps to 1) is.  A comparison against the previous pass instead of the
    # stored state fails the second call, one made before the threshold fails
    # the third.
    small, big = 2.0 ** -11, 2.0 ** -9
    o, _ = _ed_plate(nonlocal_L=0.3, thr=0.75)
    o.search_nonlocal(1.5)
    o.set("DAMAGE", np.zeros(o.N))
    o.store_state()
    o.set("DDOT_LOCAL", np.full(o.N, small))
    assert not o.update_static_damage(-1)
    np.testing.assert_array_equal(o.get("DAMAGE"), small)
    assert not o.update_static_damage(-1)          # 2 x 2^-11 = 9.8e-4 from the stored 0
    assert o.update_static_damage(-1)              # 3 x 2^-11 = 1.5e-3
    o2, _ = _ed_plate(nonlocal_L=0.3, thr=0.75)
    o2.search_nonlocal(1.5)
    o2.set("DAMAGE", np.zeros(o2.N))
    o2.store_state()
    o2.set("DDOT_LOCAL", np.full(o2.N, big))
    assert o2.update_static_damage(-1)
    o3, _ = _ed_plate(nonlocal_L=0.3, thr=0.75)
    o3.search_nonlocal(1.5)
    d0 = np.zeros(o3.N)
    d0[5] = 0.75 - small
    o3.set("DAMAGE", d0)
    o3.store_state()
    o3.set("DDOT_LOCAL", np.full(o3.N, small))
    assert o3.update_static_damage(-1)
    dmg = o3.get("DAMAGE")
    assert dmg[5] == 1.0 and np.all(np.delete(dmg, 5) == small)


def test_j2_static_damage_threshold_is_zero():
    # ParticleJ2Plasticity::setParticleProperty never sets damage_threshold
    # (declared {0}, particle_j2plasticity.h:22,155-198), so the static update
    # damage += Ddot_nonlocal >= 0 sets every J2 particle that is not of the
    # undamaged type to 1 (particle_j2plasticity.h:142-152).  A restatement
    # that took the critical_bstrain argument (0.75 here) as the threshold
    # keeps d0 + 2^-11 instead.
    o, xyz = _j2_plate(nonlocal_L=0.3, crit=0.75)
    o.search_nonlocal(1.5)
    types = np.zeros(o.N, dtype=np.int32)
    types[xyz[:, 0] < xyz[:, 0].min() + 0.25] = 4
    o.set_types(types)
    d0 = np.random.default_rng(11).uniform(0.0, 0.5, o.N)
    o.set("DAMAGE", d0)
    o.store_state()
    o.set("DDOT_LOCAL", np.full(o.N, 2.0 ** -11))
    assert o.update_static_damage(4)
    np.testing.assert_array_equal(o.get("DAMAGE"), np.where(types == 4, d0, 1.0))
    assert (types == 4).sum() > 0


# --------------------------------------------------------- bond damage and force scaling

@pytest.mark.parametrize("ptype,model,params", [
    (lat.ELASTIC_DAMAGE, P.MAT_DAMAGE, [0.0, 1, E, NU, 1.0, 1.0, 1.0, 1.0]),
    (lat.J2PLASTICITY, P.MAT_J2, [0.0, 1, E, NU, 250.0, 0.0, 2000.0, 0.0, 1.0])], ids=["elastic_damage", "j2"])
def test_bond_damage_is_max_of_bond_and_both_particles(ptype, model, params):
    # bdamage = max(bdamage, D_i, D_j): hand-set bond damages below, between
    # and above their particles' damages.  Taking only D_i, a min, or keeping
    # the old bond damage fails.  The bond force is the undamaged force times
    # (1 - bdamage) exactly (the same products, then one multiplication).
    xyz = lat.hex_plate(8, 8, R)
    rng = np.random.default_rng(3)
    strained = xyz @ np.array([[1 + 2e-4, 3e-5, 0], [1e-5, 1 - 1e-4, 0], [0, 0, 1]]).T
    o = Oracle(lat.HEXAGON2D, R, ptype, xyz)
    o.set_material(model, params)
    ref = Oracle(lat.HEXAGON2D, R, ptype, xyz)
    ref.set_material(model, params)
    D = rng.uniform(0.0, 0.6, o.N)
    bd0 = rng.uniform(0.0, 0.6, o.nbond)
    o.set("DAMAGE", D)
    o.set("BDAMAGE", bd0)
    for s in (o, ref):
        s.set("XYZ", strained)
        s.forward()
    bptr, _, p2 = o.bonds()
    p1 = np.repeat(np.arange(o.N), np.diff(bptr))
    want = np.maximum(bd0, np.maximum(D[p1], D[p2]))
    got = o.get("BDAMAGE")
    np.testing.assert_array_equal(got, want)
    assert (want == bd0).sum() > 0 and (want == D[p1]).sum() > 0 and (want == D[p2]).sum() > 0
    np.testing.assert_array_equal(o.get("BFORCE"), ref.get("BFORCE") * (1.0 - want))
    assert np.abs(ref.get("BFORCE")).max() > 0


def test_elastic_particles_keep_their_bond_damage():
    # ParticleElastic::updateBondsForce does not take the particles' damage
    # into the bonds (particle_elastic.h:38-50): bdamage stays as set
    xyz = lat.hex_plate(6, 6, R)
    o = Oracle(lat.HEXAGON2D, R, lat.ELASTIC, xyz)
    o.set_material(P.MAT_ELASTIC, [0.0, 1, E, NU, 1.0])
    bd0 = np.random.default_rng(5).uniform(0.0, 0.5, o.nbond)
    o.set("DAMAGE", np.full(o.N, 0.7))
    o.set("BDAMAGE", bd0)
    o.forward()
    np.testing.assert_array_equal(o.get("BDAMAGE"), bd0)


# --------------------------------------------------------- nonlocal neighbour search

@pytest.mark.parametrize("L,gen", [(lat.HEXAGON2D, lambda: lat.hex_plate(16, 16, R)),
                                   (lat.BCC3D, lambda: lat.bcc_block(5, 5, 5, R))], ids=["hex", "bcc"])
def test_oracle_nonlocal_search_is_the_reference_loop(L, gen):
    # Assembly::searchNonlocalNeighbors (assembly.h:117-130): for every p1,
    # every p2 of pt_sys (p1 itself too) with distanceTo < ratio * p1's
    # nonlocal_L, pushed in pt_sys order.  distanceTo is sqrt of the pow(d, 2)
    # sum (particle.h:277-280); pow(d, 2) rounds as d * d.  Checked against a
    # numpy loop with per-particle lengths: a list sorted by distance, without
    # self, or using p2's length fails.
    xyz = gen() + np.random.default_rng(2).uniform(-0.005, 0.005, (len(gen()), 3)) * (
        [1, 1, 0] if L == lat.HEXAGON2D else 1)
    o = Oracle(L, R, lat.ELASTIC_DAMAGE, xyz)
    n = o.N
    lens = np.where(np.arange(n) % 2 == 0, 0.25, 0.4)
    for nl in (0.25, 0.4):
        idx = np.nonzero(lens == nl)[0].astype(np.int32)
        o.set_material(P.MAT_DAMAGE, [nl, 1 if L < 2 else 0, 200e3, 0.3, 1e-3, 1.0, 1.0, 1.0], idx=idx)
    o.search_nonlocal(1.5)
    ptr, idx = o.nonlocal_neighbors()
    d = xyz[:, None, :] - xyz[None, :, :]
    dist = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
    for i in range(n):
        want = np.nonzero(dist[i] < 1.5 * lens[i])[0]
        np.testing.assert_array_equal(idx[ptr[i]:ptr[i + 1]], want)
