"""A fragment that broken bonds have cut loose from every constrained particle.

Its rigid motions make the stiffness singular; the GPU path finds such
fragments (connected components of the intact-bond graph without a
displacement constraint) and holds their particles in place.  Where the
fragment is in equilibrium -- as here, cut out of an unloaded plate -- the
reference's algorithm gives the same answer (its CG never leaves the
fragment-free subspace), so the result must match the oracle to the usual bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import lpmg
from lpmg import problems as P
from oracle import Oracle


def _cut_out(s, xyz, centre, radius):
    """Break every bond that crosses the circle; return the particles inside."""
    bptr, nb0, p2 = s.bonds()
    inside = np.hypot(xyz[:, 0] - centre[0], xyz[:, 1] - centre[1]) < radius
    p1 = np.repeat(np.arange(len(xyz)), np.diff(bptr))
    bd = s.get("BDAMAGE")
    bd[inside[p1] != inside[p2]] = 1.0
    s.set("BDAMAGE", bd)
    return inside


@pytest.mark.parametrize("precond", [1, 2])
def test_loose_fragment_is_held_and_rest_matches_oracle(precond):
    prob = P.hex_plate(nx=30, ny=26, steps=3)
    prob.tol = 1e-11
    xyz = prob.xyz
    centre = (xyz[:, 0].mean(), xyz[:, 1].mean())
    g, gg = P.build(lpmg.System, prob)
    o, go = P.build(Oracle, prob)
    ins_g = _cut_out(g, xyz, centre, 0.45)
    ins_o = _cut_out(o, xyz, centre, 0.45)
    assert np.array_equal(ins_g, ins_o) and 2 <= ins_g.sum() < 40
    g.forward()
    o.forward()
    P.solve_problem(g, gg, prob, dict(precond=precond, rtol=1e-24, atol=0.0))
    P.solve_problem(o, go, prob, dict(precond=0, rtol=1e-24, atol=0.0))
    np.testing.assert_array_equal(g.get("FLOATING") > 0.5, ins_g)
    x0 = o.get("XYZ_INITIAL")
    ug, uo = g.get("XYZ") - x0, o.get("XYZ") - x0
    assert np.linalg.norm(ug - uo) <= 1e-8 * np.linalg.norm(uo)
    frag = np.repeat(ins_g, 3)
    assert np.all(ug[frag] == 0.0)


def test_fragment_labels_follow_constraint_changes():
    """The labels are recomputed only when the constrained set can change: a
    displacement condition repeating one already applied keeps them, while an
    overwritten constraint field or a restart (which clears the constraints)
    makes the next assembly relabel, also when the same conditions follow."""
    prob = P.hex_plate(nx=30, ny=26, steps=3)
    xyz = prob.xyz
    centre = (xyz[:, 0].mean(), xyz[:, 1].mean())
    g, gg = P.build(lpmg.System, prob)
    ins = _cut_out(g, xyz, centre, 0.45)
    assert 2 <= ins.sum() < 40
    held = lambda: g.get("FLOATING") > 0.5

    def step_and_assemble():
        P.apply_step(g, gg, prob.steps[0])
        g.forward()
        g.assemble()

    step_and_assemble()
    np.testing.assert_array_equal(held(), ins)
    step_and_assemble()  # the same conditions again: labels kept
    np.testing.assert_array_equal(held(), ins)

    # one constrained DOF inside the fragment ties it down
    p = int(np.flatnonzero(ins)[0])
    cons = g.get("CONSTRAINT")
    cons[3 * p] = 1.0
    g.set("CONSTRAINT", cons)
    step_and_assemble()
    assert not held().any()
    cons[3 * p] = 0.0
    g.set("CONSTRAINT", cons)
    step_and_assemble()
    np.testing.assert_array_equal(held(), ins)

    # a cleared constraint field, assembled, then the same conditions again
    g.set("CONSTRAINT", np.zeros_like(cons))
    g.forward()
    g.assemble()
    assert not held().any()
    step_and_assemble()
    np.testing.assert_array_equal(held(), ins)

    # restart clears constraints and broken bonds; with nothing constrained
    # nothing is held, and the first condition afterwards relabels
    g.restart()
    assert np.array_equal(_cut_out(g, xyz, centre, 0.45), ins)
    g.forward()
    g.assemble()
    assert not held().any()
    step_and_assemble()
    np.testing.assert_array_equal(held(), ins)
