"""CPU pins of the prolongation restatement (oracle/prolong.py, e7_prolong.rs:9-47) and of the oracle's PTC from
a prolonged start: the reference's own unit test of `prolong`, closed forms of the projection out of the
colliders, and the reference's n33 prolong trace (experiments/drape7/drape-n33/base/prolong.log: the n17
equilibrium prolonged to n33 and solved by f64 PTC from h = 1)."""
import numpy as np
import pytest

import drape_amd as D
import oracle_lib as O
from contract_case import CONTRACT
from oracle import prolong as P


def test_prolong_known_answers_of_the_reference_unit_test():
    """e7_prolong.rs:72-83: a 3x3 coarse grid with x = i, y = j, z = i j."""
    coarse = np.array([[v % 3, v // 3, (v % 3) * (v // 3)] for v in range(9)], np.float64)
    fine = P.prolong(coarse, 3)
    assert fine.shape == (25, 3)
    assert list(fine[0]) == list(coarse[0])
    assert list(fine[2]) == list(coarse[1])
    assert list(fine[1]) == [0.5, 0.0, 0.0]
    assert list(fine[5 + 1]) == [0.5, 0.5, 0.25]
    assert list(fine[3 * 5 + 3]) == [1.5, 1.5, 2.25]


def test_prolong_of_a_bilinear_field_is_exact():
    """Averages of corners reproduce any field linear in (i, j) exactly at every fine vertex."""
    n = 9
    i, j = np.meshgrid(np.arange(n), np.arange(n))
    coarse = np.stack([0.25 * i.ravel(), 0.5 * j.ravel(), 0.125 * (i + j).ravel()], 1)
    fine = P.prolong(coarse, n)
    f = 2 * (n - 1) + 1
    fi, fj = np.meshgrid(np.arange(f) / 2.0, np.arange(f) / 2.0)
    want = np.stack([0.25 * fi.ravel(), 0.5 * fj.ravel(), 0.125 * (fi + fj).ravel()], 1)
    assert np.array_equal(fine, want)


def test_projection_closed_forms():
    """A vertex below the floor or inside the sphere ends at d = dhat along the collider normal; one outside
    the barrier does not move."""
    dhat = 1e-3
    packed = np.array([2, dhat, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 2, 0.5, 0.5, 0.2, 0.1, 0, 0, 0], np.float64)
    x = np.array([[0.0, 0.0, -0.01], [0.5, 0.5, 0.25], [0.1, 0.1, 0.5]])
    y, moved = P.project_out(x, packed)
    assert moved == 2
    assert y[0] == pytest.approx([0.0, 0.0, dhat], abs=1e-18)
    assert np.linalg.norm(y[1] - [0.5, 0.5, 0.2]) == pytest.approx(0.1 + dhat, rel=1e-15)
    assert list(y[2]) == list(x[2])
    assert P.infeasible(y, packed) == []
    assert P.infeasible(np.array([[0.5, 0.5, 0.2]]), packed) == [0]


def test_oracle_prolong_trace_matches_reference_log():
    """The oracle's n17 equilibrium (static energy pinned to 1e-12 against the reference's) prolonged to n33 and
    projected: its f64 r/L and the oracle's PTC from it at h = 1 reproduce the reference's prolong.log lines
    (h, r/L to 4 digits, inner count, rejections, mu) over the first steps, and the solve converges to the
    reference's static energy."""
    g = CONTRACT["prolong_n33"]
    coarse = O.Drape(O.reference_scene(17, sphere=True, placed=True))
    _, eq, st = coarse.ptc_solve()
    assert st.final_h == g["start"]["final_h"]
    fine_scene = D.reference_scene(33, sphere_on_floor=True, placed=True)
    x = P.prolong(eq, 17)
    x, moved = P.project_out(x, D.scene_colliders(fine_scene))
    assert P.infeasible(x, D.scene_colliders(fine_scene)) == []
    fine = O.Drape(O.reference_scene(33, sphere=True, placed=True))
    log, final, st2 = fine.ptc_solve(x.ravel(), st.final_h)
    mine = log.splitlines()
    print(f"prolonged start: {moved} vertices projected; first line {mine[0]}")
    assert mine[:6] == g["log"][:6]
    assert st2.reason == 0
    assert st2.static_energy == pytest.approx(g["summary"]["static_energy"], rel=1e-9)
