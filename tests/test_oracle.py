"""Pins the oracle (the C++ f64 restatement of drape's Rust oracle) before anything is checked against it:
the reference's own unit tests, its committed golden vectors (E1 PCG histories, problem structure, PTC
traces, E4-final Newton histories) and closed forms."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))


def selftest_names():
    O.build()
    out = subprocess.run([os.path.join(O.ROOT, "oracle", "_build", "selftest"), "--list"], check=True,
                         capture_output=True, text=True).stdout
    return out.split()


@pytest.mark.parametrize("name", selftest_names())
def test_reference_unit_test(name):
    """The reference's #[test]s (membrane, hinge, dual, contact, drape, newton, pcg, band, cases, e4_rules,
    e7_ptc) restated in oracle/selftest.cpp with the same inputs and tolerances."""
    r = subprocess.run([os.path.join(O.ROOT, "oracle", "_build", "selftest"), name], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("PASS"), r.stdout


E1_FAST = ["tension-n33", "dynamic-n33", "slack-n33", "tension-n65", "dynamic-n65", "slack-n65", "tension-n129",
           "dynamic-n129"]


@pytest.mark.parametrize("case", E1_FAST)
def test_e1_pcg_matches_reference_histories(case):
    """experiments/cases/<case>/reference.json: the f32 emulation reproduces every recorded relative residual
    bit for bit; f64 PCG reproduces the iteration count and the history to 1e-9 (the reference's libm is
    the Windows CRT, so f64 assembly differs from glibc's in the last bits)."""
    g = GOLDEN["e1"][case]
    kind, n = case.split("-n")
    c = O.Case(kind, int(n))
    assert (c.rows, c.blocks) == (g["rows"], g["blocks"])
    x64, s64, h64 = O.pcg(c.row_ptr, c.col_idx, c.vals, c.rhs, 1e-7, 20000, 0, 0.0, single=False)
    # slack: a nearly singular system, where f64 CG may take one more or one fewer iteration
    assert abs(s64.iterations - g["f64_iterations"]) <= (2 if kind == "slack" else 0)
    ref64 = np.array([v for _, v in g["f64_history"]])
    got64 = np.array([v for _, v in h64])
    n_hist = min(len(h64), len(g["f64_history"])) - 1
    assert [i for i, _ in h64][:n_hist] == [i for i, _ in g["f64_history"]][:n_hist]
    # CG amplifies last-bit input differences as it loses orthogonality: the first history points agree to
    # 1e-9 (3 for the nearly singular slack sheet), the iteration count exactly.
    keep = 3 if kind == "slack" else 8
    np.testing.assert_allclose(got64[:keep], ref64[:keep], rtol=1e-9)
    x32, s32, h32 = O.pcg(c.row_ptr, c.col_idx, c.vals, c.rhs, 1e-7, 20000, g["replace_every"], g["replace_above"],
                          single=True)
    assert s32.iterations == g["f32_iterations"]
    assert bool(s32.converged) == g["f32_converged"]
    # io.rs renders non-finite numbers as null: those entries must be non-finite here too
    assert [i for i, _ in h32] == [i for i, _ in g["f32_history"]]
    for (_, got), (_, want) in zip(h32, g["f32_history"]):
        assert (not np.isfinite(got)) if want is None else got == want


@pytest.mark.parametrize("case", ["tension-n33", "dynamic-n33", "dynamic-n65"])
def test_e1_band_cholesky_bandwidth_and_solution(case):
    g = GOLDEN["e1"][case]
    kind, n = case.split("-n")
    c = O.Case(kind, int(n))
    x, bw, _ = O.band_solve(c.row_ptr, c.col_idx, c.vals, c.rhs)
    assert bw == g["bandwidth"]
    x64, _, _ = O.pcg(c.row_ptr, c.col_idx, c.vals, c.rhs, 1e-12, 20000)
    assert np.linalg.norm(x64 - x) <= 1e-9 * np.linalg.norm(x)


@pytest.mark.parametrize("rel", sorted(GOLDEN["structure"]))
def test_problem_structure_matches_reference_manifests(rel):
    """Every problem.bin array length of the reference's drape problems (pattern, scatter maps, colliders)."""
    g = GOLDEN["structure"][rel]
    n = g["grid_columns"]
    sphere = g["scene"] == "sphere_on_floor"
    d = O.Drape(O.reference_scene(n, sphere=sphere, placed=g["placed"]))
    lengths = g["lengths"]
    assert 3 * d.V == lengths["rest_positions"]
    assert 3 * d.T == lengths["elem_vertices"]
    assert 4 * d.H == lengths["hinge_vertices"]
    assert d.R == lengths["row_vertex"]
    assert d.B == lengths["col_idx"]
    assert d.block_contribs == lengths["block_contrib_idx"]
    assert d.force_contribs == lengths["force_contrib_idx"]
    a = d.problem_arrays()
    for name in ("row_ptr", "col_idx", "diag_idx", "block_contrib_ptr", "block_contrib_idx", "force_contrib_ptr",
                 "force_contrib_idx", "elem_geom", "hinge_stiffness", "colliders", "contact_state", "kernel_state"):
        assert a[name].size == lengths.get(name, 0), name
    if g["kappa"] is not None and sphere:
        assert d.kappa == pytest.approx(g["kappa"], rel=1e-14)
    if g["gravity_load"] is not None:
        assert d.gravity_load == pytest.approx(g["gravity_load"], rel=1e-15)


@pytest.mark.parametrize("n,exact_steps", [(17, 25), (33, 18)])
def test_ptc_trace_matches_reference_log(n, exact_steps):
    """experiments/drape7/drape-n<n>/base/cold.log: the f64 PTC controller reproduces the reference's progress
    lines (h, r/L to 4 digits, inner Newton count, rejections, max mu) until the chaotic trajectory amplifies
    last-bit libm differences, and converges to the same equilibrium (static energy to 1e-12)."""
    g = GOLDEN["ptc"][f"n{n}"]
    d = O.Drape(O.reference_scene(n, sphere=True, placed=True))
    log, final, st = d.ptc_solve()
    mine = log.splitlines()
    assert mine[:exact_steps] == g["log"][:exact_steps]
    assert st.reason == 0 and g["summary"]["reason"] == "tolerance"
    assert st.final_relative_residual <= 1e-8
    assert st.static_energy == pytest.approx(g["summary"]["static_energy"], rel=1e-12)
    assert st.max_mu == g["summary"]["max_mu"]


def test_e4_final_newton_histories_match_reference():
    """experiments/drape4-final/drape-n33/reference.json: the f64 regularised Newton of the pinned drape
    (exact tangent, floor factor 2, tolerance 1e-10, cap 200) reproduces the first steps' residual histories."""
    steps = GOLDEN["e4_final_n33"]["steps"]
    d = O.Drape(O.reference_scene(33, sphere=False))
    k = 6
    iters, hist, _ = d.simulate(k, 1e-10, 200, True, 2.0)
    for n in range(k):
        ref = np.array(steps[n]["relative_residuals"])
        assert iters[n] == steps[n]["newton_iterations"], n
        assert len(hist[n]) == len(ref)
        np.testing.assert_allclose(hist[n][:-1], ref[:-1], rtol=1e-6, err_msg=f"step {n + 1}")
        # the last entry sits at the f64 floor (~1e-10): rounding-level agreement only
        assert hist[n][-1] <= 1e-10 or abs(hist[n][-1] - ref[-1]) <= 1e-3 * ref[-1] + 1e-12


def test_derivative_check_matches_reference_gates():
    """e3_checks.rs:144-158 at n33 (seed 53): the same FD errors as the reference's derivative_check, within
    the noise of the finite differences, and under the reference's gates."""
    ref = GOLDEN["e4_final_n33"]["derivative_check"]
    c = O.derivative_check(33)
    assert c.gradient_error < 1e-6 and c.hessian_error < 1e-5 and c.energy_change_error < 1e-12
    assert c.hinge_energy_change_error < 1e-12
    assert c.gradient_error == pytest.approx(ref["gradient_relative_error"], rel=0.5)
    assert c.hessian_error == pytest.approx(ref["hessian_relative_error"], rel=0.5)


def test_isotropic_material_maps_onto_reference_parameters():
    """E = 1e6 Pa, nu = 0.3, t = 0.5 mm: plane-stress constants in the reference's (wale, course, coupling,
    shear) form, and the energy density equals the isotropic St. Venant-Kirchhoff one."""
    E, nu, t = 1e6, 0.3, 5e-4
    wale = E * t / (1 - nu * nu)
    coupling = nu * wale
    shear = E * t / (2 * (1 + nu))
    assert wale == pytest.approx(549.4505494505494)
    assert coupling == pytest.approx(164.83516483516485)
    assert shear == pytest.approx(192.30769230769232)
    # W = t * (lambda'/2 tr(E)^2 + mu tr(E^2)) with plane-stress lambda' = E nu/(1-nu^2), mu = E/(2(1+nu))
    rng = np.random.default_rng(3)
    for _ in range(5):
        eww, ecc, ewc = rng.normal(size=3) * 1e-2
        w_ref = 0.5 * wale * eww ** 2 + 0.5 * wale * ecc ** 2 + coupling * eww * ecc + 2 * shear * ewc ** 2
        lam = E * nu / (1 - nu * nu)
        mu = E / (2 * (1 + nu))
        tr = eww + ecc
        w_iso = t * (0.5 * lam * tr * tr + mu * (eww ** 2 + ecc ** 2 + 2 * ewc ** 2))
        assert w_ref == pytest.approx(w_iso, rel=1e-12)
