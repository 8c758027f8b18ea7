"""GPU tests of the sheet's self-contact (drape_amd/csrc/kernels/selfcontact.hip, rules.self_contact) through the
C ABI, against the f64 oracle (oracle/selfcontact.py, pinned by dense sampling and finite differences) on a
folded n17 sheet whose upper half lies over its lower half:
  - the point-triangle and edge-edge pairs closer than dhat equal the oracle's brute-force sets, with the
    same distances (1e-12 relative);
  - the assembled right-hand side changes by the oracle's barrier gradient (FP32: 2e-5 of its largest entry);
  - PCG solves the full operator, the contact coupling outside the block pattern included: the true residual
    of the oracle's matrix (assembled blocks + Gauss-Newton couplings) is at the tolerance;
  - with no pair within dhat the path is bit-identical to the one without self-contact;
  - the step bound keeps every pair that can come within dhat above 0.1 of its distance, from the swept-box
    hash and, for a step too long for it, from the inflated grid; the line search's energy change includes
    the pairs' barrier change (1e-6 relative);
  - dynamics: under gravity the upper half comes to rest on the lower half with self-contact and never
    crosses it, while without self-contact it falls through to the floor.
With rules.self_contact = 0 (the default) every other test runs the unchanged path."""
import numpy as np
import pytest

import drape_amd as D
from oracle import selfcontact as S

pytestmark = pytest.mark.gpu

N = 17
DHAT = 1e-3


def folded(gap):
    """Positions [V, 3] of the folded sheet: lower half resting at dhat/2 above the floor (z = 0), rows past
    the fold mirrored onto it `gap` higher; and the scene (floor collider, negligible bending)."""
    scene = D.make_scene(N, pinned=False, colliders=[(1, 0.0, 0.0, 1.0, 0.0)], dhat=DHAT, kappa=0.0, lift=0.0,
                         bend=(1e-9, 1e-9), time_step=1.0 / 240.0)
    arrays = D.SceneArrays(scene)
    x = arrays.array("rest_positions").astype(np.float64).reshape(-1, 3).copy()
    L = x[:, 1].max()
    top = x[:, 1] > 0.5 * L + 1e-9
    x[:, 2] = 0.5 * DHAT
    x[top, 1] = L - x[top, 1]
    x[top, 2] += gap
    return arrays, x.astype(np.float32).astype(np.float64), top


def context(arrays, on):
    return D.Drape(arrays, D.default_rules(self_contact=1 if on else 0))


def assembled(arrays, x, on, h=1.0 / 240.0):
    ctx = context(arrays, on)
    xf = x.astype(np.float32).ravel()
    ctx.set_state(xf, xf, h, 0.0)
    ctx.residual()
    return ctx


def kappa_of(arrays):
    return float(arrays.array("colliders")[2])


def tri_of(arrays):
    return arrays.array("elem_vertices").reshape(-1, 3)


def test_no_pairs_leaves_the_step_unchanged():
    """A flat sheet (no two primitives within dhat, empty pair sets in both broad phases): with self-contact on,
    the assembled system, the PCG solution, the step bound and the energy change are bit-identical to the
    path without it."""
    arrays, _, _ = folded(0.5 * DHAT)
    flat = arrays.array("rest_positions").astype(np.float64).reshape(-1, 3).copy()
    flat[:, 2] = 0.5 * DHAT
    on, off = assembled(arrays, flat, True), assembled(arrays, flat, False)
    _, verts, _ = on.self_contact_pairs()
    assert len(verts) == 0
    rows = arrays.array("row_vertex")
    d = np.zeros((arrays.V, 3))
    d[:, 2] = -0.2 * DHAT
    d[:, 0] = 1e-3 * np.sin(np.arange(arrays.V))
    out = {}
    for name, c in (("on", on), ("off", off)):
        rhs, vals = c.system()
        st, sol = c.solve_once(1e-6, 2000)
        c.set_direction(d[rows].astype(np.float32).ravel())
        alpha, row, _ = c.max_step()
        e = c.energy_change(0.5 * alpha)
        out[name] = (rhs, vals, sol, st.iterations, alpha, row, e)
    a, b = out["on"], out["off"]
    print(f"[gpu] self-contact without pairs: PCG {a[3]} / {b[3]} iterations, alpha {a[4]:.4f} / {b[4]:.4f}",
          flush=True)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    assert a[3:] == b[3:]
    on.close()
    off.close()


def test_pairs_match_brute_force():
    arrays, x, _ = folded(0.5 * DHAT)
    ctx = assembled(arrays, x, True)
    npt, verts, dist = ctx.self_contact_pairs()
    quads, d_ref, kinds = S.candidates(x, tri_of(arrays), DHAT)
    print(f"[gpu] self-contact pairs: {npt} point-triangle, {len(verts) - npt} edge-edge "
          f"(oracle {int((kinds == 0).sum())}, {int((kinds == 1).sum())})", flush=True)
    got = {(0 if i < npt else 1, *map(int, v)): d for i, (v, d) in enumerate(zip(verts, dist))}
    want = {(int(k), *map(int, q)): d for q, d, k in zip(quads, d_ref, kinds)}
    # a pair within 1e-12 of dhat may fall either side
    edge = {k for k, d in want.items() if abs(d - DHAT) < 1e-12 * DHAT}
    assert set(got) - edge == set(want) - edge
    err = max(abs(got[k] - want[k]) / want[k] for k in got)
    assert err <= 1e-12
    assert len(got) > 100
    ctx.close()


def test_rhs_change_is_the_barrier_gradient():
    arrays, x, _ = folded(0.5 * DHAT)
    on, off = assembled(arrays, x, True), assembled(arrays, x, False)
    rhs_on, _ = on.system()
    rhs_off, _ = off.system()
    quads, _, kinds = S.candidates(x, tri_of(arrays), DHAT)
    g = S.gradient(x, quads, kinds, DHAT, kappa_of(arrays))
    rows = arrays.array("row_vertex")
    want = -g[rows].ravel()
    got = rhs_on.astype(np.float64) - rhs_off.astype(np.float64)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[gpu] self-contact gradient: max |g| {np.abs(want).max():.3e}, max error / max {err:.2e}", flush=True)
    assert err <= 2e-5
    on.close()
    off.close()


def test_pcg_solves_the_full_operator():
    arrays, x, _ = folded(0.5 * DHAT)
    ctx = assembled(arrays, x, True)
    rhs, vals = ctx.system()
    st, sol = ctx.solve_once(1e-6, 20000)
    assert D.PCG_STATUS[st.status] == "converged"
    R = arrays.R
    row_ptr, col = arrays.array("row_ptr"), arrays.array("col_idx")
    A = np.zeros((3 * R, 3 * R))
    blocks = vals.astype(np.float64).reshape(-1, 3, 3)
    for i in range(R):
        for b in range(row_ptr[i], row_ptr[i + 1]):
            j = col[b]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blocks[b]
    quads, _, kinds = S.candidates(x, tri_of(arrays), DHAT)
    H = S.gauss_newton(x, quads, kinds, DHAT, kappa_of(arrays), arrays.V)
    rows = arrays.array("row_vertex")
    idx = (3 * rows[:, None] + np.arange(3)[None, :]).ravel()
    Hr = H[np.ix_(idx, idx)]
    # the assembled diagonal blocks already hold the pairs' diagonal blocks
    for i in range(R):
        Hr[3 * i:3 * i + 3, 3 * i:3 * i + 3] = 0.0
    full = A + Hr
    b = rhs.astype(np.float64)
    true_rel = np.linalg.norm(b - full @ sol.astype(np.float64)) / np.linalg.norm(b)
    without = np.linalg.norm(b - A @ sol.astype(np.float64)) / np.linalg.norm(b)
    print(f"[gpu] self-contact PCG: {st.iterations} iterations; true relative residual {true_rel:.2e} "
          f"(against the block pattern alone {without:.2e})", flush=True)
    assert true_rel <= 3e-6
    assert without > 100 * true_rel
    ctx.close()


def test_step_bound_and_energy_change():
    arrays, x, top = folded(2.0 * DHAT)
    ctx = assembled(arrays, x, True)
    off = assembled(arrays, x, False)
    rows = arrays.array("row_vertex")
    # the upper half moves down by 4 dhat: through the lower half without a bound
    d = np.zeros((arrays.V, 3))
    d[top, 2] = -4.0 * DHAT
    d[:, 0] += 1e-4 * np.sin(np.arange(arrays.V))
    dr = d[rows].astype(np.float32).ravel()
    for c in (ctx, off):
        c.set_direction(dr)
    alpha, row, _ = ctx.max_step()
    alpha_off, _, _ = off.max_step()
    xs = x + alpha * d.astype(np.float32).astype(np.float64)
    quads, d0, kinds = S.candidates(x, tri_of(arrays), DHAT + 2.0 * np.abs(d).max() * 1.01)
    d1 = np.array([S.pair_geometry(xs, q, k)[0] for q, k in zip(quads, kinds)])
    ratio = (d1 / d0).min()
    print(f"[gpu] self-contact step bound: alpha {alpha:.4f} (colliders alone {alpha_off:.4f}), row {row}, "
          f"{len(quads)} pairs, min d(x + alpha p)/d(x) {ratio:.3f}", flush=True)
    # the floor alone allows 0.5625 (2.5 dhat above it, 0.1 kept); the lower half stops the upper half first
    assert alpha < alpha_off < 1.0
    assert ratio >= 0.1 * (1 - 1e-6)
    assert alpha > 0.2  # the pairs start 2 dhat apart and close at 4 dhat per unit step
    # the line search's energy change adds the pairs' barrier change
    a = 0.9 * alpha  # the upper half ends 0.38 dhat above the lower
    s_on, _ = ctx.energy_change(a)
    s_off, _ = off.energy_change(a)
    # the trial positions as the device forms them: FP32 pos + alpha dir
    xa = (x.astype(np.float32) + np.float32(a) * d.astype(np.float32)).astype(np.float64)
    kap = kappa_of(arrays)
    want = S.energy(xa, quads, kinds, DHAT, kap) - S.energy(x, quads, kinds, DHAT, kap)
    assert want > 0
    assert abs((s_on - s_off) - want) <= 1e-6 * abs(want)
    ctx.close()
    off.close()


def test_step_bound_of_a_far_step():
    """A step that moves one vertex of the upper half 52 m (its box over the step spans more cells than the
    swept-box hash takes) is bounded on the inflated grid instead: every pair within the reach of the step
    keeps at least 0.1 of its distance."""
    arrays, x, top = folded(2.0 * DHAT)
    ctx = assembled(arrays, x, True)
    rows = arrays.array("row_vertex")
    d = np.zeros((arrays.V, 3))
    d[top, 2] = -4.0 * DHAT
    w = int(np.flatnonzero(top)[len(np.flatnonzero(top)) // 2])
    d[w] += 30.0
    ctx.set_direction(d[rows].astype(np.float32).ravel())
    alpha, row, _ = ctx.max_step()
    xs = x + alpha * d.astype(np.float32).astype(np.float64)
    tri = tri_of(arrays)
    # the pairs the ordinary part of the step reaches, and every pair of the far vertex
    quads, _, kinds = S.candidates(x, tri, DHAT + 8.0 * DHAT * 1.01 + 1e-3)
    extra = [((w, *map(int, t)), 0) for t in tri if w not in t]
    extra += [((int(u), *map(int, t)), 0) for t in tri if w in t for u in range(arrays.V) if u not in t]
    edges = S.unique_edges(tri)
    for p, q in edges:
        for r, t in edges:
            if w in (p, q) and len({int(p), int(q), int(r), int(t)}) == 4:
                extra.append(((int(p), int(q), int(r), int(t)), 1))
    pairs = [(tuple(map(int, q)), int(k)) for q, k in zip(quads, kinds)] + extra
    d0 = np.array([S.pair_geometry(x, np.array(q), k)[0] for q, k in pairs])
    d1 = np.array([S.pair_geometry(xs, np.array(q), k)[0] for q, k in pairs])
    ratio = (d1 / d0).min()
    print(f"[gpu] self-contact far step: alpha {alpha:.3e}, row {row}, {len(pairs)} pairs, min d(x + alpha p)/d(x) "
          f"{ratio:.3f}", flush=True)
    assert 0.0 < alpha < 0.01  # the far vertex bounds the step (0.45 without it)
    assert ratio >= 0.1 * (1 - 1e-6)
    ctx.close()


def test_fold_comes_to_rest_without_crossing():
    arrays, x, top = folded(3.0 * DHAT)
    tri = tri_of(arrays)
    gaps = {}
    for on in (True, False):
        ctx = context(arrays, on)
        out = ctx.simulate(24, x0=x.astype(np.float32).ravel())
        xe = out["positions"].astype(np.float64).reshape(-1, 3)
        # the vertex of the lower half under each vertex of the upper half (mirror rows)
        cols = N
        idx = np.arange(arrays.V)
        r, c = idx // cols, idx % cols
        under = (N - 1 - r) * cols + c
        gaps[on] = (xe[top, 2] - xe[under[top], 2]).min()
        crossing = S.crossings(xe, tri)
        print(f"[gpu] fold, self-contact {'on' if on else 'off'}: min height of the upper half over the lower "
              f"{gaps[on] / DHAT:.3f} dhat, {crossing} edge-triangle crossings, updates "
              f"{int(out['updates'].sum())}", flush=True)
        if on:
            assert crossing == 0
        ctx.close()
    assert gaps[True] > 0.05 * DHAT
    assert gaps[False] < 0.01 * DHAT


def test_fold_equilibrium_by_ptc():
    """The folded sheet brought to static equilibrium by PTC (solveEquilibrium with self-contact): it stops by
    tolerance or stagnation with the upper half resting on the lower one and no edge crossing a triangle."""
    arrays, x, top = folded(3.0 * DHAT)
    ctx = context(arrays, True)
    out = ctx.solve_equilibrium(start=x.astype(np.float32).ravel(), cap=200)
    xe = out["positions"].astype(np.float64).reshape(-1, 3)
    idx = np.arange(arrays.V)
    under = (N - 1 - idx // N) * N + idx % N
    gap = (xe[top, 2] - xe[under[top], 2]).min()
    crossing = S.crossings(xe, tri_of(arrays))
    res = out["result"]
    print(f"[gpu] fold PTC: stop {out['stop']}, {res.accepted_steps} steps, {res.newton_iterations} Newton; min "
          f"height of the upper half {gap / DHAT:.3f} dhat, {crossing} crossings", flush=True)
    assert out["stop"] in ("tolerance", "stagnation")
    assert crossing == 0 and gap > 0.05 * DHAT
    ctx.close()
