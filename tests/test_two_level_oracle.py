"""Pins of the two-level preconditioner's oracle (oracle/two_level.py). The preconditioner has no reference
counterpart and hence no golden vector; each check below is fixed by mathematics, chosen so that a wrong
corner order, a transposed operand or a dropped term fails one of them."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle_lib as O
from oracle import two_level as TL


def grid_rows(n, pinned):
    rows = np.arange(n * n)
    if pinned:  # the reference's pinned top corners (drape.rs:148)
        rows = rows[(rows != (n - 1) * n) & (rows != n * n - 1)]
    return rows


@pytest.mark.parametrize("n,pinned", [(9, False), (17, True), (33, False), (65, True)])
def test_layout_reproduces_bilinear_fields_and_injects(n, pinned):
    rows = grid_rows(n, pinned)
    nc_count, parent, w = TL.coarse_layout(n, rows)
    nc = (n - 1) // 4 + 1
    assert nc_count == nc * nc - (2 if pinned else 0)
    # coarse coordinates of each coarse row
    coarse_xy = []
    for b in range(nc):
        for a in range(nc):
            v = 4 * b * n + 4 * a
            if v in set(rows.tolist()):
                coarse_xy.append((4 * a, 4 * b))
    coarse_xy = np.array(coarse_xy, np.float64)
    P = TL.prolongation(parent, w, nc_count)
    fine_xy = np.stack([rows % n, rows // n], 1).astype(np.float64)
    # rows whose coarse cell has four free corners
    pinned_corners = {(0, nc - 1), (nc - 1, nc - 1)} if pinned else set()
    a0 = np.minimum(fine_xy[:, 0] // 4, nc - 2).astype(int)
    b0 = np.minimum(fine_xy[:, 1] // 4, nc - 2).astype(int)
    complete = np.array([not ({(a, b), (a + 1, b), (a, b + 1), (a + 1, b + 1)} & pinned_corners)
                         for a, b in zip(a0, b0)])
    for f in (lambda x, y: 2.0 * x - 0.5 * y + 3.0, lambda x, y: x * y, lambda x, y: 1.0 + 0 * x):
        fc = f(coarse_xy[:, 0], coarse_xy[:, 1])
        got = P @ np.repeat(fc, 3)
        want = np.repeat(f(fine_xy[:, 0], fine_xy[:, 1]), 3)
        mask = np.repeat(complete, 3)
        assert np.allclose(got[mask], want[mask], rtol=0, atol=1e-12)
    # partition of unity where no parent is pinned; injection at coarse vertices
    assert np.allclose(w[complete].sum(1), 1.0)
    on_coarse = (fine_xy[:, 0] % 4 == 0) & (fine_xy[:, 1] % 4 == 0)
    assert np.all(np.sort(w[on_coarse], 1)[:, -1] == 1.0)
    assert np.all((w > 0) == (parent >= 0))


def scene_system(n, mu=0.0, bent=False, h=1.0 / 240.0):
    """The pinned sheet's Newton system at rest (SPD: K(rest) + M/h^2) or at the seed-97 bent state of
    e3_generate.rs:164-166 (indefinite where the exact tangent is), with a sagged x_hat."""
    od = O.Drape(O.reference_scene(n, sphere=False, placed=True))
    a = od.problem_arrays()
    rest = od.rest_positions()
    x = a["kernel_state"].astype(np.float64) if bent else rest
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    od.set_time_step(h)
    rhs, vals = od.assemble(x, x_hat, mu, True)
    A = TL.bsr_matrix(a["row_ptr"], a["col_idx"], vals)
    return od, a, A, rhs, vals


def test_galerkin_is_the_triple_product_and_spd():
    od, a, A, rhs, _ = scene_system(17, mu=1.0, bent=True)
    nc, parent, w = TL.coarse_layout(17, a["row_vertex"])
    P = TL.prolongation(parent, w, nc)
    Ac = TL.galerkin(A, P)
    Pd, Ad = P.toarray(), A.toarray()
    assert np.allclose(Ac, Pd.T @ Ad @ Pd, rtol=1e-12, atol=1e-12 * np.abs(Ac).max())
    assert np.allclose(Ac, Ac.T, rtol=1e-10, atol=1e-10 * np.abs(Ac).max())
    assert np.linalg.eigvalsh(0.5 * (Ac + Ac.T)).min() > 0
    # a transposed prolongation would not even conform; a dropped weight breaks the triple product
    w2 = w.copy()
    w2[5, :] = 0.0
    Ac2 = TL.galerkin(A, TL.prolongation(parent, w2, nc))
    assert not np.allclose(Ac2, Ac)


def test_two_level_pcg_solves_and_beats_block_jacobi():
    od, a, A, rhs, _ = scene_system(33, h=1.0)
    nc, parent, w = TL.coarse_layout(33, a["row_vertex"])
    P = TL.prolongation(parent, w, nc)
    direct = spla.spsolve(A.tocsc(), rhs)
    x2, it2, st2, _ = TL.two_level_pcg(A, rhs, P, 1e-10, 5000)
    x1, it1, st1, _ = TL.two_level_pcg(A, rhs, None, 1e-10, 20000)
    assert st2 == 1 and st1 == 1
    assert np.linalg.norm(x2 - direct) <= 1e-6 * np.linalg.norm(direct)
    assert it2 * 3 < it1, (it2, it1)


def test_without_coarse_space_it_is_the_block_jacobi_pcg_of_the_oracle():
    """P = None is block-Jacobi PCG: iteration count and solution agree with the C++ restatement of
    pcg.rs:60-281 in f64 (a different reduction order moves only the last bits)."""
    od, a, A, rhs, vals = scene_system(17, h=1.0)
    x1, it1, st1, _ = TL.two_level_pcg(A, rhs, None, 1e-8, 5000)
    xo, so, _ = O.pcg(a["row_ptr"].astype(np.int32), a["col_idx"].astype(np.int32), vals, rhs, 1e-8, 5000, 0, 0.0,
                      single=False)
    assert abs(it1 - so.iterations) <= 1
    assert np.linalg.norm(x1 - xo) <= 1e-7 * np.linalg.norm(xo)


def test_vcycle_error_propagation_and_symmetry():
    """The V-cycle application against the dense product of its three error-propagation factors, built
    independently from dense D^-1, P and A_c^-1; its operator is symmetric, positive definite for
    w lambda_max(D^-1 A) < 2 and not for w well above that."""
    _, a, A, _, _ = scene_system(9)
    nc, parent, w = TL.coarse_layout(9, a["row_vertex"])
    P = TL.prolongation(parent, w, nc)
    Ad = A.toarray()
    n = Ad.shape[0]
    Dd = np.zeros_like(Ad)
    for i in range(0, n, 3):
        Dd[i:i + 3, i:i + 3] = np.linalg.inv(Ad[i:i + 3, i:i + 3])
    Pd = P.toarray()
    Acinv = np.linalg.inv(Pd.T @ Ad @ Pd)
    lam = np.max(np.real(np.linalg.eigvals(Dd @ Ad)))
    for omega in (0.5 / lam, 1.0 / lam, 1.9 / lam):
        S = np.eye(n) - omega * Dd @ Ad
        E = S @ (np.eye(n) - Pd @ Acinv @ Pd.T @ Ad) @ S
        Minv = (np.eye(n) - E) @ np.linalg.inv(Ad)
        Minv_c = np.linalg.inv(np.linalg.cholesky(Pd.T @ Ad @ Pd))
        Minv_c = Minv_c.T @ Minv_c
        Dinv = TL.block_jacobi_inverse(A)
        cols = np.stack([TL.vcycle(A, Dinv, P, Minv_c, omega, e) for e in np.eye(n)], 1)
        assert np.abs(cols - Minv).max() <= 1e-9 * np.abs(Minv).max()
        assert np.abs(cols - cols.T).max() <= 1e-9 * np.abs(cols).max()
        assert np.linalg.eigvalsh(0.5 * (cols + cols.T)).min() > 0
    S = np.eye(n) - (3.0 / lam) * Dd @ Ad
    E = S @ (np.eye(n) - Pd @ Acinv @ Pd.T @ Ad) @ S
    Minv = (np.eye(n) - E) @ np.linalg.inv(Ad)
    assert np.linalg.eigvalsh(0.5 * (Minv + Minv.T)).min() < 0


def test_vcycle_pcg_solves_with_fewer_iterations():
    _, a, A, rhs, _ = scene_system(33, h=0.05)
    nc, parent, w = TL.coarse_layout(33, a["row_vertex"])
    P = TL.prolongation(parent, w, nc)
    xd = spla.spsolve(A.tocsc(), rhs)
    x1, it1, st1, _ = TL.two_level_pcg(A, rhs, P, 1e-10, 5000)
    x2, it2, st2, _ = TL.two_level_pcg(A, rhs, P, 1e-10, 5000, omega=0.5)
    assert st1 == 1 and st2 == 1
    assert np.abs(x2 - xd).max() <= 1e-7 * np.abs(xd).max()
    assert it2 < it1


@pytest.mark.parametrize("n,pinned", [(17, False), (33, True)])
def test_rotation_prolongation_reproduces_rigid_motions_of_the_shape(n, pinned):
    """With levers from a curved shape x, P reproduces t + theta x (x_i - o) exactly on every row whose parents
    are all free, from the coarse values (t + theta x (X_I - o), H theta): a lever of the wrong sign, a column
    e_a x v written as v x e_a, or a lever without its weight fails it."""
    rows = grid_rows(n, pinned)
    nc, parent, w = TL.coarse_layout(n, rows)
    cv = TL.coarse_vertices(n, rows)
    assert len(cv) == nc
    rng = np.random.default_rng(n)
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n)), -1).reshape(-1, 2) / (n - 1)
    # a draped-like shape: a bump, a twist and noise
    x = np.stack([g[:, 0] + 0.05 * np.sin(3 * g[:, 1]), g[:, 1], 0.3 * np.exp(-8 * ((g - 0.5) ** 2).sum(1))], 1)
    x += 1e-3 * rng.standard_normal(x.shape)
    H = 4.0 / (n - 1)
    lev = TL.lever_arms(parent, w, rows, cv, x.ravel(), H)
    P = TL.prolongation(parent, w, nc, lev)
    assert P.shape == (3 * len(rows), 6 * nc)
    t, theta, o = rng.standard_normal(3), rng.standard_normal(3), rng.standard_normal(3)
    X = x[cv]
    c = np.zeros((nc, 6))
    c[:, :3] = t + np.cross(theta, X - o)
    c[:, 3:] = H * theta
    got = (P @ c.ravel()).reshape(-1, 3)
    want = t + np.cross(theta, x[rows] - o)
    full = np.abs(w.sum(1) - 1.0) < 1e-12
    assert full.sum() > len(rows) // 2
    np.testing.assert_allclose(got[full], want[full], rtol=0, atol=1e-12 * np.abs(want).max())
    # the translation columns alone are the 3-unknown prolongation
    P3 = TL.prolongation(parent, w, nc)
    sel = np.concatenate([6 * np.arange(nc)[:, None] + np.arange(3)], 1).ravel()
    np.testing.assert_array_equal(P[:, sel].toarray(), P3.toarray())


def test_modified_cholesky_floor():
    """No pivot below the floor: numpy's Cholesky factor. A Gram matrix with dependent columns and an indefinite
    matrix: L L^T - A is diagonal, non-negative and nonzero exactly at the replaced pivots, and a rank
    deficiency of k replaces k pivots."""
    rng = np.random.default_rng(3)
    M = rng.standard_normal((40, 40))
    A = M @ M.T + 40 * np.eye(40)
    L, rep = TL.modified_cholesky(A, 1e-10)
    assert rep == []
    np.testing.assert_allclose(L, np.linalg.cholesky(A), rtol=0, atol=1e-12 * np.abs(L).max())
    V = rng.standard_normal((60, 30))
    V[:, 7] = V[:, 2] + 2 * V[:, 5]
    V[:, 20] = V[:, 11]
    G = V.T @ V
    L, rep = TL.modified_cholesky(G, 1e-10)
    E = L @ L.T - G
    off = E - np.diag(np.diag(E))
    assert np.abs(off).max() <= 1e-10 * np.abs(G).max()
    assert sorted(rep) == [7, 20]
    d = np.diag(E)
    assert np.all(d[rep] > 0.5 * np.diag(G)[rep])
    assert np.abs(np.delete(d, rep)).max() <= 1e-10 * np.abs(G).max()
    # indefinite with a positive diagonal: the second pivot 1 - 2^2 = -3 becomes 1
    S = np.array([[1.0, 2.0, 0.5], [2.0, 1.0, 0.0], [0.5, 0.0, 3.0]])
    L, rep = TL.modified_cholesky(S, 1e-10)
    assert rep == [1]
    E = L @ L.T - S
    np.testing.assert_allclose(E, np.diag([0.0, 4.0, 0.0]), atol=1e-14)
    # a non-positive diagonal entry is not replaced: the factorisation fails
    with pytest.raises(np.linalg.LinAlgError):
        TL.modified_cholesky(np.diag([1.0, -1.0]), 1e-10)
