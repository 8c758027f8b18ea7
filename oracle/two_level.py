"""Oracle for the two-level preconditioner: a numpy f64 statement of the coarse layout, the Galerkin
coarse operator and two-level PCG. Test infrastructure only: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may load it, and only as the checker.

The reference has no two-level preconditioner (its PCG is block-Jacobi only, engine/oracle/src/pcg.rs
and engine/gpu/src/solver.ts:109-191); the nearest code is the bilinear coarse-to-fine prolongation of
positions in engine/oracle/src/e7_prolong.rs:9-22. There is therefore no golden vector: every function
here is pinned by mathematics instead (tests/test_two_level_oracle.py):
  - the layout reproduces bilinear fields exactly and injects at coarse vertices (closed form);
  - weights of a row whose parents are all free sum to one (partition of unity);
  - A_c = P^T A P equals the dense triple product and is SPD for SPD A with full-rank P;
  - two-level PCG converges to the direct solution and in exact arithmetic never takes more iterations
    than the dimension; with a zero coarse space it is block-Jacobi PCG step for step;
  - with rotation unknowns (levers v_ik = w_ik (x_i - X_I) / H), P reproduces every rigid motion of the
    shape x exactly: translations t and rotations theta x (x - o) from the coarse values (t, theta x (X_I - o))
    and (0, H theta) (closed form, partition of unity);
  - the modified Cholesky factor is numpy's Cholesky factor when no pivot falls below the floor, and otherwise
    L L^T - A is diagonal, non-negative and nonzero only at the replaced pivots.

Layout (shared with the HIP library): coarse vertices are every f-th grid vertex in each direction of a
sheet with `columns` vertices per side ((columns - 1) % f == 0; f = 4 unless given); coarse rows are the free
coarse vertices in vertex order; each fine row has four parent slots k = 0..3 = the corners (a0, b0),
(a0 + 1, b0), (a0, b0 + 1), (a0 + 1, b0 + 1) of its coarse cell with a0 = min(x // f, nc - 2),
fx = (x - f a0) / f, weights (1 - fx)(1 - fy), fx (1 - fy), (1 - fx) fy, fx fy; a slot whose corner is
pinned, or whose weight is 0, is empty (parent -1, weight 0).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def coarse_layout(columns: int, row_vertex: np.ndarray, factor: int = 4):
    """(Nc, parent [R, 4] int32, weights [R, 4] float64) of the stride-`factor` coarse grid."""
    n, f = int(columns), int(factor)
    if f < 2 or n < 2 * f + 1 or (n - 1) % f:
        raise ValueError(f"the sheet needs (columns - 1) divisible by {f}")
    nc = (n - 1) // f + 1
    row_vertex = np.asarray(row_vertex, np.int64)
    free = np.zeros(n * n, bool)
    free[row_vertex] = True
    coarse_of = -np.ones((nc, nc), np.int64)
    count = 0
    for b in range(nc):
        for a in range(nc):
            if free[(f * b) * n + f * a]:
                coarse_of[b, a] = count
                count += 1
    R = len(row_vertex)
    parent = -np.ones((R, 4), np.int32)
    weights = np.zeros((R, 4), np.float64)
    for row, v in enumerate(row_vertex):
        x, y = int(v) % n, int(v) // n
        a0, b0 = min(x // f, nc - 2), min(y // f, nc - 2)
        fx, fy = (x - float(f) * a0) / f, (y - float(f) * b0) / f
        corners = [(a0, b0), (a0 + 1, b0), (a0, b0 + 1), (a0 + 1, b0 + 1)]
        wts = [(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy]
        for k, ((a, b), w) in enumerate(zip(corners, wts)):
            I = coarse_of[b, a]
            if I < 0 or w == 0.0:
                continue
            parent[row, k] = I
            weights[row, k] = w
    return count, parent, weights


def coarse_vertices(columns: int, row_vertex: np.ndarray, factor: int = 4) -> np.ndarray:
    """Vertex of each coarse node (coarse rows in vertex order, as coarse_layout numbers them)."""
    n, f = int(columns), int(factor)
    nc = (n - 1) // f + 1
    free = np.zeros(n * n, bool)
    free[np.asarray(row_vertex, np.int64)] = True
    out = []
    for b in range(nc):
        for a in range(nc):
            v = (f * b) * n + f * a
            if free[v]:
                out.append(v)
    return np.array(out, np.int64)


def lever_arms(parent: np.ndarray, weights: np.ndarray, row_vertex: np.ndarray, coarse_vertex: np.ndarray,
               positions: np.ndarray, spacing: float, sigma: np.ndarray | None = None) -> np.ndarray:
    """Rotation levers [R, 4, 3] of the parent slots: v_ik = w_ik sigma_ik (x_i - X_I) / H, with x the positions
    [3V] of the fine row's vertex and X those of the coarse node's vertex (zero on empty slots)."""
    x = np.asarray(positions, np.float64).reshape(-1, 3)
    R = parent.shape[0]
    v = np.zeros((R, 4, 3))
    for i in range(R):
        for k in range(4):
            I = parent[i, k]
            if I < 0:
                continue
            s = 1.0 if sigma is None else float(sigma[i, k])
            v[i, k] = float(weights[i, k]) * s * (x[row_vertex[i]] - x[coarse_vertex[I]]) / spacing
    return v


def prolongation(parent: np.ndarray, weights: np.ndarray, coarse_rows: int,
                 levers: np.ndarray | None = None) -> sp.csr_matrix:
    """P = W (x) I3 as a (3R x 3Nc) sparse matrix; with levers [R, 4, 3], the (3R x 6Nc) prolongation whose
    coarse node I has unknowns (t, r) = 6 I + (0..2, 3..5) and slot block [w I | S], S e_a = e_a x v."""
    R = parent.shape[0]
    dof = 3 if levers is None else 6
    rows, cols, vals = [], [], []
    for i in range(R):
        for k in range(4):
            I = parent[i, k]
            if I < 0:
                continue
            for d in range(3):
                rows.append(3 * i + d)
                cols.append(dof * I + d)
                vals.append(float(weights[i, k]))
            if levers is not None:
                for a in range(3):
                    col = np.cross(np.eye(3)[a], levers[i, k])
                    for d in range(3):
                        if col[d] != 0.0:
                            rows.append(3 * i + d)
                            cols.append(dof * I + 3 + a)
                            vals.append(float(col[d]))
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * R, dof * coarse_rows))


def modified_cholesky(A: np.ndarray, floor: float):
    """L with L L^T = A + E, E >= 0 diagonal: column by column, a pivot d_j <= floor * A_jj (or not finite, or
    negative) is replaced by A_jj, the factorisation's pivot floor. Returns (L, replaced column indices)."""
    A = np.array(A, np.float64)
    n = A.shape[0]
    L = np.zeros_like(A)
    replaced = []
    for j in range(n):
        d = A[j, j] - L[j, :j] @ L[j, :j]
        if not (d > floor * A[j, j]) and np.isfinite(d) and A[j, j] > 0:
            d = A[j, j]
            replaced.append(j)
        if not (d > 0 and np.isfinite(d)):
            raise np.linalg.LinAlgError(f"pivot {j} is not positive")
        L[j, j] = np.sqrt(d)
        L[j + 1:, j] = (A[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return L, replaced


def bsr_matrix(row_ptr, col_idx, vals) -> sp.csr_matrix:
    """The 3x3 BSR system (vals [9B], row-major blocks) as a scalar CSR matrix in f64."""
    R = len(row_ptr) - 1
    data = np.asarray(vals, np.float64).reshape(-1, 3, 3)
    return sp.bsr_matrix((data, np.asarray(col_idx, np.int64), np.asarray(row_ptr, np.int64)),
                         shape=(3 * R, 3 * R)).tocsr()


def galerkin(A: sp.spmatrix, P: sp.spmatrix) -> np.ndarray:
    return (P.T @ (A @ P)).toarray()


def block_jacobi_inverse(A: sp.csr_matrix) -> sp.csr_matrix:
    R = A.shape[0] // 3
    blocks = np.zeros((R, 3, 3))
    Ad = A.tocsr()
    for i in range(R):
        blocks[i] = Ad[3 * i:3 * i + 3, 3 * i:3 * i + 3].toarray()
    inv = np.linalg.inv(blocks)
    return sp.bsr_matrix((inv, np.arange(R), np.arange(R + 1)), shape=A.shape).tocsr()


def vcycle(A, Dinv, P, Minv_c, omega, r):
    """The symmetric multiplicative two-level application (V(1,1) with damped block-Jacobi smoothing):
    z1 = w D^-1 r; z2 = z1 + P A_c^-1 P^T (r - A z1); z = z2 + w D^-1 (r - A z2). Its error propagation
    is I - M^-1 A = (I - w D^-1 A)(I - P A_c^-1 P^T A)(I - w D^-1 A)."""
    z1 = omega * (Dinv @ r)
    z2 = z1 + P @ (Minv_c @ (P.T @ (r - A @ z1)))
    return z2 + omega * (Dinv @ (r - A @ z2))


def two_level_pcg(A: sp.csr_matrix, b: np.ndarray, P: sp.spmatrix | None, tolerance: float,
                  max_iterations: int = 20000, omega: float = 0.0, pivot_floor: float = 0.0):
    """f64 PCG with M^-1 = D^-1 + P A_c^-1 P^T (block-Jacobi alone when P is None), or with the V-cycle of
    `vcycle` when omega > 0, x0 = 0, stop when |r|^2 <= tol^2 |r0|^2 (the recursive residual,
    pcg.rs:247-259 order of tests), breakdown when p^T A p <= 0 or alpha is not finite. Returns (x,
    iterations, status, history of |r|/|r0|). pivot_floor > 0 factors A_c by modified_cholesky with that floor
    (the coarse operator of a space with rotation unknowns can be near-dependent or indefinite)."""
    Dinv = block_jacobi_inverse(A)
    if P is not None:
        Ac = galerkin(A, P)
        L = modified_cholesky(Ac, pivot_floor)[0] if pivot_floor > 0 else np.linalg.cholesky(Ac)
        Linv = np.linalg.inv(L)
        Minv_c = Linv.T @ Linv

        if omega > 0:
            def apply(r):
                return vcycle(A, Dinv, P, Minv_c, omega, r)
        else:
            def apply(r):
                return Dinv @ r + P @ (Minv_c @ (P.T @ r))
    else:
        def apply(r):
            return Dinv @ r
    b = np.asarray(b, np.float64)
    x = np.zeros_like(b)
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz = r @ z
    rr0 = r @ r
    history = [1.0]
    if rr0 <= 0:
        return x, 0, 1, history
    for it in range(1, max_iterations + 1):
        q = A @ p
        pq = p @ q
        alpha = rz / pq if pq != 0 else np.inf
        if not (pq > 0 and np.isfinite(alpha)):
            return x, it - 1, 3, history
        x = x + alpha * p
        r = r - alpha * q
        z = apply(r)
        rz_new = r @ z
        rr = r @ r
        history.append(float(np.sqrt(rr / rr0)))
        if rr <= tolerance * tolerance * rr0:
            return x, it, 1, history
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iterations, 2, history
