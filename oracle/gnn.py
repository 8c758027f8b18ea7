"""Oracle for the learned prolongation of the two-level preconditioner: a numpy f64 statement of the graph
network's features, forward pass, the two-level loss and its gradient. Test infrastructure only: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker.

There is no reference counterpart: the reference has no two-level preconditioner. Pins, all by mathematics
(tests/test_gnn_oracle.py):
  - with a zero output layer the network returns the bilinear weights exactly;
  - the analytic gradient (Galerkin orthogonality makes the coarse solve drop out of it) matches central
    finite differences of the loss;
  - the loss is the A-norm of the two-level coarse-projection error, 0 for errors in range(P) and 1 for
    A-orthogonal ones;
  - the training loss (pcg_loss: the A-norm error after k two-level PCG iterations) is 1 at k = 0 and equals
    the training script's value; that script's autograd gradient of it matches finite differences here.

Model (shared with the HIP inference kernels and the training script):
  features f_i (13): the four bilinear parent weights, the position (fx, fy) inside the coarse cell, the
  diagonal block divided by s_i = trace(A_ii)/3 minus the identity (xx, yy, zz, and the symmetric parts
  xy, xz, yz), and the coupling strength sum_j |A_ij|_F / (3 s_i) over the off-diagonal blocks;
  edge weights c_ij = |A_ij|_F / (3 sqrt(s_i s_j)) over the off-diagonal blocks;
  h0 = relu(W0 f + b0); three layers h <- relu(Ws h_i + Wn m_i + b), m_i = sum_j c_ij h_j / (sum_j c_ij + 1e-6);
  delta = Wo h + bo (4); weights = mass_i softmax(log w_bilinear + delta) over the parent slots with a
  non-zero bilinear weight, mass_i their bilinear sum; with the rotation head (a coarse space with rotation
  unknowns, two_level.prolongation with levers), sigma = exp(clip(Wr h + br, -3, 3)) scales each slot's lever
  v = w sigma (x_i - X_I) / H, so a zero head keeps the levers of the shape.
Parameters are one flat float array: W0 [H x 13], b0 [H], then per layer Ws [H x H], Wn [H x H], b [H], then
Wo [4 x H], bo [4], and optionally Wr [4 x H], br [4] (row-major), H = 32.
With rotations the gradient also runs through the levers: dL/dv = -2 (A eps)_i x c_r / |e|_A^2 per slot, which
adds dL/dv . v / w to dL/dw and gives dL/dsigma = dL/dv . v / sigma (its finite-difference pin is in
tests/test_gnn_oracle.py).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import two_level as TL

F_IN = 13
HIDDEN = 32
LAYERS = 3


def param_count(hidden: int = HIDDEN, rotation_head: bool = False) -> int:
    base = hidden * F_IN + hidden + LAYERS * (2 * hidden * hidden + hidden) + 4 * hidden + 4
    return base + (4 * hidden + 4 if rotation_head else 0)


def unpack(theta: np.ndarray, hidden: int = HIDDEN) -> dict:
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0

    def take(name, shape):
        nonlocal o
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n

    take("W0", (hidden, F_IN))
    take("b0", (hidden,))
    for l in range(LAYERS):
        take(f"Ws{l}", (hidden, hidden))
        take(f"Wn{l}", (hidden, hidden))
        take(f"bl{l}", (hidden,))
    take("Wo", (4, hidden))
    take("bo", (4,))
    if theta.size > o:
        take("Wr", (4, hidden))
        take("br", (4,))
    assert o == theta.size, (o, theta.size)
    return out


def pack(p: dict, hidden: int = HIDDEN) -> np.ndarray:
    parts = [p["W0"], p["b0"]]
    for l in range(LAYERS):
        parts += [p[f"Ws{l}"], p[f"Wn{l}"], p[f"bl{l}"]]
    parts += [p["Wo"], p["bo"]]
    if "Wr" in p:
        parts += [p["Wr"], p["br"]]
    return np.concatenate([np.ravel(x) for x in parts])


def init_params(seed: int = 0, hidden: int = HIDDEN, scale: float = 1.0, rotation_head: bool = False) -> np.ndarray:
    """He-style random hidden layers, zero output layers (the network starts at the bilinear weights and, with
    the rotation head, at the levers of the shape)."""
    rng = np.random.default_rng(seed)
    p = {"W0": rng.standard_normal((hidden, F_IN)) * scale * np.sqrt(2.0 / F_IN), "b0": np.zeros(hidden)}
    for l in range(LAYERS):
        p[f"Ws{l}"] = rng.standard_normal((hidden, hidden)) * scale * np.sqrt(1.0 / hidden)
        p[f"Wn{l}"] = rng.standard_normal((hidden, hidden)) * scale * np.sqrt(1.0 / hidden)
        p[f"bl{l}"] = np.zeros(hidden)
    p["Wo"] = np.zeros((4, hidden))
    p["bo"] = np.zeros(4)
    if rotation_head:
        p["Wr"] = np.zeros((4, hidden))
        p["br"] = np.zeros(4)
    return pack(p, hidden)


class Graph:
    """Features and edge weights of one assembled system (row_ptr, col_idx, vals [9B]) on a grid sheet."""

    def __init__(self, columns, row_vertex, row_ptr, col_idx, vals, positions=None, spacing=None):
        """positions [3V] and the coarse spacing H give the coarse space rotation unknowns (levers from them)."""
        row_vertex = np.asarray(row_vertex, np.int64)
        row_ptr = np.asarray(row_ptr, np.int64)
        col_idx = np.asarray(col_idx, np.int64)
        blocks = np.asarray(vals, np.float64).reshape(-1, 3, 3)
        R = len(row_ptr) - 1
        self.R = R
        self.coarse_rows, self.parent, self.w_bil = TL.coarse_layout(columns, row_vertex)
        n = int(columns)
        nc = (n - 1) // 4 + 1
        x, y = row_vertex % n, row_vertex // n
        a0, b0 = np.minimum(x // 4, nc - 2), np.minimum(y // 4, nc - 2)
        fx, fy = (x - 4.0 * a0) / 4.0, (y - 4.0 * b0) / 4.0
        rows = np.repeat(np.arange(R), np.diff(row_ptr))
        diag = rows == col_idx
        Aii = np.zeros((R, 3, 3))
        Aii[rows[diag]] = blocks[diag]
        s = np.trace(Aii, axis1=1, axis2=2) / 3.0
        Dn = Aii / s[:, None, None]
        nb = np.sqrt((blocks ** 2).sum(axis=(1, 2)))
        off = ~diag
        strength = np.zeros(R)
        np.add.at(strength, rows[off], nb[off])
        strength /= 3.0 * s
        self.f = np.stack([self.w_bil[:, 0], self.w_bil[:, 1], self.w_bil[:, 2], self.w_bil[:, 3], fx, fy,
                           Dn[:, 0, 0] - 1.0, Dn[:, 1, 1] - 1.0, Dn[:, 2, 2] - 1.0,
                           0.5 * (Dn[:, 0, 1] + Dn[:, 1, 0]), 0.5 * (Dn[:, 0, 2] + Dn[:, 2, 0]),
                           0.5 * (Dn[:, 1, 2] + Dn[:, 2, 1]), strength], axis=1)
        c = np.where(off, nb / (3.0 * np.sqrt(s[rows] * s[col_idx])), 0.0)
        self.c = c
        norm = np.zeros(R)
        np.add.at(norm, rows, c)
        # row-normalised message matrix C: m = C h
        self.C = sp.csr_matrix((c / (norm[rows] + 1e-6), (rows, col_idx)), shape=(R, R))
        self.A = TL.bsr_matrix(row_ptr, col_idx, blocks.reshape(-1))
        self.valid = self.w_bil > 0
        self.mass = self.w_bil.sum(axis=1)
        self.d = None  # unweighted lever directions (x_i - X_I) / H [R, 4, 3] of a coarse space with rotations
        if positions is not None:
            cv = TL.coarse_vertices(columns, row_vertex)
            self.d = TL.lever_arms(self.parent, np.where(self.valid, 1.0, 0.0), row_vertex, cv, positions, spacing)


def forward(theta, g: Graph, hidden: int = HIDDEN, keep: bool = False):
    """Prolongation weights [R, 4] (0 on empty slots); with keep, also the activations for backward."""
    p = unpack(theta, hidden)
    u0 = g.f @ p["W0"].T + p["b0"]
    h = np.maximum(u0, 0.0)
    cache = {"u": [u0], "h": [h], "m": []}
    for l in range(LAYERS):
        m = g.C @ h
        u = h @ p[f"Ws{l}"].T + m @ p[f"Wn{l}"].T + p[f"bl{l}"]
        h = np.maximum(u, 0.0)
        cache["m"].append(m)
        cache["u"].append(u)
        cache["h"].append(h)
    delta = h @ p["Wo"].T + p["bo"]
    z = np.where(g.valid, np.log(np.where(g.valid, g.w_bil, 1.0)) + delta, -np.inf)
    zmax = z.max(axis=1, keepdims=True)
    ez = np.where(g.valid, np.exp(z - zmax), 0.0)
    soft = ez / ez.sum(axis=1, keepdims=True)
    w = g.mass[:, None] * soft
    sigma = None
    if "Wr" in p and g.d is not None:
        o = h @ p["Wr"].T + p["br"]
        sigma = np.exp(np.clip(o, -3.0, 3.0))
        cache["o"] = o
    if keep:
        cache["soft"] = soft
        cache["sigma"] = sigma
        return w, cache
    return w if sigma is None else (w, sigma)


def smooth_vectors(g: Graph, count: int, sweeps: int = 5, omega: float = 0.6, seed: int = 0) -> np.ndarray:
    """Error vectors left by damped block-Jacobi sweeps on random starts (what the coarse space must catch)."""
    Dinv = TL.block_jacobi_inverse(g.A)
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((3 * g.R, count))
    for _ in range(sweeps):
        E = E - omega * (Dinv @ (g.A @ E))
    return E


def levers(w: np.ndarray, g: Graph, sigma: np.ndarray | None = None) -> np.ndarray | None:
    """The slot levers v = w sigma d [R, 4, 3] of a graph with rotations (None without)."""
    if g.d is None:
        return None
    s = np.ones_like(w) if sigma is None else sigma
    return (w * s)[:, :, None] * g.d


def loss_and_weight_gradient(w: np.ndarray, g: Graph, E: np.ndarray, sigma: np.ndarray | None = None,
                             with_sigma: bool = False):
    """L = mean_e |e - P c|_A^2 / |e|_A^2 with c = A_c^-1 P^T A e, and dL/dw [R, 4]. By Galerkin
    orthogonality (P^T A (e - P c) = 0) the derivative through c vanishes: dL/dP = -2 (A eps) c^T / |e|_A^2.
    With rotations (g.d) the levers v = w sigma d carry the gradient too; with_sigma also returns dL/dsigma."""
    lev = levers(w, g, sigma)
    P = TL.prolongation(g.parent, w, g.coarse_rows, lev)
    dof = 3 if lev is None else 6
    Ac = (P.T @ (g.A @ P)).toarray()
    # the device's factorisation: with rotations, the modified Cholesky of its pivot floor
    L = np.linalg.cholesky(Ac) if lev is None else TL.modified_cholesky(0.5 * (Ac + Ac.T), 1e-10)[0]
    AE = g.A @ E
    rhs = P.T @ AE
    Cc = np.linalg.solve(L.T, np.linalg.solve(L, rhs))
    eps = E - P @ Cc
    Aeps = g.A @ eps
    num = np.einsum("ij,ij->j", eps, Aeps)
    den = np.einsum("ij,ij->j", E, AE)
    loss = float(np.mean(num / den))
    grad = np.zeros_like(w)
    gsig = np.zeros_like(w)
    R = g.R
    Aeps3 = Aeps.reshape(R, 3, -1)
    C3 = Cc.reshape(-1, dof, Cc.shape[1])
    for k in range(4):
        I = g.parent[:, k]
        ok = I >= 0
        cc = C3[np.where(ok, I, 0)]
        contrib = -2.0 * np.einsum("idn,idn->in", Aeps3, cc[:, :3]) / den[None, :]
        gk = contrib.mean(axis=1)
        if lev is not None:
            gv = (-2.0 * np.cross(Aeps3, cc[:, 3:], axis=1) / den[None, None, :]).mean(axis=2)
            gvv = np.einsum("id,id->i", gv, lev[:, k])
            gk = gk + gvv / np.where(ok, w[:, k], 1.0)
            s = np.ones(R) if sigma is None else sigma[:, k]
            gsig[:, k] = np.where(ok, gvv / s, 0.0)
        grad[:, k] = np.where(ok, gk, 0.0)
    if with_sigma:
        return loss, grad, gsig
    return loss, grad


def loss(theta, g: Graph, E: np.ndarray, hidden: int = HIDDEN) -> float:
    out = forward(theta, g, hidden)
    w, sigma = out if isinstance(out, tuple) else (out, None)
    return loss_and_weight_gradient(w, g, E, sigma)[0]


def backward(theta, g: Graph, E: np.ndarray, hidden: int = HIDDEN):
    """(loss, dL/dtheta) by hand: softmax, output layer, three message-passing layers, input layer."""
    p = unpack(theta, hidden)
    w, cache = forward(theta, g, hidden, keep=True)
    sigma = cache["sigma"]
    L, gw, gs = loss_and_weight_gradient(w, g, E, sigma, with_sigma=True)
    soft = cache["soft"]
    gw = np.where(g.valid, gw, 0.0)
    # w = mass softmax(z): dz_j = mass p_j (gw_j - sum_k gw_k p_k)
    gz = g.mass[:, None] * soft * (gw - (gw * soft).sum(axis=1, keepdims=True))
    gz = np.where(g.valid, gz, 0.0)
    grads = {"Wo": gz.T @ cache["h"][-1], "bo": gz.sum(axis=0)}
    gh = gz @ p["Wo"]
    if "Wr" in p:
        if sigma is None:  # a rotation head without a rotation coarse space: unused
            grads["Wr"], grads["br"] = np.zeros_like(p["Wr"]), np.zeros_like(p["br"])
        else:
            o = cache["o"]
            go = np.where((o > -3.0) & (o < 3.0), gs * sigma, 0.0)
            grads["Wr"], grads["br"] = go.T @ cache["h"][-1], go.sum(axis=0)
            gh = gh + go @ p["Wr"]
    for l in reversed(range(LAYERS)):
        gu = gh * (cache["u"][l + 1] > 0)
        h_prev, m = cache["h"][l], cache["m"][l]
        grads[f"Ws{l}"] = gu.T @ h_prev
        grads[f"Wn{l}"] = gu.T @ m
        grads[f"bl{l}"] = gu.sum(axis=0)
        gh = gu @ p[f"Ws{l}"] + g.C.T @ (gu @ p[f"Wn{l}"])
    gu0 = gh * (cache["u"][0] > 0)
    grads["W0"] = gu0.T @ g.f
    grads["b0"] = gu0.sum(axis=0)
    return L, pack(grads, hidden)


def pcg_loss(w: np.ndarray, g: Graph, b: np.ndarray, xstar: np.ndarray, k: int) -> float:
    """|x* - x_k|_A^2 / |x*|_A^2 after k iterations of f64 PCG from x = 0 on A x = b with the additive
    two-level preconditioner M^-1 = D^-1 + P A_c^-1 P^T (the quantity the learned prolongation is trained to
    lower; x* the system's solution)."""
    P = TL.prolongation(g.parent, w, g.coarse_rows)
    Ac = (P.T @ (g.A @ P)).toarray()
    Ac = 0.5 * (Ac + Ac.T)
    L = np.linalg.cholesky(Ac)
    Dinv = TL.block_jacobi_inverse(g.A)

    def apply(r):
        return Dinv @ r + P @ np.linalg.solve(L.T, np.linalg.solve(L, P.T @ r))

    x = np.zeros_like(b)
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz = r @ z
    for _ in range(k):
        q = g.A @ p
        alpha = rz / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        z = apply(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    e = xstar - x
    return float((e @ (g.A @ e)) / (xstar @ (g.A @ xstar)))

