"""Training of the learned prolongation (run on a GPU: `python -m drape_amd.learned.train --out FILE`).

Data: Newton systems assembled by the library along pseudo-transient continuation runs of sphere drapes whose
parameters are drawn from the configured ranges (E log-uniform in [1e5, 1e7] Pa, bending stiffness log-uniform
in [1e-6, 1e-4] N m, sphere radius uniform in [0.1, 0.25] m, release height uniform in [0.02, 0.1] m), at n65
and n129. For each system: the first Newton system of the next PTC step (x_hat = x + h^2 g, M/h^2, mu = 0) and its
solution by FP64-refined PCG to 1e-9.

Loss: the convergence the preconditioner is for -- the A-norm error |x* - x_k|_A^2 / |x*|_A^2 after k PCG
iterations (from x = 0, with M^-1 = D^-1 + P A_c^-1 P^T and A_c = P^T A P) on the system's own right-hand side,
x* its FP64-refined solution; differentiated through the iterations with torch autograd in float64, Adam. The
network starts at the bilinear prolongation (zero output layer).

The network, features and weight map are those of the HIP inference kernels (drape_amd/csrc/kernels/gnn.hip).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import drape_amd as D  # noqa: E402
from drape_amd.learned import F_IN, HIDDEN, LAYERS, param_count  # noqa: E402


# ---- one system as torch tensors -------------------------------------------------------------------
class System:
    def __init__(self, columns, row_vertex, row_ptr, col_idx, vals, parent, w_bil, coarse_rows, device,
                 dtype=torch.float64):
        self.device, self.dtype = device, dtype
        row_vertex = np.asarray(row_vertex, np.int64)
        row_ptr = np.asarray(row_ptr, np.int64)
        col_idx = np.asarray(col_idx, np.int64)
        blocks = np.asarray(vals, np.float64).reshape(-1, 3, 3)
        R = len(row_ptr) - 1
        self.R, self.Nc = R, int(coarse_rows)
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
        parent = np.asarray(parent, np.int64).reshape(R, 4)
        w_bil = np.asarray(w_bil, np.float64).reshape(R, 4)
        f = np.stack([w_bil[:, 0], w_bil[:, 1], w_bil[:, 2], w_bil[:, 3], fx, fy,
                      Dn[:, 0, 0] - 1.0, Dn[:, 1, 1] - 1.0, Dn[:, 2, 2] - 1.0,
                      0.5 * (Dn[:, 0, 1] + Dn[:, 1, 0]), 0.5 * (Dn[:, 0, 2] + Dn[:, 2, 0]),
                      0.5 * (Dn[:, 1, 2] + Dn[:, 2, 1]), strength], axis=1)
        c = nb / (3.0 * np.sqrt(s[rows] * s[col_idx]))
        norm = np.zeros(R)
        np.add.at(norm, rows[off], c[off])
        t = lambda a: torch.as_tensor(a, device=device)
        self.f = t(f).to(dtype)
        self.msg_rows, self.msg_cols = t(rows[off]), t(col_idx[off])
        self.msg_w = t(c[off] / (norm[rows[off]] + 1e-6)).to(dtype)
        self.parent = t(parent)
        self.w_bil = t(w_bil).to(dtype)
        self.valid = self.w_bil > 0
        self.mass = self.w_bil.sum(dim=1)
        # the fine matrix (for products) and the Galerkin contribution lists
        ii = np.repeat(np.arange(3), 3)
        jj = np.tile(np.arange(3), 3)
        rr = (3 * rows[:, None] + ii[None, :]).ravel()
        cc = (3 * col_idx[:, None] + jj[None, :]).ravel()
        self.Acsr = torch.sparse_coo_tensor(torch.stack([t(rr), t(cc)]), t(blocks.reshape(-1, 9).ravel()).to(dtype),
                                            size=(3 * R, 3 * R)).coalesce().to_sparse_csr()
        k, l = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
        k, l = k.ravel(), l.ravel()
        pi = parent[rows][:, k]          # [B, 16]
        pj = parent[col_idx][:, l]
        ok = (pi >= 0) & (pj >= 0)
        bsel, ksel = np.nonzero(ok)
        self.g_block = t(bsel)
        self.g_wi = t(rows[bsel] * 4 + k[ksel])
        self.g_wj = t(col_idx[bsel] * 4 + l[ksel])
        self.g_dst = t(pi[bsel, ksel] * self.Nc + pj[bsel, ksel])
        self.blocks = t(blocks).to(dtype)


def forward(theta: torch.Tensor, s: System) -> torch.Tensor:
    """Prolongation weights [R, 4] from the flat parameter vector (float64 or float32 tensor)."""
    H = HIDDEN
    o = 0

    def take(n, shape):
        nonlocal o
        v = theta[o:o + n].reshape(shape)
        o += n
        return v

    W0, b0 = take(H * F_IN, (H, F_IN)), take(H, (H,))
    h = torch.relu(s.f @ W0.T + b0)
    for _ in range(LAYERS):
        Ws, Wn, b = take(H * H, (H, H)), take(H * H, (H, H)), take(H, (H,))
        m = torch.zeros_like(h).index_add_(0, s.msg_rows, s.msg_w[:, None] * h[s.msg_cols])
        h = torch.relu(h @ Ws.T + m @ Wn.T + b)
    Wo, bo = take(4 * H, (4, H)), take(4, (4,))
    delta = h @ Wo.T + bo
    z = torch.where(s.valid, torch.log(torch.where(s.valid, s.w_bil, torch.ones_like(s.w_bil))) + delta,
                    torch.full_like(delta, -math.inf))
    soft = torch.softmax(z, dim=1)
    return s.mass[:, None] * torch.where(s.valid, soft, torch.zeros_like(soft))


def smooth_vectors(s: System, count: int, sweeps: int = 5, omega: float = 0.6, seed: int = 0) -> torch.Tensor:
    g = torch.Generator(device="cpu").manual_seed(seed)
    E = torch.randn(3 * s.R, count, generator=g, dtype=s.dtype).to(s.device)
    Dinv = torch.linalg.inv(s.diag)
    for _ in range(sweeps):
        AE = (s.Acsr @ E).reshape(s.R, 3, count)
        E = E - omega * (Dinv @ AE).reshape(3 * s.R, count)
    return E


def loss(w: torch.Tensor, s: System, E: torch.Tensor, AE: torch.Tensor, den: torch.Tensor) -> torch.Tensor:
    Nc, R, K = s.Nc, s.R, E.shape[1]
    wf = w.reshape(-1)
    coef = wf[s.g_wi] * wf[s.g_wj]
    contrib = coef[:, None, None] * s.blocks[s.g_block]
    Ac = torch.zeros(Nc * Nc, 3, 3, dtype=s.dtype, device=s.device).index_add_(0, s.g_dst, contrib)
    Ac = Ac.reshape(Nc, Nc, 3, 3).permute(0, 2, 1, 3).reshape(3 * Nc, 3 * Nc)
    Ac = 0.5 * (Ac + Ac.T)
    L, info = torch.linalg.cholesky_ex(Ac)
    if int(info) != 0:
        raise np.linalg.LinAlgError("coarse operator not positive definite")
    # P^T (A e)
    slots = s.parent.reshape(-1)
    okslot = slots >= 0
    rows_of_slot = torch.arange(4 * R, device=s.device) // 4
    AE3 = AE.reshape(R, 3, K)
    rhs = torch.zeros(Nc, 3, K, dtype=s.dtype, device=s.device).index_add_(
        0, slots[okslot], wf[okslot][:, None, None] * AE3[rows_of_slot[okslot]])
    c = torch.cholesky_solve(rhs.reshape(3 * Nc, K), L).reshape(Nc, 3, K)
    Pc = torch.zeros(R, 3, K, dtype=s.dtype, device=s.device).index_add_(
        0, rows_of_slot[okslot], wf[okslot][:, None, None] * c[slots[okslot]])
    eps = E - Pc.reshape(3 * R, K)
    num = (eps * (s.Acsr @ eps)).sum(dim=0)
    return (num / den).mean()


def pcg_loss(w: torch.Tensor, s: System, b: torch.Tensor, xstar: torch.Tensor, k: int) -> torch.Tensor:
    """|x* - x_k|_A^2 / |x*|_A^2 after k iterations of PCG from 0 with M^-1 = D^-1 + P A_c^-1 P^T: the
    convergence the preconditioner is for, differentiated through the iterations."""
    Nc, R = s.Nc, s.R
    wf = w.reshape(-1)
    coef = wf[s.g_wi] * wf[s.g_wj]
    contrib = coef[:, None, None] * s.blocks[s.g_block]
    Ac = torch.zeros(Nc * Nc, 3, 3, dtype=s.dtype, device=s.device).index_add_(0, s.g_dst, contrib)
    Ac = Ac.reshape(Nc, Nc, 3, 3).permute(0, 2, 1, 3).reshape(3 * Nc, 3 * Nc)
    Ac = 0.5 * (Ac + Ac.T)
    L, info = torch.linalg.cholesky_ex(Ac)
    if int(info) != 0:
        raise np.linalg.LinAlgError("coarse operator not positive definite")
    slots = s.parent.reshape(-1)
    ok = slots >= 0
    srow = torch.arange(4 * R, device=s.device) // 4
    ws, sl, sr = wf[ok], slots[ok], srow[ok]

    def apply(r):
        r3 = r.reshape(R, 3)
        rc = torch.zeros(Nc, 3, dtype=s.dtype, device=s.device).index_add_(0, sl, ws[:, None] * r3[sr])
        yc = torch.cholesky_solve(rc.reshape(3 * Nc, 1), L).reshape(Nc, 3)
        py = torch.zeros(R, 3, dtype=s.dtype, device=s.device).index_add_(0, sr, ws[:, None] * yc[sl])
        return (s.dinv @ r3[:, :, None]).reshape(R, 3).reshape(-1) + py.reshape(-1)

    x = torch.zeros_like(b)
    r = b.clone()
    z = apply(r)
    p = z
    rz = r @ z
    for _ in range(k):
        q = s.Acsr @ p
        alpha = rz / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        z = apply(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    e = xstar - x
    return (e @ (s.Acsr @ e)) / (xstar @ (s.Acsr @ xstar))


# ---- data --------------------------------------------------------------------------------------------
def draw(rng: np.random.Generator):
    return {"E": float(10 ** rng.uniform(5, 7)), "bend": float(10 ** rng.uniform(-6, -4)),
            "radius": float(rng.uniform(0.10, 0.25)), "height": float(rng.uniform(0.02, 0.10))}


def generate(columns: int, params: dict, states: int, device, log=print):
    """Newton systems along a PTC run of one drape (the library assembles them on the GPU)."""
    scene = D.drape_config_scene(columns, radius=params["radius"], clearance=params["height"], E=params["E"],
                                 bend=params["bend"])
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays, D.default_rules(preconditioner=1))
    solver = D.Drape(arrays, D.default_rules(preconditioner=1, refine_fp64=1))
    nc, parent, w_bil, _ = ctx.two_level_layout()
    row_ptr, col_idx, row_vertex = arrays.array("row_ptr"), arrays.array("col_idx"), arrays.array("row_vertex")
    x, h = arrays.array("rest_positions"), ctx.rules.h0
    out = []
    for k in range(states):
        run = ctx.solve_equilibrium(start=x, h_start=h, cap=1)
        x, h = run["positions"], run["result"].final_h
        x_hat = x.copy()
        x_hat[2::3] -= h * h * scene.gravity
        solver.set_state(x, x_hat, h, 0.0)
        solver.residual()
        rhs, vals = solver.system()
        st, _ = solver.solve_once(1e-9, 20000)
        if D.PCG_STATUS[st.status] != "converged":
            continue   # no reference solution for this system
        sysm = System(columns, row_vertex, row_ptr, col_idx, vals, parent, w_bil, nc, device)
        sysm.b = torch.as_tensor(rhs.astype(np.float64), device=device)
        sysm.xstar = torch.as_tensor(solver.solution_fp64(), device=device)
        out.append(sysm)
    ctx.close()
    solver.close()
    log(f"[train] n{columns} {params}: {len(out)} systems")
    return out


def attach(s: System, count: int, seed: int):
    """Smooth test vectors, A e and |e|_A^2 of one system."""
    R = s.R
    diag = torch.zeros(R, 3, 3, dtype=s.dtype, device=s.device)
    # diagonal blocks from the matrix: entries (3i + a, 3i + b)
    A = s.Acsr.to_sparse_coo().coalesce()
    idx, val = A.indices(), A.values()
    same = (idx[0] // 3) == (idx[1] // 3)
    diag.index_put_((idx[0][same] // 3, idx[0][same] % 3, idx[1][same] % 3), val[same], accumulate=True)
    s.diag = diag
    s.dinv = torch.linalg.inv(diag)
    E = smooth_vectors(s, count, seed=seed)
    AE = s.Acsr @ E
    s.E, s.AE, s.den = E, AE, (E * AE).sum(dim=0)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes65", type=int, default=8)
    ap.add_argument("--scenes129", type=int, default=4)
    ap.add_argument("--states", type=int, default=6)
    ap.add_argument("--vectors", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--budget-s", type=float, default=420.0)
    ap.add_argument("--k", type=int, default=20, help="PCG iterations unrolled in the loss")
    ap.add_argument("--log-loss", action="store_true", help="minimise the mean log of the A-norm error ratio")
    args = ap.parse_args()
    device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    t0 = time.time()
    rng = np.random.default_rng(args.seed)
    systems, scenes = [], []
    for columns, count in ((65, args.scenes65), (129, args.scenes129)):
        for _ in range(count):
            p = draw(rng)
            scenes.append({"columns": columns, **p})
            systems += generate(columns, p, args.states, device)
    kept = []
    for i, s in enumerate(systems):
        attach(s, args.vectors, seed=1000 + i)
        try:   # systems whose bilinear coarse operator is not positive definite teach nothing here
            pcg_loss(s.w_bil, s, s.b, s.xstar, 1)
            kept.append(s)
        except np.linalg.LinAlgError:
            pass
    print(f"[train] kept {len(kept)} of {len(systems)} systems", flush=True)
    systems = kept
    # hold out every fourth system for validation
    val = systems[3::4]
    train = [s for i, s in enumerate(systems) if i % 4 != 3]
    print(f"[train] {len(train)} training and {len(val)} validation systems, {time.time() - t0:.1f} s", flush=True)

    hidden_init = np.random.default_rng(args.seed)
    theta0 = np.zeros(param_count())
    o = 0
    for shape, fan in [((HIDDEN, F_IN), F_IN), ((HIDDEN,), 0)] + \
            [((HIDDEN, HIDDEN), HIDDEN), ((HIDDEN, HIDDEN), HIDDEN), ((HIDDEN,), 0)] * LAYERS + [((4, HIDDEN), 0), ((4,), 0)]:
        n = int(np.prod(shape))
        if fan:
            theta0[o:o + n] = hidden_init.standard_normal(n) * math.sqrt((2.0 if shape[1] == F_IN else 1.0) / fan)
        o += n
    theta = torch.tensor(theta0, dtype=torch.float64, device=device, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=args.lr)

    def mean_loss(group, th):
        with torch.no_grad():
            vals = [pcg_loss(forward(th, s), s, s.b, s.xstar, args.k).item() for s in group]
            return float(np.exp(np.mean(np.log(vals)))) if args.log_loss else float(np.mean(vals))

    zero = torch.zeros_like(theta)
    base_train, base_val = mean_loss(train, zero), mean_loss(val, zero)
    print(f"[train] bilinear loss (A-norm error after {args.k} PCG iterations): train {base_train:.4e} "
          f"validation {base_val:.4e}", flush=True)
    history = []
    order = np.random.default_rng(args.seed + 1)
    for epoch in range(args.epochs):
        for i in order.permutation(len(train)):
            s = train[i]
            opt.zero_grad()
            try:
                L = pcg_loss(forward(theta, s), s, s.b, s.xstar, args.k)
                if args.log_loss:
                    L = torch.log(L)
            except np.linalg.LinAlgError:
                continue
            L.backward()
            opt.step()
        lt, lv = mean_loss(train, theta.detach()), mean_loss(val, theta.detach())
        history.append((epoch, lt, lv))
        print(f"[train] epoch {epoch}: train {lt:.4e} validation {lv:.4e} ({time.time() - t0:.0f} s)", flush=True)
        if time.time() - t0 > args.budget_s:
            break
    meta = {"hidden": HIDDEN, "layers": LAYERS, "features": F_IN,
            "params": [float(v) for v in theta.detach().cpu().numpy().astype(np.float32)],
            "training": {"scenes": scenes, "states_per_scene": args.states, "test_vectors": args.vectors,
                         "epochs": len(history), "lr": args.lr, "seed": args.seed, "pcg_iterations_in_loss": args.k,
                         "log_loss": bool(args.log_loss),
                         "loss_bilinear": {"train": base_train, "validation": base_val},
                         "loss_learned": {"train": history[-1][1], "validation": history[-1][2]}}}
    with open(args.out, "w") as f:
        json.dump(meta, f)
    print(f"[train] wrote {args.out}: validation loss {base_val:.4f} -> {history[-1][2]:.4f}", flush=True)


if __name__ == "__main__":
    main()
