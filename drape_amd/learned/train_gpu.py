"""Training of the learned prolongation on the GPU by the library's own kernels
(`python -m drape_amd.learned.train_gpu --out FILE`).

Data: Newton systems the library assembles along pseudo-transient continuation runs of sphere drapes with
parameters drawn from the configured ranges (E log-uniform in [1e5, 1e7] Pa, bending stiffness log-uniform in
[1e-6, 1e-4] N m, sphere radius uniform in [0.1, 0.25] m, release height uniform in [0.02, 0.1] m): the first
Newton system of each PTC step (x_hat = x + h^2 g, M/h^2, mu = 0). Each system stays assembled in a context of
its own, with its error vectors made on the device (random starts, five damped block-Jacobi sweeps).

Loss: the projection loss of the coarse space, mean_e |e - P c_e|_A^2 / |e|_A^2 with c_e = A_c^-1 P^T A e,
and its parameter gradient, both from the device's forward and backward kernels
(drape_prolongation_gradient); Adam on the host over the mean gradient of the systems. The network starts at
the bilinear prolongation (zero output layer). With --rotations the coarse space has three rotation unknowns
per node (levers from the shape of each system) and the network a rotation head, which starts at the levers of
the shape. With --select the kept parameters are those of the fewest two-level PCG iterations on separate
validation drapes among the checkpoints, the start included.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import drape_amd as D  # noqa: E402
from drape_amd.learned import F_IN, HIDDEN, LAYERS, ROTATION_RULES, param_count  # noqa: E402

RULES = {}  # coarse-space rules of every training context (ROTATION_RULES under --rotations)


def init_params(seed: int = 0, rotation_head: bool = False) -> np.ndarray:
    """He-style random hidden layers and zero output layers: the network starts at the bilinear weights (and at
    the levers of the shape)."""
    rng = np.random.default_rng(seed)
    H = HIDDEN
    parts = [rng.standard_normal((H, F_IN)) * np.sqrt(2.0 / F_IN), np.zeros(H)]
    for _ in range(LAYERS):
        parts += [rng.standard_normal((H, H)) * np.sqrt(1.0 / H), rng.standard_normal((H, H)) * np.sqrt(1.0 / H),
                  np.zeros(H)]
    parts += [np.zeros((4, H)), np.zeros(4)]
    if rotation_head:
        parts += [np.zeros((4, H)), np.zeros(4)]
    theta = np.concatenate([np.ravel(p) for p in parts])
    assert theta.size == param_count(rotation_head=rotation_head)
    return theta


def draw(rng: np.random.Generator) -> dict:
    return {"E": float(10 ** rng.uniform(5, 7)), "bend": float(10 ** rng.uniform(-6, -4)),
            "radius": float(rng.uniform(0.10, 0.25)), "height": float(rng.uniform(0.02, 0.10))}


class TrainingSystem:
    """One assembled Newton system in a context of its own, with its error vectors on the device."""

    def __init__(self, arrays: D.SceneArrays, x: np.ndarray, x_hat: np.ndarray, h: float, vectors: int,
                 seed: int, sweeps: int = 5):
        self.ctx = D.Drape(arrays, D.default_rules(preconditioner=1, **RULES))
        self.ctx.set_state(x, x_hat, h, 0.0)
        self.ctx.residual()
        self.ctx.training_vectors(count=vectors, sweeps=sweeps, omega=0.6, seed=seed)

    def gradient(self, theta: np.ndarray):
        return self.ctx.prolongation_gradient(theta)

    def pcg_iterations(self, theta, tol: float = 1e-6, cap: int = 20000) -> int:
        """Two-level PCG iterations on the system's own right-hand side (theta None: bilinear weights, with the
        training coarse space)."""
        self.ctx.set_preconditioner("two_level")
        self.ctx.set_prolongation_model(None if theta is None else theta)
        st, _ = self.ctx.solve_once(tol, cap)
        return st.iterations

    def translations_iterations(self, tol: float = 1e-6, cap: int = 20000) -> int:
        """Two-level PCG iterations with the bilinear translations-only coarse space (the reference point of
        the learned prolongation), on the same system."""
        x, x_hat, h = self.state
        c = D.Drape(self.arrays, D.default_rules(preconditioner=1))
        c.set_state(x, x_hat, h, 0.0)
        c.residual()
        st, _ = c.solve_once(tol, cap)
        c.close()
        return st.iterations

    def close(self):
        self.ctx.close()


def collect(columns: int, params: dict, states: int, vectors: int, seed: int, sweeps: int = 5, log=print,
            skip: int = 0) -> list:
    """Training systems along a PTC run of one drape: those of `states` steps after the first `skip`."""
    scene = D.drape_config_scene(columns, radius=params["radius"], clearance=params["height"], E=params["E"],
                                 bend=params["bend"])
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays, D.default_rules(preconditioner=1))
    x, h = arrays.array("rest_positions"), ctx.rules.h0
    out = []
    for k in range(skip + states):
        run = ctx.solve_equilibrium(start=x, h_start=h, cap=1)
        x, h = run["positions"], run["result"].final_h
        if k < skip:
            continue
        x_hat = x.copy()
        x_hat[2::3] -= h * h * scene.gravity
        system = TrainingSystem(arrays, x, x_hat, h, vectors, seed + k, sweeps)
        system.h = h
        system.arrays = arrays
        system.state = (x, x_hat, h)
        # an indefinite system is left out: its bilinear coarse operator is not positive definite, or PCG meets
        # a direction of non-positive curvature on it
        try:
            system.gradient(init_params(0, bool(RULES)).astype(np.float32))
            system.ctx.set_prolongation_model(None)
            st, _ = system.ctx.solve_once(1e-6, 20000)
            spd = D.PCG_STATUS[st.status] == "converged"
        except D.DrapeError:
            spd = False
        if not spd:
            system.close()
            continue
        out.append(system)
    ctx.close()
    if log:
        log(f"[train] n{columns} {params}: {len(out)} systems, h {[round(s.h, 4) for s in out]}")
    return out


def mean_gradient(systems, theta):
    L, g = 0.0, np.zeros(theta.size)
    for s in systems:
        l, gr = s.gradient(theta)
        L += l
        g += gr
    return L / len(systems), g / len(systems)


def adam(systems, theta0, steps: int, lr: float = 2e-3, betas=(0.9, 0.999), eps: float = 1e-8, log=print,
         log_every: int = 25, budget_s: float | None = None, select=None, select_every: int = 10):
    """Adam over the mean device gradient. Returns the parameters kept and the loss history: with `select`
    (parameters -> score, lower is better), those of the best score among the checkpoints every
    `select_every` steps; otherwise those of the lowest mean loss seen."""
    theta = np.asarray(theta0, np.float64).copy()
    m, v = np.zeros_like(theta), np.zeros_like(theta)
    best, best_score, history = theta.copy(), np.inf, []
    if select is not None:  # the start is a candidate too
        best_score = select(theta.astype(np.float32))
    t0 = time.perf_counter()
    for t in range(1, steps + 1):
        L, g = mean_gradient(systems, theta.astype(np.float32))
        history.append(L)
        if select is None and L < best_score:
            best, best_score = theta.copy(), L
        m = betas[0] * m + (1 - betas[0]) * g
        v = betas[1] * v + (1 - betas[1]) * g * g
        mh, vh = m / (1 - betas[0] ** t), v / (1 - betas[1] ** t)
        theta = theta - lr * mh / (np.sqrt(vh) + eps)
        if select is not None and t % select_every == 0:
            score = select(theta.astype(np.float32))
            if score < best_score:
                best, best_score = theta.copy(), score
        if log and (t % log_every == 0 or t == 1):
            log(f"[train] step {t}: loss {L:.5f} (best {best_score:.5f}), {time.perf_counter() - t0:.1f} s")
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    if select is None:
        L, _ = mean_gradient(systems, theta.astype(np.float32))
        history.append(L)
        if L < best_score:
            best = theta.copy()
    return best.astype(np.float32), history


def pcg_objective(systems, theta, k: int) -> float:
    """Mean over the systems of log10 of the relative residual after k two-level PCG iterations on each system's
    own right-hand side (lower is better): what the preconditioner is for, and unlike an iteration count a
    continuous function of the parameters."""
    total = 0.0
    for s in systems:
        s.ctx.set_preconditioner("two_level")
        s.ctx.set_prolongation_model(theta)
        st, _ = s.ctx.solve_once(1e-12, k)
        total += np.log10(max(st.recursive_relative_residual, 1e-30))
    return total / len(systems)


def es_head(systems, theta0, steps: int, k: int, pairs: int = 16, noise: float = 0.05, lr: float = 0.02,
            seed: int = 0, log=print, select=None, select_every: int = 5, budget_s: float | None = None):
    """Antithetic evolution strategies (Salimans et al. 2017) on the network's output layers (the weight and
    rotation heads, over the hidden features of the fixed layers) against pcg_objective, with Adam on the
    estimated gradient. Returns the parameters of the best `select` score among the checkpoints (the start
    included), or of the lowest objective without `select`."""
    from drape_amd.learned import param_count as pc
    rng = np.random.default_rng(seed)
    head = slice(pc() - 4 * HIDDEN - 4, theta0.size)
    theta = np.asarray(theta0, np.float64).copy()
    n = theta[head].size
    m, v = np.zeros(n), np.zeros(n)
    best = theta.copy()
    best_score = select(theta.astype(np.float32)) if select else pcg_objective(systems, theta.astype(np.float32), k)
    t0 = time.perf_counter()
    for t in range(1, steps + 1):
        eps = rng.standard_normal((pairs, n))
        diff = np.zeros(pairs)
        for j in range(pairs):
            tp, tm = theta.copy(), theta.copy()
            tp[head] += noise * eps[j]
            tm[head] -= noise * eps[j]
            diff[j] = pcg_objective(systems, tp.astype(np.float32), k) - pcg_objective(systems, tm.astype(np.float32), k)
        g = (diff[:, None] * eps).sum(axis=0) / (2.0 * pairs * noise)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        theta[head] -= lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        if t % select_every == 0 or t == steps:
            score = select(theta.astype(np.float32)) if select else pcg_objective(systems, theta.astype(np.float32), k)
            if score < best_score:
                best, best_score = theta.copy(), score
            if log:
                obj = pcg_objective(systems, theta.astype(np.float32), k)
                log(f"[train] es step {t}: objective {obj:.4f}, score {score:.2f} (best {best_score:.2f}), "
                    f"{time.perf_counter() - t0:.1f} s")
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    return best.astype(np.float32)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes65", type=int, default=6)
    ap.add_argument("--scenes129", type=int, default=2)
    ap.add_argument("--states", type=int, default=4)
    ap.add_argument("--vectors", type=int, default=8)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--budget-s", type=float, default=300.0)
    ap.add_argument("--sweeps", type=int, default=5, help="block-Jacobi sweeps that smooth the error vectors")
    ap.add_argument("--skip", type=int, default=2, help="PTC steps of each drape before its systems are taken")
    ap.add_argument("--select", action="store_true",
                    help="keep the checkpoint with the fewest PCG iterations on the validation systems")
    ap.add_argument("--rotations", action="store_true",
                    help="a coarse space with rotation unknowns and a network with the rotation head")
    ap.add_argument("--es-steps", type=int, default=0,
                    help="after the gradient steps, evolution-strategy steps on the output layers against the PCG "
                         "residual after --es-k iterations")
    ap.add_argument("--es-k", type=int, default=15)
    ap.add_argument("--es-budget-s", type=float, default=300.0)
    args = ap.parse_args()
    if args.rotations:
        RULES.update(ROTATION_RULES)
    rng = np.random.default_rng(args.seed)
    train, valid, held = [], [], []
    for columns, count in ((65, args.scenes65), (129, args.scenes129)):
        for k in range(count):
            train += collect(columns, draw(rng), args.states, args.vectors, 100 * k, args.sweeps, skip=args.skip)
        valid += collect(columns, draw(rng), 2, args.vectors, 998, args.sweeps, skip=args.skip)
        held += collect(columns, draw(rng), 2, args.vectors, 999, args.sweeps, skip=args.skip)
    theta0 = init_params(args.seed, args.rotations).astype(np.float32)

    def pcg_score(theta):  # checkpoint selection: two-level PCG iterations on the validation systems
        return float(sum(s.pcg_iterations(theta) for s in valid))

    theta, history = adam(train, theta0, args.steps, args.lr, budget_s=args.budget_s,
                          select=pcg_score if args.select else None)
    if args.es_steps > 0:
        theta = es_head(train, theta, args.es_steps, args.es_k, seed=args.seed, budget_s=args.es_budget_s,
                        select=pcg_score if args.select else None)
    report = {}
    for name, group in (("train", train), ("validation", valid), ("held_out", held)):
        l0 = mean_gradient(group, theta0)[0]
        l1 = mean_gradient(group, theta)[0]
        it0 = [s.pcg_iterations(None) for s in group]
        it1 = [s.pcg_iterations(theta) for s in group]
        report[name] = {"loss_bilinear": l0, "loss_trained": l1, "pcg_bilinear": it0, "pcg_trained": it1}
        if args.rotations:
            # the start (bilinear weights, levers of the shape) against the translations-only bilinear space
            itt = [s.translations_iterations() for s in group]
            report[name]["pcg_bilinear_translations"] = itt
        print(f"[train] {name}: loss {l0:.5f} -> {l1:.5f}; PCG iterations (1e-6) start {sum(it0)} -> trained "
              f"{sum(it1)}" + (f"; bilinear translations only {sum(itt)}" if args.rotations else ""), flush=True)
    meta = {"features": F_IN, "hidden": HIDDEN, "layers": LAYERS,
            "loss": "projection" + (f", then ES on the PCG residual after {args.es_k} iterations" if args.es_steps else ""),
            "trainer": "device",
            "rotations": bool(args.rotations), "gradient_steps": len(history), "history": history[::10],
            "report": report, "command": "python -m drape_amd.learned.train_gpu " + " ".join(sys.argv[1:]),
            "params": [float(x) for x in theta]}
    with open(args.out, "w") as f:
        json.dump(meta, f)
    for s in train + valid + held:
        s.close()


if __name__ == "__main__":
    main()
