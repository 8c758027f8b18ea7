"""Writes tests/golden/ptc_probes_n<N>.npz: the oracle's f64 pseudo-transient-continuation solve of the
quasi-static sphere drape (drape_config_scene(N): isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm, 0.2 kg/m^2,
hinge 1e-5 N m, sphere r = 0.15 m, dhat = 1 mm), its equilibrium numbers and the start states of its probe
steps, each with the oracle's own step from that start.

The probe steps follow the reference's rule (e7_run.rs:81-91 probe_steps): step 1, the first step whose end
state has a free vertex within dhat of the sphere (e5_checks.rs:24-47), and the first steps whose end state has
r/L <= 1e-1 and <= 1e-2 (r the static gradient norm, L the gravity load). For each probe step s the fixture
holds the f32 start state x_{s-1} and the f32 step size h_{s-1} (the reference's probes.bin, e7_run.rs:93-103),
the oracle's next state x_s of the trajectory, and the oracle's one PTC attempt from the f32-rounded start at
the f32 step size (the step the GPU repeats), both as f32 displacements from the f32 start (the one-step gate
is relative to the displacement, so 2^-24 of it is far below the 1e-3 gate), plus the step's representation
floor (the f64 step residual of the f32-rounded oracle step over the step's scale, e7_compare.rs:180-190).

The sheet rests on a frictionless sphere with no floor, so its symmetric equilibrium is unstable: the f64 solve
comes close to it and then slides off (n33: r/L 8.9e-6 at step 20, free fall from step ~70; n129: r/L 4.2e-5
at step 18, 0.44 after step 19). The solve therefore stops at the first state with r/L <= --stop-below
(1e-5), or once it has left the equilibrium (r/L above 1000 times its least value after that fell below
1e-3), and the equilibrium numbers are those of the state of least r/L.

The oracle runs one accepted step per call (its step-size rule depends only on the last step), so the solve
checkpoints to --work after every step and resumes from there.

    python tests/golden/make_probe_fixtures.py --columns 129 [--work /tmp/drape_probes_n129]
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402

f64p = C.POINTER(C.c_double)


def oracle_scene(columns):
    import drape_amd as D

    ps = D.drape_config_scene(columns)
    s = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(ps, name))
    return s


def bind():
    L = O.lib()
    L.or_ptc_steps.restype = C.c_int32
    L.or_ptc_steps.argtypes = [C.c_void_p, f64p, C.c_double, C.c_int32, f64p, C.c_int32, f64p,
                               C.POINTER(O.PtcStats)]
    L.or_static_floor.restype = C.c_double
    L.or_static_floor.argtypes = [C.c_void_p, f64p]
    return L


def one_step(L, od, x, h):
    """One accepted oracle PTC step from x at step size h (rejections halve it): (next, row, stats)."""
    row = np.zeros(5)
    nxt = np.zeros(3 * od.V)
    st = O.PtcStats()
    n = L.or_ptc_steps(od.h, O.ptr(np.ascontiguousarray(x), f64p), h, 1, O.ptr(row, f64p), 1, O.ptr(nxt, f64p),
                       C.byref(st))
    return n, nxt, row, st


def sphere_contacts(x, scene, row_vertex):
    c = np.array(scene.colliders[1:4])
    r = scene.colliders[4]
    p = x.reshape(-1, 3)[row_vertex]
    return int(np.count_nonzero(np.linalg.norm(p - c, axis=1) - r < scene.dhat))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--columns", type=int, default=129)
    ap.add_argument("--work", default=None)
    ap.add_argument("--max-steps", type=int, default=200)
    ap.add_argument("--stop-below", type=float, default=1e-5)
    args = ap.parse_args()
    n = args.columns
    work = args.work or f"/tmp/drape_probes_n{n}"
    os.makedirs(work, exist_ok=True)
    ckpt = os.path.join(work, "trajectory.npz")
    L = bind()
    scene = oracle_scene(n)
    od = O.Drape(scene)
    load = od.gravity_load
    row_vertex = od.problem_arrays()["row_vertex"].astype(np.int64)
    h0 = 1.0 / 240.0
    if os.path.exists(ckpt):
        z = np.load(ckpt)
        states, rows = list(z["states"]), [r for r in z["rows"]]
        h, done, reason = float(z["h_next"]), bool(z["done"]), int(z["reason"])
        print(f"resuming at {len(rows)} steps", flush=True)
    else:
        states, rows, h, done, reason = [od.rest_positions()], [], h0, False, -1
    def left(rows):
        least = min(r[1] for r in rows)
        return least <= 1e-3 and rows[-1][1] > 1e3 * least

    if rows and not done and left(rows):
        done, reason = True, 6
    t0 = time.time()
    while not done and len(rows) < args.max_steps:
        k, nxt, row, st = one_step(L, od, states[-1], h)
        if k == 0:
            done, reason = True, st.reason
        else:
            states.append(nxt)
            rows.append(row)
            h = st.final_h
            print(f"step {len(rows)} h={row[0]:.3e} r/L={row[1]:.3e} inner={int(row[2])} rej={int(row[3])} "
                  f"mu={row[4]:g} next r/L={st.final_relative_residual:.3e} ({time.time() - t0:.0f} s)", flush=True)
            if st.final_relative_residual <= args.stop_below:
                done, reason = True, 5
            elif left(rows + [[0.0, st.final_relative_residual]]):
                done, reason = True, 6
        np.savez(ckpt, states=np.array(states), rows=np.array(rows).reshape(-1, 5), h_next=h, done=done,
                 reason=reason)
    if not done:
        print(f"the oracle did not stop within {args.max_steps} steps", flush=True)
    states = np.array(states)
    rows = np.array(rows).reshape(-1, 5)
    rl = np.array([od.static_residual(x) / load for x in states])
    contacts = [sphere_contacts(x, scene, row_vertex) for x in states]
    first = lambda pred: next((s for s in range(1, len(states)) if pred(s)), None)  # noqa: E731
    picks = [1, first(lambda s: contacts[s] > 0), first(lambda s: rl[s] <= 1e-1), first(lambda s: rl[s] <= 1e-2)]
    steps = sorted({s for s in picks if s is not None})
    print(f"oracle: {len(rows)} steps, {int(rows[:, 2].sum())} Newton, stop reason {reason} (5: r/L below "
          f"--stop-below, 6: left the equilibrium), final r/L {rl[-1]:.3e}; probe steps {steps}", flush=True)
    best = int(np.argmin(rl))
    eq = states[best]
    eq32 = eq.astype(np.float32).astype(np.float64)
    out = {"columns": n, "gravity_load": load, "steps": rows[:, 0].size, "newton": rows[:, 2].sum(),
           "reason": reason, "final_relative": rl[-1], "equilibrium_step": best, "equilibrium_relative": rl[best],
           "representation_floor": od.static_residual(eq32) / load,
           "static_energy": od.static_energy(eq), "rows": rows, "relative": rl,
           "probe_steps": np.array(steps, np.int32)}
    starts, hs, disp_traj, disp_step, floors, inner = [], [], [], [], [], []
    for s in steps:
        start = states[s - 1].astype(np.float32)
        h32 = np.float32(rows[s - 1, 0])
        x = start.astype(np.float64)
        status, its, nxt = od.inner_solve(x, float(h32))
        od.set_time_step(float(h32))
        x_hat = x.copy()
        x_hat[3 * row_vertex + 2] -= float(h32) ** 2 * scene.gravity
        scale = np.linalg.norm(od.assemble(x, x_hat)[0])
        rep = np.linalg.norm(od.assemble(nxt.astype(np.float32).astype(np.float64), x_hat)[0]) / scale
        starts.append(start)
        hs.append(h32)
        disp_traj.append((states[s] - x).astype(np.float32))
        disp_step.append((nxt - x).astype(np.float32))
        floors.append(rep)
        inner.append(its)
        print(f"probe step {s}: h {h32:.3e}, oracle step {its} Newton (status {status}), |x_s - x_step| / |x_step - "
              f"x_start| = {np.linalg.norm(states[s] - nxt) / np.linalg.norm(nxt - x):.2e}, floor {rep:.3e}",
              flush=True)
    out.update(probe_start=np.array(starts), probe_h=np.array(hs, np.float32), probe_disp_trajectory=np.array(disp_traj),
               probe_disp_step=np.array(disp_step), probe_floor=np.array(floors), probe_inner=np.array(inner, np.int32))
    path = os.path.join(HERE, f"ptc_probes_n{n}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)", flush=True)


if __name__ == "__main__":
    main()
