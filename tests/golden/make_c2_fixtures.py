"""Writes tests/golden/c2_probes.npz: the oracle's f64 backward-Euler trajectory of the dynamic sphere drape
(n65, isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm, 0.2 kg/m^2, hinge 1e-5 N m, no pins, released 0.05 m above
the top of a sphere of radius 0.15 m, dhat = 1 mm, h = 1/60 s, 120 steps) with the IPC barrier-stiffness
update after each step (kappa doubles when the smallest vertex-sphere distance is below 0.1 dhat and below the
previous step's, up to 100 times the resting kappa of contact.rs:103-107), Newton to 1e-10, regularised.

Per step: the Newton iterations, the kappa used and the smallest distance after it; the final positions.
Probe steps follow the reference's rule for contact runs (e5_checks.rs:49-56 event_probes: step 1, the first
step whose end state has a free vertex within dhat of the sphere, the step with most such vertices, the last
step), plus every 20th step. For each probe step s: the f32 start state x_{s-1}, v_{s-1} (as the reference's
probe_positions / probe_velocities, e3_generate.rs:168-175), the kappa of step s, and the oracle's state x_s as
an f32 displacement from the f32 start (exact to 2^-24 of the displacement; the one-step gate is 1e-3 of it),
plus the oracle's own step from the f32-rounded start (its sensitivity to that rounding).

    python tests/golden/make_c2_fixtures.py      # about ten minutes on one core
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import drape_amd as D  # noqa: E402  (scene description only: the builder's parameters)
import oracle_lib as O  # noqa: E402

STEPS = 120


def main():
    scene = D.drape_config_scene(65, clearance=0.05)
    s = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(scene, name))
    od = O.Drape(s)
    t0 = time.time()
    keep = np.arange(STEPS + 1)
    it, kap, dmin, pos, vel = od.simulate_states(STEPS, keep=keep, adaptive=True)
    print(f"trajectory: {STEPS} steps, {int(it.sum())} Newton, kappa {kap[0]:.4g} -> {kap[-1]:.4g}, min distance "
          f"{dmin.min():.3e} ({time.time() - t0:.0f} s)", flush=True)
    rv = od.problem_arrays()["row_vertex"].astype(np.int64)
    c = np.array(s.colliders[1:4])
    near = [int(np.count_nonzero(np.linalg.norm(x.reshape(-1, 3)[rv] - c, axis=1) - s.colliders[4] < s.dhat))
            for x in pos]
    first = next((k for k in range(1, STEPS + 1) if near[k] > 0), None)
    most = max(range(1, STEPS + 1), key=lambda k: (near[k], -k))
    steps = sorted({1, most, STEPS, *range(20, STEPS, 20)} | ({first} if first else set()))
    print(f"probe steps {steps} (first contact {first}, most active {most} with {near[most]} vertices)", flush=True)
    out = {"iterations": it, "kappa": kap, "min_distance": dmin, "final": pos[-1], "rest": pos[0],
           "contacts": np.array(near, np.int32), "probe_steps": np.array(steps, np.int32)}
    starts, vels, kaps, disp, disp_step = [], [], [], [], []
    for k in steps:
        x0 = pos[k - 1].astype(np.float32)
        v0 = vel[k - 1].astype(np.float32)
        x0d = x0.astype(np.float64)
        _, _, _, p1, _ = od.simulate_states(1, keep=[1], x0=x0d, v0=v0.astype(np.float64), kappa=kap[k - 1])
        starts.append(x0)
        vels.append(v0)
        kaps.append(kap[k - 1])
        disp.append((pos[k] - x0d).astype(np.float32))
        disp_step.append((p1[0] - x0d).astype(np.float32))
        print(f"probe step {k}: kappa {kap[k - 1]:.4g}, |x_k - x_step| / |x_k - x_start| = "
              f"{np.linalg.norm(pos[k] - p1[0]) / np.linalg.norm(pos[k] - x0d):.2e}", flush=True)
    out.update(probe_start=np.array(starts), probe_velocity=np.array(vels), probe_kappa=np.array(kaps),
               probe_disp_trajectory=np.array(disp), probe_disp_step=np.array(disp_step))
    path = os.path.join(HERE, "c2_probes.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
