"""Writes tests/golden/c5_probes.npz: the oracle's first PTC steps of the irregular n257 sphere drape
(drape_config_scene(257, jitter=0.2, flip_fraction=0.1, mesh_seed=5): interior vertices jittered by up to 0.2
of the spacing, 10% of the cells split along the other diagonal; isotropic E = 1e6 Pa, sphere r = 0.15 m,
dhat = 1 mm), for per-step probes of the GPU at that size.

The f64 oracle needs about a minute per Newton iteration at n257 (band Cholesky of 198,147 unknowns), so the
probes are the first accepted PTC steps from the flat start and their f32 start states are chained from the
fixture itself: start 1 is the f32 rest state, start s+1 = f32(start s + displacement s), where displacement s is
the oracle's accepted state s (from its f64 trajectory) minus start s, stored in f32. The chained start differs
from the f32 rounding of the oracle's state by at most an ulp; its SHA-256 is stored so that the test rebuilds
exactly these starts. The displacements are stored with their bytes shuffled (byte k of every float together),
which lets zlib compress them to about half.

    python tests/golden/make_c5_fixtures.py [--steps 3]
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import drape_amd as D  # noqa: E402  (scene description only)
import oracle_lib as O  # noqa: E402


def c5_oracle():
    ps = D.drape_config_scene(257, jitter=0.2, flip_fraction=0.1, mesh_seed=5)
    s = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(ps, name))
    return O.Drape(s)


def shuffle(a):
    """f32 array -> [4, n] uint8, byte k of every float in row k."""
    return np.ascontiguousarray(np.asarray(a, np.float32).view(np.uint8).reshape(-1, 4).T)


def unshuffle(b):
    return np.ascontiguousarray(np.asarray(b, np.uint8).T).view(np.float32).ravel()


def chained_starts(rest, shuffled_displacements):
    """The f32 start of every probe step: start 1 = f32(rest), start s+1 = f32(start s + displacement s)."""
    starts = [np.asarray(rest, np.float32)]
    for d in shuffled_displacements[:-1]:
        starts.append((starts[-1].astype(np.float64) + unshuffle(d).astype(np.float64)).astype(np.float32))
    return starts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args()
    t0 = time.time()
    od = c5_oracle()
    states, rows, reason = od.ptc_trajectory(max_steps=args.steps)
    print(f"oracle: {len(rows)} steps, inner {rows[:, 2].astype(int).tolist()}, r/L {rows[:, 1].round(4).tolist()} "
          f"({time.time() - t0:.0f} s)", flush=True)
    n = len(rows)
    sha, disp = [], []
    start = states[0].astype(np.float32)
    for s in range(1, n + 1):
        sha.append(hashlib.sha256(start.tobytes()).hexdigest())
        d = (states[s] - start.astype(np.float64)).astype(np.float32)
        disp.append(shuffle(d))
        print(f"probe step {s}: |start - x_f64| / |step| = "
              f"{np.linalg.norm(start - states[s - 1]) / np.linalg.norm(states[s] - states[s - 1]):.2e}", flush=True)
        start = (start.astype(np.float64) + d.astype(np.float64)).astype(np.float32)
    out = {"probe_steps": np.arange(1, n + 1, dtype=np.int32), "probe_h": rows[:, 0].astype(np.float32),
           "probe_inner": rows[:, 2].astype(np.int32), "relative": rows[:, 1], "start_sha256": np.array(sha),
           "probe_disp_trajectory": np.array(disp)}
    assert [hashlib.sha256(x.tobytes()).hexdigest() for x in chained_starts(states[0], disp)] == sha
    path = os.path.join(HERE, "c5_probes.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
