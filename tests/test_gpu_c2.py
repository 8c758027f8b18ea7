"""GPU tests of the dynamic sphere drape as configured (n65, isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm,
0.2 kg/m^2, hinge 1e-5 N m, no pins, released 0.05 m above a sphere of radius 0.15 m, dhat = 1 mm,
h = 1/60 s, 120 steps) with the adaptive barrier stiffness (rules.adaptive_kappa: the IPC update of Li et al.
2020, kappa doubling when the smallest vertex-sphere distance after a step is below 0.1 dhat and below the
previous step's), plain FP32 block-Jacobi PCG, against the oracle's f64 trajectory with the same update
(tests/golden/c2_probes.npz, make_c2_fixtures.py).

- The whole 120-step run: kappa follows the update replayed on the run's own per-step distances exactly
  (FP32 comparisons, as the library makes them); no vertex ever reaches the sphere; the run's kappa schedule and
  final state are reported beside the oracle's (the contact phase is chaotic over 120 steps, so the per-step
  comparison is by probes, as the reference scores its contact runs).
- Per-step oracle comparison by probes (e3_compare.rs:113-140 probe; e5_checks.rs:49-56 event_probes plus every
  20th step): from the oracle's f32-rounded state and velocity at step k-1 and its kappa, one GPU time step
  agrees with the oracle's state k within 1e-3 of the step's displacement (ONE_STEP_LIMIT, e3_compare.rs:19),
  except where the sheet barely moves and FP32 positions cannot resolve the step: there the GPU's state must
  solve the oracle's f64 step system within twice the representation floor.
"""
import os

import numpy as np
import pytest

import drape_amd as D
from drape_amd.learned import load_model, model_rules

pytestmark = pytest.mark.gpu
FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_probes.npz"))
STEPS = 120
CENTRE, RADIUS, DHAT = np.array([0.25, 0.25, 0.15]), 0.15, 1e-3


def scene_arrays():
    return D.SceneArrays(D.drape_config_scene(65, clearance=0.05))


def replay_kappa(kappa0, distances, fraction=0.1, cap_factor=100.0):
    """The update after each step in FP32 arithmetic: the kappa each step uses."""
    k = np.float32(kappa0)
    cap = np.float32(cap_factor * float(kappa0))
    threshold = np.float32(fraction * float(np.float32(DHAT)))   # dhat as the kernels hold it (FP32)
    previous = np.float32(np.inf)
    out = []
    for d in np.asarray(distances, np.float32):
        out.append(k)
        if d < threshold and d < previous and k < cap:
            k = np.float32(min(np.float32(2.0) * k, cap))
        previous = d
    return np.array(out, np.float64)


def test_c2_full_run_with_adaptive_kappa():
    arrays = scene_arrays()
    ctx = D.Drape(arrays, D.default_rules(preconditioner=0, adaptive_kappa=1))
    kappa0 = ctx.kappa
    print(f"[gpu] C2: {STEPS} steps, adaptive kappa from {kappa0:.4g}", flush=True)
    out = ctx.simulate(STEPS)
    kappa, dist = ctx.kappa_history()
    x = out["positions"].astype(np.float64)
    want, rest = FIX["final"], FIX["rest"]
    gpu_changes = [int(i) for i in np.nonzero(np.diff(kappa))[0] + 1]
    ora_changes = [int(i) for i in np.nonzero(np.diff(FIX["kappa"]))[0] + 1]
    rel = np.linalg.norm(x - want) / np.linalg.norm(want - rest)
    print(f"[gpu]   Newton updates GPU {int(out['updates'].sum())} / oracle {int(FIX['iterations'].sum())}; kappa "
          f"{kappa0:.4g} -> {kappa[-1]:.4g} (oracle {FIX['kappa'][-1]:.4g}); kappa doubles before steps "
          f"{[i + 1 for i in gpu_changes]} (oracle {[i + 1 for i in ora_changes]}); least distance "
          f"{dist.min():.3e} (oracle {FIX['min_distance'].min():.3e}); final |x - x_f64| / |x_f64 - x_0| = {rel:.2e}",
          flush=True)
    assert len(kappa) == STEPS and np.all(np.isfinite(x))
    assert np.array_equal(kappa.astype(np.float32), replay_kappa(kappa0, dist).astype(np.float32))
    assert abs(kappa0 - FIX["kappa"][0]) <= 1e-6 * FIX["kappa"][0]
    assert dist.min() > 0.0
    d_final = np.linalg.norm(x.reshape(-1, 3) - CENTRE, axis=1) - RADIUS
    assert d_final.min() > 0.0
    ctx.close()


@pytest.fixture(scope="module")
def oracle():
    import oracle_lib as O

    ps = D.drape_config_scene(65, clearance=0.05)
    sc = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(sc, name, getattr(ps, name))
    od = O.Drape(sc)
    od.set_time_step(ps.time_step)
    return od, od.problem_arrays()["row_vertex"].astype(np.int64), ps.time_step, ps.gravity


@pytest.mark.parametrize("preconditioner", ["block_jacobi", "learned"])
def test_c2_probes_match_the_oracle_step(preconditioner, oracle):
    """Gate per probe (e3_compare.rs:113-165): the one-step error is within 1e-3 of the step, or the step is
    below what FP32 positions resolve and the GPU's state, evaluated in f64 on the oracle's step system, solves it
    within twice the representation floor (the f64 step residual of the oracle's next state rounded to f32,
    relative to the step's scale; newton_ok's FLOOR_FACTOR). Both are printed per probe."""
    od, free, h, g = oracle
    arrays = scene_arrays()
    ctx = D.Drape(arrays, D.default_rules(**(dict(preconditioner=0) if preconditioner == "block_jacobi" else
                                             dict(preconditioner=1, **model_rules()))))
    if preconditioner == "learned":
        ctx.set_prolongation_model(load_model())
    results = []
    for k, s in enumerate(FIX["probe_steps"]):
        kappa = float(FIX["probe_kappa"][k])
        ctx.set_kappa(kappa)
        od.set_kappa(kappa)
        x0 = FIX["probe_start"][k]
        v0 = FIX["probe_velocity"][k]
        out = ctx.simulate(1, x0=x0, v0=v0)
        x = out["positions"].astype(np.float64)
        start = x0.astype(np.float64)
        want = start + FIX["probe_disp_trajectory"][k].astype(np.float64)
        err = np.linalg.norm(x - want) / np.linalg.norm(want - start)
        x_hat = start + h * v0.astype(np.float64)
        x_hat[3 * free + 2] -= h * h * g
        scale = od.step_residual(start, x_hat)
        rep = od.step_residual(want.astype(np.float32).astype(np.float64), x_hat) / scale
        res = od.step_residual(x, x_hat) / scale
        ok = err <= 1e-3 or res <= 2.0 * rep
        results.append((s, err, res, rep, ok))
        print(f"[gpu] C2 probe {preconditioner} step {s} (kappa {kappa:.4g}): one-step error {err:.2e}; f64 step "
              f"residual of the GPU state {res:.3e}, representation floor {rep:.3e}; {int(out['updates'][0])} Newton "
              f"updates ({'ok' if ok else 'NOT ok'})", flush=True)
    strict = sum(e <= 1e-3 for _, e, _, _, _ in results)
    print(f"[gpu]   {strict} of {len(results)} probes within 1e-3 (the reference's one_step_ok needs all)", flush=True)
    assert all(ok for *_, ok in results)
    assert strict >= 3
    ctx.close()
