"""GPU tests of the dynamic loop (drape_simulate: backward Euler, one robust Newton solve per step, as the
reference's simulate, drape.rs:426-439) on the two dynamic workloads, against the oracle's f64 trajectories
(tests/golden/dynamic_trajectories.npz, written by make_dynamic_fixtures.py): the final positions agree to
1e-3 of the trajectory's displacement for the pinned sheet and 1e-2 through the sphere contact (FP32 states
against f64 at Newton tolerance 1e-10), and with the sphere no vertex ends inside the collider.

The pinned sheet runs as configured: Newton to a relative gradient norm of 1e-8, each linear solve FP32 PCG
inside the FP64 true-residual refinement loop (rules.refine_fp64). FP32 positions cannot represent a residual
that small, so a step ends by the reference's update rule (robust-newton.ts:153-154: the update below 1e-4 of
the step or twice the f32 spacing) when the FP32 floor is reached first."""
import os

import numpy as np
import pytest

import drape_amd as D

pytestmark = pytest.mark.gpu
FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dynamic_trajectories.npz"))


@pytest.mark.parametrize("name,steps,preconditioner", [("c1", 60, 0), ("c2", 30, 1)])
def test_dynamic_trajectory_matches_f64(name, steps, preconditioner):
    if name == "c1":
        scene = D.make_scene(33, material=D.isotropic_material(), bend=(1e-5, 1e-5), density=0.2,
                             time_step=1.0 / 60.0, pinned=True)
    else:
        scene = D.drape_config_scene(65, clearance=0.05)
    arrays = D.SceneArrays(scene)
    rules = dict(newton_tolerance=1e-8, refine_fp64=1) if name == "c1" else {}
    ctx = D.Drape(arrays, D.default_rules(preconditioner=preconditioner, **rules))
    print(f"[gpu] dynamic {name}: {steps} steps", flush=True)
    out = ctx.simulate(steps)
    x = out["positions"].astype(np.float64)
    want, rest = FIX[f"{name}_positions"], FIX[f"{name}_rest"]
    rel = np.linalg.norm(x - want) / np.linalg.norm(want - rest)
    print(f"[gpu]   Newton updates GPU {int(out['updates'].sum())} / oracle {int(FIX[name + '_iterations'].sum())}; "
          f"final |x - x_f64| / |x_f64 - x_0| = {rel:.2e}", flush=True)
    assert np.all(np.isfinite(x))
    assert rel <= (1e-3 if name == "c1" else 1e-2)
    if name == "c2":
        c = np.array([0.25, 0.25, 0.15])
        d = np.linalg.norm(x.reshape(-1, 3) - c, axis=1) - 0.15
        assert d.min() > 0.0
    ctx.close()
