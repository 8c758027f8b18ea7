"""GPU tests of the learned prolongation's inference (drape_amd/csrc/kernels/gnn.hip) through the C ABI: the
parent-slot weights the device computes from its own assembled system against the f64 numpy network
(oracle/gnn.py, pinned by finite differences) on the same FP32 matrix -- 1e-4 of the largest weight for
random parameters (three FP32 layers against f64), 1e-6 at a zero output layer (the bilinear weights)."""
import numpy as np
import pytest

import drape_amd as D
from oracle import gnn as G

pytestmark = pytest.mark.gpu


def assembled(columns, scene):
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays)
    rest = arrays.array("rest_positions")
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    ctx.set_state(rest, x_hat, 0.05, 0.0)
    ctx.residual()
    return arrays, ctx


@pytest.mark.parametrize("columns", [33, 129])
@pytest.mark.parametrize("output", ["zero", "random"])
def test_device_weights_match_f64_network(columns, output):
    scene = D.reference_scene(33, sphere_on_floor=False, placed=True) if columns == 33 else D.drape_config_scene(129)
    arrays, ctx = assembled(columns, scene)
    _, vals = ctx.system()
    g = G.Graph(columns, arrays.array("row_vertex"), arrays.array("row_ptr"), arrays.array("col_idx"),
                vals.astype(np.float64))
    theta = G.init_params(seed=21)
    if output == "random":
        theta = theta + 0.2 * np.random.default_rng(22).standard_normal(theta.size)
    theta32 = theta.astype(np.float32)
    ctx.set_preconditioner("two_level")
    ctx.set_prolongation_model(theta32)
    st, _ = ctx.solve_once(1e-2, 500)
    w_dev = ctx.two_level_weights().astype(np.float64)
    w_ref = G.forward(theta32.astype(np.float64), g)
    err = np.abs(w_dev - w_ref).max()
    print(f"[gpu] learned weights n{columns} ({output} output layer): max |w_dev - w_f64| = {err:.2e}, "
          f"PCG {st.iterations} iterations", flush=True)
    assert err <= (1e-6 if output == "zero" else 1e-4)
    assert np.all(w_dev[~g.valid] == 0.0)
    ctx.close()
