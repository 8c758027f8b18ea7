"""GPU tests of the FP64 iterative refinement around the FP32 PCG (rules.refine_fp64; no reference
counterpart: the reference's only relative is FP32 residual replacement, pcg.rs:82-86). On a system whose
FP32 PCG stalls at its round-off floor, the refined solution's residual, computed on the host in f64 with the
same FP32 matrix, reaches the requested tolerance for the FP64 iterate; the statistics report that FP64
residual, and the FP32 solution is that iterate rounded (its own residual sits at the FP32 floor)."""
import numpy as np
import pytest

import drape_amd as D
from oracle import two_level as TL

pytestmark = pytest.mark.gpu


def pinned(n, h, rules):
    ps = D.reference_scene(n, sphere_on_floor=False, placed=True)
    arrays = D.SceneArrays(ps)
    ctx = D.Drape(arrays, rules)
    rest = arrays.array("rest_positions")
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    ctx.set_state(rest, x_hat, h, 0.0)
    ctx.residual()
    return arrays, ctx


@pytest.mark.parametrize("preconditioner,h", [(0, 0.1), (1, 0.1)])
def test_refinement_reaches_tolerances_below_the_fp32_floor(preconditioner, h):
    tol = 1e-8
    arrays, plain = pinned(65, h, D.default_rules(preconditioner=preconditioner))
    rhs, vals = plain.system()
    A = TL.bsr_matrix(arrays.array("row_ptr"), arrays.array("col_idx"), vals.astype(np.float64))
    b = rhs.astype(np.float64)
    st0, x0 = plain.solve_once(tol, 20000)
    r0 = np.linalg.norm(b - A @ x0.astype(np.float64)) / np.linalg.norm(b)
    _, ref = pinned(65, h, D.default_rules(preconditioner=preconditioner, refine_fp64=1))
    st, x = ref.solve_once(tol, 20000)
    x64 = ref.solution_fp64()
    r = np.linalg.norm(b - A @ x64) / np.linalg.norm(b)
    print(f"[gpu] refinement ({'two-level' if preconditioner else 'block-Jacobi'}): FP32 PCG {st0.iterations} its, "
          f"f64 residual {r0:.2e}; refined {st.refinements} passes, {st.iterations} its, reported "
          f"{st.fp64_relative_residual:.2e}, host f64 {r:.2e}", flush=True)
    assert r0 > 1e3 * tol                     # the FP32 solve alone cannot get there
    assert D.PCG_STATUS[st.status] == "converged" and st.refinements >= 1
    assert st.fp64_relative_residual <= tol
    # the device's FP64 residual is the host's, and the FP32 solution is the FP64 iterate rounded
    assert r <= 2 * tol and r == pytest.approx(st.fp64_relative_residual, rel=1e-3, abs=1e-12)
    assert np.array_equal(x, x64.astype(np.float32))
    plain.close()
    ref.close()
