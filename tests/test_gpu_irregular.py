"""GPU tests on the irregular n257 sheet (interior vertices jittered by up to 0.2 of the spacing, 10% of the
cell diagonals flipped, seeded): the element kernels against f64 at the seed-97 bent state (the reference's
kernel gate, 1e-5 relative L2, e3_compare.rs:14,92-111), the assembled system against f64 (1e-5), the FP32
PCG on that system bit-identical to the oracle's f32 emulation, and the two-level preconditioner on it."""
import numpy as np
import pytest

import drape_amd as D
import oracle_lib as O

pytestmark = pytest.mark.gpu


def rel_l2(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return np.linalg.norm(got - want) / np.linalg.norm(want)


@pytest.fixture(scope="module")
def irregular():
    scene = D.drape_config_scene(257, jitter=0.2, flip_fraction=0.1, mesh_seed=5)
    s = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(scene, name))
    od = O.Drape(s)
    arrays = D.SceneArrays(scene)
    return scene, arrays, od


def test_irregular_element_kernels_and_assembly_match_f64(irregular):
    scene, arrays, od = irregular
    ctx = D.Drape(arrays)
    a = od.problem_arrays()
    state = a["kernel_state"]
    print(f"[gpu] irregular n257: V {arrays.V} T {arrays.T} H {arrays.H} B {arrays.B}", flush=True)
    g, hs = ctx.eval_elements(state)
    x = state.astype(np.float64)
    mg, mh = od.membrane_elements(x, True)
    hg, hh = od.hinge_elements(x, True)
    T, H = od.T, od.H
    errors = [rel_l2(g[:9 * T], mg), rel_l2(g[9 * T:9 * T + 12 * H], hg), rel_l2(hs[:81 * T], mh),
              rel_l2(hs[81 * T:81 * T + 144 * H], hh)]
    print(f"[gpu]   element errors (membrane grad, hinge grad, membrane hess, hinge hess): {errors}", flush=True)
    assert max(errors) <= 1e-5
    rest = arrays.array("rest_positions")
    ctx.set_state(state, rest, 0.1, 0.37)
    od.set_time_step(0.1)
    ctx.residual()
    rhs, vals = ctx.system()
    orhs, ovals = od.assemble(x, rest.astype(np.float64), 0.37, True)
    e_rhs, e_vals = rel_l2(rhs, orhs), rel_l2(vals, ovals)
    print(f"[gpu]   assembled: rhs {e_rhs:.2e} vals {e_vals:.2e}", flush=True)
    assert e_rhs <= 1e-5 and e_vals <= 1e-5
    st, xs = ctx.solve_once(1e-4, 1000)
    xo, so, _ = O.pcg(a["row_ptr"].astype(np.int32), a["col_idx"].astype(np.int32), vals.astype(np.float64),
                      rhs.astype(np.float64), 1e-4, 1000, 50, 1e-2, single=True, breakdown_check=True)
    print(f"[gpu]   PCG {st.iterations} iterations (oracle f32 emulation {so.iterations})", flush=True)
    assert st.iterations == so.iterations
    assert np.array_equal(xs.view(np.uint32), xo.astype(np.float32).view(np.uint32))
    ctx.close()


def test_irregular_two_level_solves_with_fewer_iterations(irregular):
    scene, arrays, od = irregular
    ctx = D.Drape(arrays)
    rest = arrays.array("rest_positions")
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    ctx.set_state(rest, x_hat, 0.05, 0.0)
    ctx.residual()
    st_bj, _ = ctx.solve_once(1e-4, 20000)
    ctx.set_preconditioner("two_level")
    st, _ = ctx.solve_once(1e-4, 20000)
    counters = ctx.two_level_layout()[3]
    print(f"[gpu] irregular n257 two-level: {st.iterations} iterations vs block-Jacobi {st_bj.iterations}, "
          f"counters {counters}", flush=True)
    assert D.PCG_STATUS[st.status] == "converged" and counters[1] == 0
    assert st.iterations * 2 < st_bj.iterations
    ctx.close()
