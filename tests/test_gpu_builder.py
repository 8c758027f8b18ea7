"""GPU tests of the device problem builder (drape_amd/csrc/kernels/builder.hip): the scene arrays built on the
device (hinges, pattern, element geometry, hinge stiffness, masses) equal the host builder's bit for bit, its
diagonal positions and contribution lists equal the oracle's restatement of the reference generator
(e3_generate.rs:111-161) entry for entry, and so do the scatter maps every context now builds on the device at
drape_create. Regular reference scenes (pinned sheet, sphere on floor) and an irregular sheet (jittered, 10%
of the diagonals flipped) are covered; the large-grid case reports the build times."""
import time

import numpy as np
import pytest

import drape_amd as D
import oracle_lib as O

pytestmark = pytest.mark.gpu

ARRAYS = ["rest_positions", "elem_vertices", "elem_geom", "hinge_vertices", "hinge_stiffness", "row_vertex",
          "row_unit_mass", "row_ptr", "col_idx", "colliders"]
MAPS = ["block_contrib_ptr", "block_contrib_idx", "force_contrib_ptr", "force_contrib_idx"]


def scenes():
    irregular = D.make_scene(65, pinned=False, jitter=0.2, flip_fraction=0.1, mesh_seed=11)
    osc = O.reference_scene(65)
    osc.pinned = 0
    osc.jitter, osc.flip_fraction, osc.mesh_seed = 0.2, 0.1, 11
    return {
        "n17_sphere": (D.reference_scene(17, sphere_on_floor=True, placed=True),
                       O.reference_scene(17, sphere=True, placed=True)),
        "n33_pinned": (D.reference_scene(33, sphere_on_floor=False), O.reference_scene(33, sphere=False)),
        "n65_irregular": (irregular, osc),
    }


@pytest.mark.parametrize("name", ["n17_sphere", "n33_pinned", "n65_irregular"])
def test_device_build_matches_host_and_oracle(name):
    ps, osc = scenes()[name]
    print(f"[gpu] device builder {name}", flush=True)
    host = D.SceneArrays(ps)
    dev = D.SceneArrays(ps, device=True)
    assert (dev.V, dev.T, dev.H, dev.R, dev.B) == (host.V, host.T, host.H, host.R, host.B)
    for a in ARRAYS:
        x, y = dev.array(a), host.array(a)
        assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32)), a
    assert dev.gravity_load == host.gravity_load
    oa = O.Drape(osc).problem_arrays()
    maps = dev.contributions()
    for a in ["diag_idx"] + MAPS:
        assert np.array_equal(maps[a], oa[a]), a
    print(f"[gpu]   {dev.H} hinges, {dev.B} blocks, {maps['block_contrib_idx'].size} block and "
          f"{maps['force_contrib_idx'].size} force contributions: equal", flush=True)


@pytest.mark.parametrize("name", ["n17_sphere", "n33_pinned", "n65_irregular"])
def test_context_scatter_maps_match_oracle(name):
    ps, osc = scenes()[name]
    ctx = D.Drape(D.SceneArrays(ps))
    got = ctx.contributions()
    oa = O.Drape(osc).problem_arrays()
    for a in MAPS:
        assert np.array_equal(got[a], oa[a]), a
    ctx.close()


def test_large_grid_build_times():
    """n1025 (1,050,625 vertices): the device build equals the host build on the pattern, and the times of
    both and of the context creation are reported."""
    ps = D.drape_config_scene(1025)
    t0 = time.perf_counter()
    host = D.SceneArrays(ps)
    t1 = time.perf_counter()
    dev = D.SceneArrays(ps, device=True)
    t2 = time.perf_counter()
    ctx = D.Drape(dev)
    t3 = time.perf_counter()
    for a in ("hinge_vertices", "row_ptr", "col_idx", "hinge_stiffness", "row_unit_mass", "elem_geom"):
        x, y = dev.array(a), host.array(a)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), a
    print(f"[gpu]   n1025: host build {t1 - t0:.2f} s, device build {t2 - t1:.2f} s, context creation "
          f"{t3 - t2:.2f} s", flush=True)
    ctx.close()
