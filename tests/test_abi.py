"""CPU checks of the C ABI library: it loads, exports every entry point include/drape_hip.h declares, its
scene builder writes the same problem arrays as the oracle (and the reference's manifests), and malformed
inputs are rejected with DRAPE_E_INVALID before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import drape_amd as D
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "drape_hip.h")).read()
    return sorted(set(re.findall(r"\b(drape_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = D.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_rules_are_the_reference_constants():
    r = D.default_rules()
    assert (r.newton_tolerance, r.newton_max_iterations, r.armijo, r.max_halvings) == (1e-6, 30, 1e-4, 30)
    assert (r.pcg_tolerance, r.pcg_max_iterations, r.replace_every, r.replace_above) == (1e-7, 20000, 50, 1e-2)
    assert (r.eta_start, r.eta_min, r.eta_max, r.eta_gamma, r.eta_safeguard) == (1e-2, 1e-7, 1e-2, 0.9, 0.1)
    assert (r.update_stop, r.resolution_factor, r.mu_start, r.mu_cap) == (1e-4, 2.0, 1e-3, 1e3)
    assert (r.h0, r.h_max, r.h_min_factor, r.growth, r.shrink) == (1.0 / 240.0, 1.0, 0.1, 2.0, 0.5)
    assert (r.inner_tolerance, r.inner_max_iterations, r.max_rejections) == (0.1, 10, 12)
    assert (r.equilibrium_f32, r.stagnation_window, r.stagnation_factor) == (1e-3, 10, 0.95)
    assert (r.floor_stop, r.forcing, r.regularize, r.noise_aware) == (1, 1, 1, 1)


ARRAYS = ["rest_positions", "elem_vertices", "elem_geom", "hinge_vertices", "hinge_stiffness", "row_vertex",
          "row_unit_mass", "row_ptr", "col_idx", "colliders"]


@pytest.mark.parametrize("n,sphere", [(9, False), (17, True), (33, False), (33, True), (65, True)])
def test_scene_builder_matches_oracle_arrays_bit_for_bit(n, sphere):
    a = D.SceneArrays(D.reference_scene(n, sphere_on_floor=sphere, placed=True))
    od = O.Drape(O.reference_scene(n, sphere=sphere, placed=True))
    oa = od.problem_arrays()
    for name in ARRAYS:
        x, y = a.array(name), oa[name]
        assert x.shape == y.shape, name
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    assert a.gravity_load == od.gravity_load


def test_irregular_scene_matches_oracle():
    """The jittered, partly flipped mesh (no reference counterpart) agrees between library and oracle."""
    s = D.make_scene(17, pinned=False, jitter=0.2, flip_fraction=0.1, mesh_seed=11)
    a = D.SceneArrays(s)
    osc = O.reference_scene(17)
    osc.pinned = 0
    osc.jitter, osc.flip_fraction, osc.mesh_seed = 0.2, 0.1, 11
    od = O.Drape(osc)
    oa = od.problem_arrays()
    for name in ARRAYS[:-1]:
        x, y = a.array(name), oa[name]
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name


def test_config_material_is_the_isotropic_mapping():
    wale, course, coupling, shear = D.isotropic_material(1e6, 0.3, 5e-4)
    assert wale == course == pytest.approx(549.4505494505494)
    assert coupling == pytest.approx(164.83516483516485)
    assert shear == pytest.approx(192.30769230769232)


def test_malformed_problems_are_rejected_before_device_work():
    a = D.SceneArrays(D.reference_scene(9, sphere_on_floor=False))
    p = D.Problem.from_buffer_copy(a.problem)
    h = C.c_void_p()
    r = D.default_rules()
    bad = D.Problem.from_buffer_copy(p)
    bad.rows = 0
    assert D.lib().drape_create(C.byref(bad), C.byref(r), C.byref(h)) == -1
    bad = D.Problem.from_buffer_copy(p)
    bad.row_ptr = None
    assert D.lib().drape_create(C.byref(bad), C.byref(r), C.byref(h)) == -1
    assert b"missing" in D.lib().drape_last_error()
    # a triangle that references a vertex beyond V
    tri = np.ctypeslib.as_array(p.elem_vertices, shape=(3 * p.triangles,)).copy()
    tri[5] = p.vertices + 7
    bad = D.Problem.from_buffer_copy(p)
    bad.elem_vertices = tri.ctypes.data_as(D.u32p)
    assert D.lib().drape_create(C.byref(bad), C.byref(r), C.byref(h)) == -1
    assert b"out of range" in D.lib().drape_last_error()
    # unsorted columns in a linear system
    rp = np.array([0, 2, 3], np.uint32)
    ci = np.array([1, 0, 1], np.uint32)
    vals = np.ones(27, np.float32)
    rhs = np.ones(6, np.float32)
    assert D.lib().drape_linear_create(2, 3, rp.ctypes.data_as(D.u32p), ci.ctypes.data_as(D.u32p),
                                       vals.ctypes.data_as(D.f32p), rhs.ctypes.data_as(D.f32p), C.byref(r),
                                       C.byref(h)) in (-1, -2)


def test_product_never_imports_the_oracle():
    """The shipped package has no path to the oracle (the checker) or to any CPU fallback."""
    pkg = os.path.join(ROOT, "drape_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".hpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle_lib" not in text and "liboracle" not in text, f
