"""drape_amd: the implicit Newton step of drape's cloth solver on MI355X (gfx950).

A thin ctypes layer over the C ABI in include/drape_hip.h (libdrape_hip.so, built in-tree by
drape_amd/build.py). It mirrors the reference runner's entry points for the step:

    solve_once          <- solveOnce          (engine/gpu/src/solver.ts:179)
    solve_step_robust   <- solveStepRobust    (engine/gpu/src/robust-newton.ts:231)
    solve_equilibrium   <- solveEquilibrium   (engine/gpu/src/drape-e7.ts:156)

The HIP library is the only compute path: importing this module on a machine without the built
library raises, and no call falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdrape_hip.so")

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)


class DrapeError(RuntimeError):
    pass


class Problem(C.Structure):
    _fields_ = [("vertices", C.c_int32), ("triangles", C.c_int32), ("hinges", C.c_int32), ("rows", C.c_int32),
                ("blocks", C.c_int32), ("rest_positions", f32p), ("elem_vertices", u32p), ("elem_geom", f32p),
                ("hinge_vertices", u32p), ("hinge_stiffness", f32p), ("row_vertex", u32p), ("row_unit_mass", f32p),
                ("row_ptr", u32p), ("col_idx", u32p), ("colliders", f32p), ("collider_floats", C.c_int32),
                ("wale", C.c_double), ("course", C.c_double), ("coupling", C.c_double), ("shear", C.c_double),
                ("gravity", C.c_double), ("time_step", C.c_double), ("grid_columns", C.c_int32)]


class Rules(C.Structure):
    _fields_ = [("newton_tolerance", C.c_double), ("newton_max_iterations", C.c_int32), ("armijo", C.c_double),
                ("max_halvings", C.c_int32), ("pcg_tolerance", C.c_double), ("pcg_max_iterations", C.c_int32),
                ("replace_every", C.c_int32), ("replace_above", C.c_double)] + \
               [(k, C.c_double) for k in ("eta_start", "eta_min", "eta_max", "eta_gamma", "eta_safeguard",
                                          "update_stop", "resolution_factor", "mu_start", "mu_cap", "h0", "h_max",
                                          "h_min_factor", "growth", "shrink", "inner_tolerance")] + \
               [("inner_max_iterations", C.c_int32), ("max_rejections", C.c_int32), ("equilibrium_f32", C.c_double),
                ("stagnation_window", C.c_int32), ("stagnation_factor", C.c_double), ("gravity_load", C.c_double),
                ("floor_stop", C.c_int32), ("forcing", C.c_int32), ("regularize", C.c_int32),
                ("noise_aware", C.c_int32), ("exact_tangent", C.c_int32), ("refine_fp64", C.c_int32),
                ("preconditioner", C.c_int32), ("coarse_reuse", C.c_int32),
                ("fallback", C.c_int32), ("coarse_shift", C.c_double), ("smoother", C.c_double),
                ("coarse_half", C.c_int32), ("refresh_rate", C.c_double), ("refresh_on_breakdown", C.c_int32),
                ("self_contact", C.c_int32), ("coarse_async", C.c_int32), ("block_jacobi_truncate", C.c_int32),
                ("adaptive_kappa", C.c_int32), ("kappa_distance", C.c_double), ("kappa_max_factor", C.c_double),
                ("coarse_factor", C.c_int32), ("coarse_split", C.c_int32),
                ("mu_carry", C.c_int32), ("coarse_rotations", C.c_int32), ("coarse_pivot_floor", C.c_double)]


class PcgStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("recursive_relative_residual", C.c_double),
                ("true_relative_residual", C.c_double), ("chunks", C.c_int32), ("refinements", C.c_int32),
                ("fp64_relative_residual", C.c_double)]


RECORD_FIELDS = ["step", "iteration", "relative_residual", "pcg_iterations", "pcg_status", "linear_residual",
                 "step_length", "halvings", "status"]
EXTRA_FIELDS = ["eta", "mu", "update_inf", "step_inf", "retries", "resolution", "alpha_max", "limiting_vertex",
                "min_distance", "delta_sum", "delta_abs", "noise_only", "slope", "slope_exact", "norm_squared",
                "norm_squared_exact"]


class Record(C.Structure):
    _fields_ = [(k, C.c_double) for k in RECORD_FIELDS + EXTRA_FIELDS]


class PtcStep(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("step", "h", "residual", "inner", "rejected", "pcg_iterations",
                                          "min_distance", "curvature_dropped")]


class PtcAttempt(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("step", "attempt", "h", "reason", "record_start", "record_end")]


class PtcResult(C.Structure):
    _fields_ = [("stop", C.c_int32), ("accepted_steps", C.c_int32), ("newton_iterations", C.c_int32),
                ("pcg_iterations", C.c_int64), ("records", C.c_int32), ("attempts", C.c_int32),
                ("final_h", C.c_double), ("final_residual", C.c_double), ("gravity_load", C.c_double),
                ("seconds", C.c_double), ("phase_ms", C.c_double * 7), ("fallbacks", C.c_int64),
                ("read_bytes", C.c_int64), ("truncations", C.c_int64), ("breakdowns", C.c_int64)]


PHASES = ["assembly", "pcg", "line_search", "ccd", "reduce", "io", "other"]


class Scene(C.Structure):
    _fields_ = [("columns", C.c_int32), ("sheet_size", C.c_double), ("rest_factor", C.c_double),
                ("wale", C.c_double), ("course", C.c_double), ("coupling", C.c_double), ("shear", C.c_double),
                ("bend_wale", C.c_double), ("bend_course", C.c_double), ("density", C.c_double),
                ("gravity", C.c_double), ("time_step", C.c_double), ("pinned", C.c_int32),
                ("n_colliders", C.c_int32), ("colliders", C.c_double * 32), ("dhat", C.c_double),
                ("kappa", C.c_double), ("lift", C.c_double), ("exact_tangent", C.c_int32), ("jitter", C.c_double),
                ("flip_fraction", C.c_double), ("mesh_seed", C.c_uint64)]


STOP_RULES = ["tolerance", "stagnation", "cap", "stalled", "time_budget"]
PCG_STATUS = ["running", "converged", "iteration_limit", "breakdown", "negative_curvature"]

_lib = None


def lib():
    """The HIP library; raises DrapeError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DrapeError(f"{LIB_PATH} is missing: build it with `python drape_amd/build.py` "
                             "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.drape_last_error.restype = C.c_char_p
        L.drape_default_rules.argtypes = [C.POINTER(Rules)]
        L.drape_create.argtypes = [C.POINTER(Problem), C.POINTER(Rules), C.POINTER(vp)]
        L.drape_linear_create.argtypes = [C.c_int32, C.c_int32, u32p, u32p, f32p, f32p, C.POINTER(Rules),
                                          C.POINTER(vp)]
        L.drape_destroy.argtypes = [vp]
        L.drape_set_state.argtypes = [vp, f32p, f32p, C.c_double, C.c_float]
        L.drape_get_positions.argtypes = [vp, f32p]
        L.drape_set_barrier_curvature.argtypes = [vp, C.c_int32]
        L.drape_residual.argtypes = [vp, f32p, f64p]
        L.drape_pcg.argtypes = [vp, C.c_double, C.c_int32, C.POINTER(PcgStats), f32p]
        L.drape_slope.argtypes = [vp, f64p]
        L.drape_set_direction.argtypes = [vp, f32p]
        L.drape_energy_change.argtypes = [vp, C.c_float, f64p, f64p]
        L.drape_max_step.argtypes = [vp, f32p, i32p, f32p]
        L.drape_newton_step.argtypes = [vp, f32p, C.c_float, C.c_int32, f32p, C.POINTER(Record), C.c_int32, i32p]
        L.drape_ptc_solve.argtypes = [vp, f32p, C.c_double, C.c_int32, C.c_double, C.POINTER(PtcResult),
                                      C.POINTER(PtcStep), C.c_int32, C.POINTER(PtcAttempt), C.c_int32,
                                      C.POINTER(Record), C.c_int32, f32p]
        L.drape_set_colliders.argtypes = [vp, C.c_int32]
        L.drape_set_row_mass.argtypes = [vp, f32p]
        L.drape_set_tangent.argtypes = [vp, C.c_int32]
        L.drape_device_name.argtypes = [C.c_char_p, C.c_int32]
        L.drape_prolong_state.argtypes = [C.c_int32, f32p, f64p, C.c_int32, f32p, i32p, i32p]
        L.drape_ptc_keep_states.argtypes = [vp, C.c_int32]
        L.drape_ptc_states.argtypes = [vp, f32p, C.c_int32, i32p]
        L.drape_eval_elements.argtypes = [vp, f32p, f32p, f32p]
        L.drape_get_system.argtypes = [vp, f32p, f32p]
        L.drape_pcg_timing.argtypes = [vp, f64p, f64p, i32p]
        L.drape_set_preconditioner.argtypes = [vp, C.c_int32, f32p]
        L.drape_two_level_layout.argtypes = [vp, i32p, i32p, f32p, C.POINTER(C.c_int64)]
        L.drape_two_level_coarse.argtypes = [vp, f64p]
        L.drape_copy_timing.argtypes = [C.c_int64, C.c_int32, f64p]
        L.drape_set_prolongation_model.argtypes = [vp, f32p, C.c_int32]
        L.drape_two_level_weights.argtypes = [vp, f32p]
        L.drape_self_contact_pairs.argtypes = [vp, i32p, i32p, i32p, f64p]
        L.drape_training_vectors_set.argtypes = [vp, f32p, C.c_int32]
        L.drape_training_vectors_make.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, f32p]
        L.drape_prolongation_gradient.argtypes = [vp, f32p, C.c_int32, f64p, f32p]
        L.drape_solution_fp64.argtypes = [vp, f64p]
        L.drape_simulate.argtypes = [vp, C.c_int32, f32p, f32p, f32p, f32p, i32p, f32p, C.POINTER(Record), C.c_int32,
                                     i32p]
        L.drape_two_level_inverse_factor.argtypes = [vp, f32p]
        L.drape_two_level_ordering.argtypes = [vp, i32p]
        L.drape_two_level_rotations.argtypes = [vp, i32p, f32p, C.POINTER(C.c_int64)]
        L.drape_ptc_probe.argtypes = [vp, f32p, C.c_double, C.c_int32, C.POINTER(Record), C.c_int32, i32p, f32p,
                                      f32p]
        L.drape_kernel_timing.argtypes = [vp, C.c_int32, C.c_int32, f64p, f64p]
        L.drape_set_kappa.argtypes = [vp, C.c_double]
        L.drape_kappa.argtypes = [vp, f64p]
        L.drape_kappa_history.argtypes = [vp, f64p, f32p, C.c_int32, i32p]
        L.drape_scene_build.argtypes = [C.POINTER(Scene), C.POINTER(vp)]
        L.drape_scene_build_device.argtypes = [C.POINTER(Scene), C.POINTER(vp)]
        L.drape_scene_contributions.argtypes = [vp, C.POINTER(C.c_int64), u32p, u32p, u32p, u32p, u32p]
        L.drape_get_contributions.argtypes = [vp, C.POINTER(C.c_int64), u32p, u32p, u32p, u32p]
        L.drape_scene_problem.restype = C.POINTER(Problem)
        L.drape_scene_problem.argtypes = [vp, f64p]
        L.drape_scene_free.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise DrapeError(f"drape error {rc}: {lib().drape_last_error().decode()}")


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def copy_timing(nbytes: int, launches: int = 10) -> float:
    """Average ms of a plain device copy of nbytes (the byte-counter calibration kernel)."""
    ms = C.c_double()
    _check(lib().drape_copy_timing(nbytes, launches, C.byref(ms)))
    return ms.value


def device_name() -> str:
    """Name, architecture and compute units of the current device."""
    buf = C.create_string_buffer(256)
    _check(lib().drape_device_name(buf, 256))
    return buf.value.decode()


def scene_colliders(scene: Scene) -> np.ndarray:
    """The scene's colliders packed as [count, dhat, kappa, 0, (kind, a, b, c, w, 0, 0, 0) per collider] in f64."""
    k = scene.n_colliders
    out = np.zeros(4 + 8 * k)
    out[:4] = (k, scene.dhat, scene.kappa, 0.0)
    for c in range(k):
        out[4 + 8 * c:4 + 8 * c + 5] = scene.colliders[8 * c:8 * c + 5]
    return out


def prolong_state(coarse, coarse_columns, colliders=None):
    """A coarse f32 state prolonged to the (2 coarse_columns - 1)^2 grid and projected out of the colliders
    (e7_prolong.rs:9-47): (fine positions, vertices moved to d = dhat, vertices left infeasible)."""
    coarse = _f32(coarse).ravel()
    if coarse.size != 3 * coarse_columns * coarse_columns:
        raise DrapeError(f"coarse state has {coarse.size // 3} vertices, not {coarse_columns}^2")
    nf = 2 * (coarse_columns - 1) + 1
    fine = np.zeros(3 * nf * nf, np.float32)
    cols = None if colliders is None else np.ascontiguousarray(colliders, np.float64)
    moved, bad = C.c_int32(), C.c_int32()
    _check(lib().drape_prolong_state(coarse_columns, _ptr(coarse, f32p), None if cols is None else _ptr(cols, f64p),
                                     0 if cols is None else cols.size, _ptr(fine, f32p), C.byref(moved),
                                     C.byref(bad)))
    return fine, moved.value, bad.value


def default_rules(**overrides) -> Rules:
    r = Rules()
    lib().drape_default_rules(C.byref(r))
    for k, v in overrides.items():
        setattr(r, k, v)
    return r


def isotropic_material(E=1e6, nu=0.3, thickness=5e-4):
    """(wale, course, coupling, shear) of an isotropic plane-stress sheet in the reference's
    orthotropic parameters: W = wale/2 Eww^2 + course/2 Ecc^2 + coupling Eww Ecc + 2 shear Ewc^2."""
    membrane = E * thickness / (1.0 - nu * nu)
    return membrane, membrane, nu * membrane, E * thickness / (2.0 * (1.0 + nu))


def make_scene(columns, *, sheet_size=0.5, material=(1000.0, 300.0, 100.0, 50.0), bend=(2e-6, 1e-6), density=0.13,
               gravity=9.81, time_step=1.0 / 240.0, pinned=True, colliders=(), dhat=0.0, kappa=0.0, lift=0.0,
               rest_factor=1.0, jitter=0.0, flip_fraction=0.0, mesh_seed=0) -> Scene:
    s = Scene()
    s.columns = columns
    s.sheet_size = sheet_size
    s.rest_factor = rest_factor
    s.wale, s.course, s.coupling, s.shear = material
    s.bend_wale, s.bend_course = bend
    s.density = density
    s.gravity = gravity
    s.time_step = time_step
    s.pinned = int(pinned)
    s.n_colliders = len(colliders)
    for k, c in enumerate(colliders):
        for j, v in enumerate(c):
            s.colliders[8 * k + j] = v
    s.dhat, s.kappa, s.lift = dhat, kappa, lift
    s.exact_tangent = 1
    s.jitter, s.flip_fraction, s.mesh_seed = jitter, flip_fraction, mesh_seed
    return s


def reference_scene(columns, sphere_on_floor=True, material_factor=1.0, radius_factor=1.0, rest_factor=1.0,
                    placed=True) -> Scene:
    """The reference's Scene::SphereOnFloor / PinnedCorners with SceneParams (drape.rs:53-155)."""
    sheet = 0.5
    mat = tuple(material_factor * v for v in (1000.0, 300.0, 100.0, 50.0))
    bend = (material_factor * 2e-6, material_factor * 1e-6)
    if not sphere_on_floor:
        return make_scene(columns, material=mat, bend=bend, pinned=True, rest_factor=rest_factor)
    activation = 1e-3 * sheet
    radius = 0.25 * sheet * radius_factor
    clearance = 2.0 * activation if placed else 0.02 * sheet
    return make_scene(columns, material=mat, bend=bend, pinned=False, rest_factor=rest_factor,
                      colliders=[(1, 0.0, 0.0, 1.0, 0.0), (2, 0.5 * sheet, 0.5 * sheet, radius, radius)],
                      dhat=activation, kappa=0.0, lift=2.0 * radius + clearance)


def drape_config_scene(columns, *, radius=0.15, clearance=None, dhat=1e-3, time_step=1.0 / 60.0, E=1e6, bend=1e-5,
                       jitter=0.0, flip_fraction=0.0, mesh_seed=0):
    """The quasi-static / dynamic sphere drape of the configured workloads: isotropic E (1e6 Pa), nu = 0.3,
    t = 0.5 mm, areal density 0.2 kg/m^2, hinge bending `bend` (1e-5 N m), a sphere of the given radius
    centred under the 0.5 m sheet (resting on z = 0, no floor), barrier dhat, sheet released flat
    `clearance` above the sphere top (2 dhat when None); jitter / flip_fraction > 0 give an irregular mesh."""
    clearance = 2.0 * dhat if clearance is None else clearance
    return make_scene(columns, material=isotropic_material(E=E), bend=(bend, bend), density=0.2, time_step=time_step,
                      jitter=jitter, flip_fraction=flip_fraction, mesh_seed=mesh_seed,
                      pinned=False, colliders=[(2, 0.25, 0.25, radius, radius)], dhat=dhat, kappa=0.0,
                      lift=2.0 * radius + clearance)


class SceneArrays:
    """Problem arrays built by the library from a scene (problem.bin's arrays)."""

    def __init__(self, scene: Scene, device: bool = False):
        """device: build hinges, pattern, geometry, masses and the contribution lists on the GPU."""
        h = C.c_void_p()
        build = lib().drape_scene_build_device if device else lib().drape_scene_build
        _check(build(C.byref(scene), C.byref(h)))
        self.h = h
        load = C.c_double()
        self.problem_ptr = lib().drape_scene_problem(self.h, C.byref(load))
        self.problem = self.problem_ptr.contents
        self.gravity_load = load.value
        p = self.problem
        self.V, self.T, self.H, self.R, self.B = p.vertices, p.triangles, p.hinges, p.rows, p.blocks

    def array(self, name):
        p = self.problem
        n = {"rest_positions": 3 * self.V, "elem_vertices": 3 * self.T, "elem_geom": 7 * self.T,
             "hinge_vertices": 4 * self.H, "hinge_stiffness": self.H, "row_vertex": self.R,
             "row_unit_mass": self.R, "row_ptr": self.R + 1, "col_idx": self.B,
             "colliders": p.collider_floats}[name]
        ptr = getattr(p, name)
        if n == 0 or not ptr:
            return np.zeros(0, np.float32)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def contributions(self):
        """diag_idx, block_contrib_ptr, block_contrib_idx, force_contrib_ptr, force_contrib_idx of a
        device-built scene (problem.bin's arrays)."""
        sizes = (C.c_int64 * 2)()
        _check(lib().drape_scene_contributions(self.h, sizes, None, None, None, None, None))
        out = [np.zeros(self.R, np.uint32), np.zeros(self.B + 1, np.uint32), np.zeros(sizes[0], np.uint32),
               np.zeros(self.R + 1, np.uint32), np.zeros(sizes[1], np.uint32)]
        _check(lib().drape_scene_contributions(self.h, sizes, *[_ptr(a, u32p) for a in out]))
        return dict(zip(["diag_idx", "block_contrib_ptr", "block_contrib_idx", "force_contrib_ptr",
                         "force_contrib_idx"], out))

    def __del__(self):
        if getattr(self, "h", None):
            lib().drape_scene_free(self.h)
            self.h = None


class Drape:
    """A device context for one problem (replaces prepareDrapeGpu, drape-gpu.ts:141-162)."""

    def __init__(self, arrays: SceneArrays, rules: Rules | None = None):
        self.arrays = arrays
        self.rules = rules if rules is not None else default_rules()
        if self.rules.gravity_load <= 0:
            self.rules.gravity_load = arrays.gravity_load
        h = C.c_void_p()
        _check(lib().drape_create(arrays.problem_ptr, C.byref(self.rules), C.byref(h)))
        self.h = h
        self.V, self.T, self.H, self.R, self.B = arrays.V, arrays.T, arrays.H, arrays.R, arrays.B

    def close(self):
        if getattr(self, "h", None):
            lib().drape_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- single device operations (NewtonOps, newton-ops.ts:32-36) ----
    def set_state(self, positions=None, x_hat=None, h=0.0, mu=0.0):
        p = None if positions is None else _ptr(_f32(positions), f32p)
        xh = None if x_hat is None else _ptr(_f32(x_hat), f32p)
        _check(lib().drape_set_state(self.h, p, xh, h, mu))

    def positions(self):
        out = np.zeros(3 * self.V, np.float32)
        _check(lib().drape_get_positions(self.h, _ptr(out, f32p)))
        return out

    def contributions(self):
        """The context's device-built scatter maps: block_contrib_ptr/idx and force_contrib_ptr/idx."""
        sizes = (C.c_int64 * 2)()
        _check(lib().drape_get_contributions(self.h, sizes, None, None, None, None))
        out = [np.zeros(self.B + 1, np.uint32), np.zeros(sizes[0], np.uint32), np.zeros(self.R + 1, np.uint32),
               np.zeros(sizes[1], np.uint32)]
        _check(lib().drape_get_contributions(self.h, sizes, *[_ptr(a, u32p) for a in out]))
        return dict(zip(["block_contrib_ptr", "block_contrib_idx", "force_contrib_ptr", "force_contrib_idx"], out))

    def set_colliders(self, enabled: bool):
        """withoutColliders (contact-host.ts:64-73): False zeroes the collider count the kernels read."""
        _check(lib().drape_set_colliders(self.h, 1 if enabled else 0))

    def set_row_mass(self, row_mass):
        _check(lib().drape_set_row_mass(self.h, _ptr(_f32(row_mass), f32p)))

    def set_tangent(self, exact: bool):
        _check(lib().drape_set_tangent(self.h, 1 if exact else 0))

    def set_barrier_curvature(self, drop: bool):
        _check(lib().drape_set_barrier_curvature(self.h, int(drop)))

    def residual(self):
        n = C.c_float()
        n2 = C.c_double()
        _check(lib().drape_residual(self.h, C.byref(n), C.byref(n2)))
        return n.value, n2.value

    def solve_once(self, tolerance, max_iterations=20000):
        st = PcgStats()
        x = np.zeros(3 * self.R, np.float32)
        _check(lib().drape_pcg(self.h, tolerance, max_iterations, C.byref(st), _ptr(x, f32p)))
        return st, x

    def slope(self):
        s = C.c_double()
        _check(lib().drape_slope(self.h, C.byref(s)))
        return s.value

    def set_direction(self, rows=None):
        _check(lib().drape_set_direction(self.h, None if rows is None else _ptr(_f32(rows), f32p)))

    def energy_change(self, alpha):
        s, a = C.c_double(), C.c_double()
        _check(lib().drape_energy_change(self.h, alpha, C.byref(s), C.byref(a)))
        return s.value, a.value

    def max_step(self):
        a, v, d = C.c_float(), C.c_int32(), C.c_float()
        _check(lib().drape_max_step(self.h, C.byref(a), C.byref(v), C.byref(d)))
        return a.value, v.value, d.value

    def eval_elements(self, positions=None):
        g = np.zeros(9 * self.T + 12 * self.H + 3 * self.R, np.float32)
        hs = np.zeros(81 * self.T + 144 * self.H + 9 * self.R, np.float32)
        p = None if positions is None else _ptr(_f32(positions), f32p)
        _check(lib().drape_eval_elements(self.h, p, _ptr(g, f32p), _ptr(hs, f32p)))
        return g, hs

    def system(self):
        rhs = np.zeros(3 * self.R, np.float32)
        vals = np.zeros(9 * self.B, np.float32)
        _check(lib().drape_get_system(self.h, _ptr(rhs, f32p), _ptr(vals, f32p)))
        return rhs, vals

    def spmv_timing(self, launches=200):
        spmv, solve = C.c_double(), C.c_double()
        n = C.c_int32(launches)
        _check(lib().drape_pcg_timing(self.h, C.byref(spmv), C.byref(solve), C.byref(n)))
        return spmv.value, solve.value

    KERNELS = {"k_pcg_spmv": 0, "k_pcg_dir_spmv": 1, "k_tri_gemv_lower": 2, "k_tri_gemv_upper": 3, "k_gram": 4,
               "coarse_factorisation": 5}

    def kernel_timing(self, kernel, launches=200):
        """(average ms per launch, algorithmic bytes per launch) of one PCG kernel (drape_kernel_timing);
        for "coarse_factorisation", (ms per factorisation, FP64 flops of its Cholesky factorisation and inverse)."""
        ms, nbytes = C.c_double(), C.c_double()
        _check(lib().drape_kernel_timing(self.h, self.KERNELS[kernel], launches, C.byref(ms), C.byref(nbytes)))
        return ms.value, nbytes.value

    def set_preconditioner(self, kind, weights=None):
        """'block_jacobi' or 'two_level' (optionally with parent-slot weights [4R])."""
        k = {"block_jacobi": 0, "two_level": 1}[kind] if isinstance(kind, str) else int(kind)
        w = None if weights is None else _f32(weights)
        if w is not None and w.size != 4 * self.R:
            raise ValueError("weights must have 4 entries per row")
        _check(lib().drape_set_preconditioner(self.h, k, None if w is None else _ptr(w, f32p)))
        self.rules.preconditioner = k

    def two_level_layout(self):
        """(coarse rows, parent [R, 4], bilinear weights [R, 4], counters (factorisations, failures,
        fallbacks, microseconds factorising))."""
        nc = C.c_int32()
        parent = np.zeros(4 * self.R, np.int32)
        w = np.zeros(4 * self.R, np.float32)
        cnt = (C.c_int64 * 4)()
        _check(lib().drape_two_level_layout(self.h, C.byref(nc), _ptr(parent, i32p), _ptr(w, f32p), cnt))
        return nc.value, parent.reshape(-1, 4), w.reshape(-1, 4), tuple(cnt)

    def two_level_rotations(self):
        """(unknowns per coarse node, rotation levers [R, 4, 3] of the current factorisation, pivots replaced by
        the pivot floor so far)."""
        dof = C.c_int32()
        rep = C.c_int64()
        lv = np.zeros(16 * self.R, np.float32)
        _check(lib().drape_two_level_rotations(self.h, C.byref(dof), _ptr(lv, f32p), C.byref(rep)))
        return dof.value, lv.reshape(self.R, 4, 4)[:, :, :3].copy(), rep.value

    def coarse_unknowns(self):
        """Unknowns of the coarse space (unknowns per node x coarse nodes)."""
        dof = C.c_int32()
        _check(lib().drape_two_level_rotations(self.h, C.byref(dof), None, None))
        return dof.value * self.two_level_layout()[0]

    def two_level_coarse(self):
        """The Galerkin coarse operator P^T A P of the assembled system (dense, f64)."""
        n = self.coarse_unknowns()
        out = np.zeros((n, n))
        _check(lib().drape_two_level_coarse(self.h, _ptr(out, f64p)))
        return out

    def solution_fp64(self):
        """The FP64 iterate of the last refined solve (rules.refine_fp64)."""
        out = np.zeros(3 * self.R)
        _check(lib().drape_solution_fp64(self.h, _ptr(out, f64p)))
        return out

    def set_prolongation_model(self, params=None):
        """The learned prolongation (flat float32 parameters; None: bilinear weights)."""
        if params is None:
            _check(lib().drape_set_prolongation_model(self.h, None, 0))
            return
        theta = _f32(params)
        _check(lib().drape_set_prolongation_model(self.h, _ptr(theta, f32p), theta.size))

    def self_contact_pairs(self):
        """Self-contact pairs of the last evaluation: (point-triangle count, vertices [n, 4] int32, FP64
        distances [n]); the point-triangle pairs come first."""
        npt, nee = np.zeros(1, np.int32), np.zeros(1, np.int32)
        _check(lib().drape_self_contact_pairs(self.h, _ptr(npt, i32p), _ptr(nee, i32p), None, None))
        n = int(npt[0] + nee[0])
        verts = np.zeros((max(n, 1), 4), np.int32)
        dist = np.zeros(max(n, 1), np.float64)
        _check(lib().drape_self_contact_pairs(self.h, None, None, _ptr(verts, i32p), _ptr(dist, f64p)))
        return int(npt[0]), verts[:n], dist[:n]

    def training_vectors(self, vectors=None, count=16, sweeps=5, omega=0.6, seed=0):
        """The error vectors of the prolongation loss, kept on the device: `vectors` [count, 3R] when given,
        else made on the device (random starts, damped block-Jacobi sweeps on the assembled system) and
        returned as float32 [count, 3R]."""
        if vectors is not None:
            v = np.ascontiguousarray(np.asarray(vectors, np.float32).reshape(-1, 3 * self.R))
            _check(lib().drape_training_vectors_set(self.h, _ptr(v, f32p), v.shape[0]))
            return v
        out = np.zeros((count, 3 * self.R), np.float32)
        _check(lib().drape_training_vectors_make(self.h, count, sweeps, omega, seed, _ptr(out, f32p)))
        return out

    def prolongation_gradient(self, params):
        """(loss, dloss/dparams) of the learned prolongation at `params` over the training vectors, on the
        device; `params` becomes the context's prolongation model."""
        theta = _f32(params)
        grad = np.zeros(theta.size, np.float32)
        loss = C.c_double()
        _check(lib().drape_prolongation_gradient(self.h, _ptr(theta, f32p), theta.size, C.byref(loss),
                                                 _ptr(grad, f32p)))
        return loss.value, grad

    def two_level_weights(self):
        """Parent-slot weights [R, 4] of the current coarse factorisation."""
        out = np.zeros(4 * self.R, np.float32)
        _check(lib().drape_two_level_weights(self.h, _ptr(out, f32p)))
        return out.reshape(-1, 4)

    def two_level_inverse_factor(self):
        """B = L^-1 of the current coarse factorisation (FP32, dense lower triangular) in the factor's ordering
        (two_level_ordering)."""
        n = self.coarse_unknowns()
        out = np.zeros((n, n), np.float32)
        _check(lib().drape_two_level_inverse_factor(self.h, _ptr(out, f32p)))
        return out

    def two_level_ordering(self):
        """The natural coarse unknown dof I + q at each position of the factor's ordering [dof Nc]."""
        out = np.zeros(self.coarse_unknowns(), np.int32)
        _check(lib().drape_two_level_ordering(self.h, _ptr(out, i32p)))
        return out

    # ---- the step and its caller ----
    def solve_step_robust(self, x, scale, step, origin, cap=1024):
        x = _f32(x).copy()
        origin = _f32(origin)
        recs = (Record * cap)()
        n = C.c_int32()
        _check(lib().drape_newton_step(self.h, _ptr(x, f32p), scale, step, _ptr(origin, f32p), recs, cap,
                                       C.byref(n)))
        return x, [recs[i] for i in range(min(n.value, cap))]

    def simulate(self, steps, x0=None, v0=None, keep=20000):
        """Backward-Euler trajectory (drape.rs simulate): final positions and velocities, per-step Newton
        updates and final relative residuals, records."""
        x = np.zeros(3 * self.V, np.float32)
        v = np.zeros(3 * self.V, np.float32)
        up = np.zeros(max(steps, 1), np.int32)
        rel = np.zeros(max(steps, 1), np.float32)
        recs = (Record * keep)()
        n = C.c_int32()
        _check(lib().drape_simulate(self.h, steps, None if x0 is None else _ptr(_f32(x0), f32p),
                                    None if v0 is None else _ptr(_f32(v0), f32p), _ptr(x, f32p), _ptr(v, f32p),
                                    _ptr(up, i32p), _ptr(rel, f32p), recs, keep, C.byref(n)))
        return {"positions": x, "velocities": v, "updates": up[:steps], "relative": rel[:steps],
                "records": [recs[i] for i in range(min(n.value, keep))]}

    def ptc_probe(self, start, h, step=1, cap=64, iterates=False):
        """One PTC attempt from `start` at step size h (runPtcProbes, drape-e7.ts:183-196): the last iterate,
        the records and, with `iterates`, the positions at each record [n, 3V]."""
        recs = (Record * cap)()
        n = C.c_int32()
        final = np.zeros(3 * self.V, np.float32)
        its = np.zeros((cap, 3 * self.V), np.float32) if iterates else None
        _check(lib().drape_ptc_probe(self.h, _ptr(_f32(start), f32p), h, step, recs, cap, C.byref(n),
                                     None if its is None else _ptr(its, f32p), _ptr(final, f32p)))
        k = min(n.value, cap)
        return {"positions": final, "records": [recs[i] for i in range(k)],
                "iterates": None if its is None else its[:k].copy()}

    @property
    def kappa(self):
        out = C.c_double()
        _check(lib().drape_kappa(self.h, C.byref(out)))
        return out.value

    def set_kappa(self, kappa):
        _check(lib().drape_set_kappa(self.h, kappa))

    def kappa_history(self):
        """(kappa of each step of the last simulate, smallest collider distance after it)."""
        n = C.c_int32()
        _check(lib().drape_kappa_history(self.h, None, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        dist = np.zeros(n.value, np.float32)
        _check(lib().drape_kappa_history(self.h, _ptr(out, f64p), _ptr(dist, f32p), n.value, C.byref(n)))
        return out, dist

    def solve_equilibrium(self, start=None, h_start=None, cap=400, budget_s=0.0, keep=20000, keep_states=False):
        """solveEquilibrium (drape-e7.ts:156-181). "states" holds the start and every accepted state when
        keep_states, else the start and the last."""
        _check(lib().drape_ptc_keep_states(self.h, 1 if keep_states else 0))
        res = PtcResult()
        steps = (PtcStep * keep)()
        attempts = (PtcAttempt * keep)()
        recs = (Record * keep)()
        final = np.zeros(3 * self.V, np.float32)
        sp = None if start is None else _ptr(_f32(start), f32p)
        h0 = self.rules.h0 if h_start is None else h_start
        _check(lib().drape_ptc_solve(self.h, sp, h0, cap, budget_s, C.byref(res), steps, keep, attempts, keep,
                                     recs, keep, _ptr(final, f32p)))
        return {
            "result": res,
            "stop": STOP_RULES[res.stop] if 0 <= res.stop < len(STOP_RULES) else str(res.stop),
            "steps": [steps[i] for i in range(min(res.accepted_steps, keep))],
            "attempts": [attempts[i] for i in range(min(res.attempts, keep))],
            "records": [recs[i] for i in range(min(res.records, keep))],
            "positions": final,
            "states": self.kept_states(),
        }

    def kept_states(self):
        n = C.c_int32()
        _check(lib().drape_ptc_states(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 3 * self.V), np.float32)
        _check(lib().drape_ptc_states(self.h, _ptr(out, f32p), n.value, C.byref(n)))
        return out


class LinearSystem:
    """A context over a plain BSR system (the E1 cases): solve_once mirrors solveOnce (solver.ts:179)."""

    def __init__(self, row_ptr, col_idx, vals, rhs, rules: Rules | None = None):
        self.rules = rules if rules is not None else default_rules()
        row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        col_idx = np.ascontiguousarray(col_idx, np.uint32)
        self.R, self.B = len(row_ptr) - 1, len(col_idx)
        h = C.c_void_p()
        _check(lib().drape_linear_create(self.R, self.B, _ptr(row_ptr, u32p), _ptr(col_idx, u32p),
                                         _ptr(_f32(vals), f32p), _ptr(_f32(rhs), f32p), C.byref(self.rules),
                                         C.byref(h)))
        self.h = h

    def solve_once(self, tolerance, max_iterations=20000):
        st = PcgStats()
        x = np.zeros(3 * self.R, np.float32)
        _check(lib().drape_pcg(self.h, tolerance, max_iterations, C.byref(st), _ptr(x, f32p)))
        return st, x

    def spmv_timing(self, launches=200):
        spmv, solve = C.c_double(), C.c_double()
        n = C.c_int32(launches)
        _check(lib().drape_pcg_timing(self.h, C.byref(spmv), C.byref(solve), C.byref(n)))
        return spmv.value, solve.value

    def close(self):
        if getattr(self, "h", None):
            lib().drape_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
