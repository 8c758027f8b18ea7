"""ctypes binding of the oracle library (oracle/liboracle.so) for the test-suite and the CPU
baseline leg of bench.py. The oracle is the checker: nothing in drape_amd/ imports this module."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class Scene(C.Structure):
    _fields_ = [
        ("columns", C.c_int32),
        ("sheet_size", C.c_double), ("rest_factor", C.c_double),
        ("wale", C.c_double), ("course", C.c_double), ("coupling", C.c_double), ("shear", C.c_double),
        ("bend_wale", C.c_double), ("bend_course", C.c_double), ("density", C.c_double),
        ("gravity", C.c_double), ("time_step", C.c_double),
        ("pinned", C.c_int32), ("n_colliders", C.c_int32),
        ("colliders", C.c_double * 32),
        ("dhat", C.c_double), ("kappa", C.c_double), ("lift", C.c_double),
        ("exact_tangent", C.c_int32),
        ("jitter", C.c_double), ("flip_fraction", C.c_double),
        ("mesh_seed", C.c_uint64),
    ]


class PtcStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("newton_iterations", C.c_int32), ("rejected_attempts", C.c_int32),
                ("reason", C.c_int32), ("final_relative_residual", C.c_double), ("final_h", C.c_double),
                ("static_energy", C.c_double), ("max_mu", C.c_double), ("seconds", C.c_double)]


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("status", C.c_int32),
                ("recursive_relative_residual", C.c_double), ("true_relative_residual", C.c_double),
                ("seconds", C.c_double)]


class DerivativeCheck(C.Structure):
    _fields_ = [("gradient_error", C.c_double), ("hessian_error", C.c_double),
                ("hinge_energy_change_error", C.c_double), ("energy_change_error", C.c_double)]


def build() -> None:
    """Compile the oracle with its own Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_reference_scene.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                         C.POINTER(Scene)]
        L.or_drape_create.restype = C.c_void_p
        L.or_drape_create.argtypes = [C.POINTER(Scene)]
        L.or_drape_destroy.argtypes = [C.c_void_p]
        L.or_drape_sizes.argtypes = [C.c_void_p, i64p]
        L.or_drape_set_time_step.argtypes = [C.c_void_p, C.c_double]
        L.or_drape_kappa.restype = C.c_double
        L.or_drape_kappa.argtypes = [C.c_void_p]
        L.or_drape_gravity_load.restype = C.c_double
        L.or_drape_gravity_load.argtypes = [C.c_void_p]
        L.or_drape_rest_positions.argtypes = [C.c_void_p, f64p]
        L.or_problem_arrays.argtypes = [C.c_void_p] + [C.c_void_p] * 18
        L.or_membrane_elements.argtypes = [C.c_void_p, f64p, C.c_int32, f64p, f64p]
        L.or_hinge_elements.argtypes = [C.c_void_p, f64p, C.c_int32, f64p, f64p]
        L.or_row_terms.argtypes = [C.c_void_p, f64p, f64p, C.c_double, C.c_int32, f64p, f64p]
        L.or_assemble.argtypes = [C.c_void_p, f64p, f64p, C.c_double, C.c_int32, f64p, f64p]
        L.or_energy_change.restype = C.c_double
        L.or_energy_change.argtypes = [C.c_void_p, f64p, f64p, f64p, C.c_double]
        L.or_max_step.restype = C.c_double
        L.or_max_step.argtypes = [C.c_void_p, f64p, f64p, i32p]
        L.or_static_residual.restype = C.c_double
        L.or_static_residual.argtypes = [C.c_void_p, f64p]
        L.or_static_energy.restype = C.c_double
        L.or_static_energy.argtypes = [C.c_void_p, f64p]
        L.or_ptc_solve.argtypes = [C.c_void_p, f64p, C.c_double, C.c_int32, C.c_char_p, C.c_int64, f64p,
                                   C.POINTER(PtcStats)]
        L.or_simulate.restype = C.c_int32
        L.or_simulate.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_double, i32p,
                                  f64p, C.c_int64, i64p, f64p]
        L.or_drape_set_kappa.argtypes = [C.c_void_p, C.c_double]
        L.or_ptc_steps.restype = C.c_int32
        L.or_ptc_steps.argtypes = [C.c_void_p, f64p, C.c_double, C.c_int32, f64p, C.c_int32, f64p,
                                   C.POINTER(PtcStats)]
        L.or_static_floor.restype = C.c_double
        L.or_static_floor.argtypes = [C.c_void_p, f64p]
        L.or_simulate_states.restype = C.c_int32
        L.or_simulate_states.argtypes = [C.c_void_p, C.c_int32, f64p, f64p, C.c_double, C.c_double, C.c_int32,
                                         C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double, i32p, f64p,
                                         f64p, i32p, C.c_int32, f64p, f64p]
        L.or_ptc_inner_solve.restype = C.c_int32
        L.or_ptc_inner_solve.argtypes = [C.c_void_p, f64p, C.c_double, f64p, i32p]
        L.or_time_newton_iteration.restype = C.c_double
        L.or_time_newton_iteration.argtypes = [C.c_void_p, f64p, C.c_double, i32p]
        L.or_drape_derivative_check.argtypes = [C.c_int32, C.POINTER(DerivativeCheck)]
        L.or_case_create.restype = C.c_void_p
        L.or_case_create.argtypes = [C.c_int32, C.c_int32]
        L.or_case_destroy.argtypes = [C.c_void_p]
        L.or_case_sizes.argtypes = [C.c_void_p, i64p, i64p]
        L.or_case_arrays.argtypes = [C.c_void_p, i32p, i32p, f64p, f64p]
        L.or_pcg.argtypes = [C.c_int64, C.c_int64, i32p, i32p, f64p, f64p, C.c_double, C.c_int32, C.c_int32,
                             C.c_double, C.c_int32, C.c_int32, C.c_int32, f64p, f64p, C.c_int64, i64p,
                             C.POINTER(PcgStats)]
        L.or_band_solve.restype = C.c_int32
        L.or_band_solve.argtypes = [C.c_int64, C.c_int64, i32p, i32p, f64p, f64p, f64p, i32p, f64p]
        _lib = L
    return _lib


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def reference_scene(columns, sphere=False, material_factor=1.0, radius_factor=1.0, rest_factor=1.0, placed=False):
    s = Scene()
    lib().or_reference_scene(columns, int(sphere), material_factor, radius_factor, rest_factor, int(placed), C.byref(s))
    return s


class Drape:
    """An oracle drape (drape.rs:88-106) built from a Scene."""

    def __init__(self, scene: Scene):
        self.scene = scene
        self.h = lib().or_drape_create(C.byref(scene))
        sizes = np.zeros(7, np.int64)
        lib().or_drape_sizes(self.h, ptr(sizes, i64p))
        self.V, self.T, self.H, self.R, self.B, self.block_contribs, self.force_contribs = (int(v) for v in sizes)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_drape_destroy(self.h)
            self.h = None

    def set_time_step(self, h):
        lib().or_drape_set_time_step(self.h, h)

    @property
    def kappa(self):
        return lib().or_drape_kappa(self.h)

    def set_kappa(self, kappa):
        lib().or_drape_set_kappa(self.h, kappa)

    @property
    def gravity_load(self):
        return lib().or_drape_gravity_load(self.h)

    def rest_positions(self):
        x = np.zeros(3 * self.V)
        lib().or_drape_rest_positions(self.h, ptr(x, f64p))
        return x

    def problem_arrays(self) -> dict:
        V, T, H, R, B = self.V, self.T, self.H, self.R, self.B
        ncoll = self.scene.n_colliders
        a = {
            "rest_positions": np.zeros(3 * V, np.float32), "elem_vertices": np.zeros(3 * T, np.uint32),
            "elem_geom": np.zeros(7 * T, np.float32), "hinge_vertices": np.zeros(4 * H, np.uint32),
            "hinge_stiffness": np.zeros(H, np.float32), "row_vertex": np.zeros(R, np.uint32),
            "row_mass": np.zeros(R, np.float32), "row_unit_mass": np.zeros(R, np.float32),
            "row_ptr": np.zeros(R + 1, np.uint32), "col_idx": np.zeros(B, np.uint32),
            "diag_idx": np.zeros(R, np.uint32), "block_contrib_ptr": np.zeros(B + 1, np.uint32),
            "block_contrib_idx": np.zeros(self.block_contribs, np.uint32),
            "force_contrib_ptr": np.zeros(R + 1, np.uint32),
            "force_contrib_idx": np.zeros(self.force_contribs, np.uint32),
            "kernel_state": np.zeros(3 * V, np.float32),
            "colliders": np.zeros(4 + 8 * ncoll if ncoll else 0, np.float32),
            "contact_state": np.zeros(3 * V if ncoll else 0, np.float32),
        }
        order = ["rest_positions", "elem_vertices", "elem_geom", "hinge_vertices", "hinge_stiffness", "row_vertex",
                 "row_mass", "row_unit_mass", "row_ptr", "col_idx", "diag_idx", "block_contrib_ptr",
                 "block_contrib_idx", "force_contrib_ptr", "force_contrib_idx", "kernel_state", "colliders",
                 "contact_state"]
        args = [a[k].ctypes.data_as(C.c_void_p) if a[k].size else None for k in order]
        lib().or_problem_arrays(self.h, *args)
        return a

    def membrane_elements(self, x, exact=True):
        x = np.ascontiguousarray(x, np.float64)
        g = np.zeros(9 * self.T)
        hs = np.zeros(81 * self.T)
        lib().or_membrane_elements(self.h, ptr(x, f64p), int(exact), ptr(g, f64p), ptr(hs, f64p))
        return g, hs

    def hinge_elements(self, x, exact=True):
        x = np.ascontiguousarray(x, np.float64)
        g = np.zeros(12 * self.H)
        hs = np.zeros(144 * self.H)
        lib().or_hinge_elements(self.h, ptr(x, f64p), int(exact), ptr(g, f64p), ptr(hs, f64p))
        return g, hs

    def row_terms(self, x, x_hat, mu=0.0, curvature=True):
        x = np.ascontiguousarray(x, np.float64)
        x_hat = np.ascontiguousarray(x_hat, np.float64)
        g = np.zeros(3 * self.R)
        hs = np.zeros(9 * self.R)
        lib().or_row_terms(self.h, ptr(x, f64p), ptr(x_hat, f64p), mu, int(curvature), ptr(g, f64p), ptr(hs, f64p))
        return g, hs

    def assemble(self, x, x_hat, mu=0.0, curvature=True):
        x = np.ascontiguousarray(x, np.float64)
        x_hat = np.ascontiguousarray(x_hat, np.float64)
        rhs = np.zeros(3 * self.R)
        vals = np.zeros(9 * self.B)
        lib().or_assemble(self.h, ptr(x, f64p), ptr(x_hat, f64p), mu, int(curvature), ptr(rhs, f64p), ptr(vals, f64p))
        return rhs, vals

    def energy_change(self, x, x_hat, d, alpha):
        x, x_hat, d = (np.ascontiguousarray(v, np.float64) for v in (x, x_hat, d))
        return lib().or_energy_change(self.h, ptr(x, f64p), ptr(x_hat, f64p), ptr(d, f64p), alpha)

    def max_step(self, x, d):
        x, d = (np.ascontiguousarray(v, np.float64) for v in (x, d))
        vertex = C.c_int32(-1)
        a = lib().or_max_step(self.h, ptr(x, f64p), ptr(d, f64p), C.byref(vertex))
        return a, vertex.value

    def static_residual(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return lib().or_static_residual(self.h, ptr(x, f64p))

    def static_energy(self, x):
        x = np.ascontiguousarray(x, np.float64)
        return lib().or_static_energy(self.h, ptr(x, f64p))

    def ptc_solve(self, start=None, h_start=1.0 / 240.0, max_steps=0):
        log = C.create_string_buffer(1 << 20)
        final = np.zeros(3 * self.V)
        stats = PtcStats()
        sp = None if start is None else ptr(np.ascontiguousarray(start, np.float64), f64p)
        lib().or_ptc_solve(self.h, sp, h_start, max_steps, log, len(log), ptr(final, f64p), C.byref(stats))
        return log.value.decode(), final, stats

    def ptc_trajectory(self, h_start=1.0 / 240.0, max_steps=1000, stop_below=0.0):
        """The oracle's PTC solve (e7_ptc.rs:163-189) one accepted step at a time: the states [n+1, 3V], the step
        rows [n, 5] = (h, r/L at the step's start, inner iterations, rejections, max mu) and the stop reason
        (0 tolerance, 1 floor, 2 cap, 3 stalled, 5 r/L below stop_below)."""
        states, rows = [self.rest_positions()], []
        h, reason = h_start, 2
        while len(rows) < max_steps:
            row = np.zeros(5)
            nxt = np.zeros(3 * self.V)
            st = PtcStats()
            n = lib().or_ptc_steps(self.h, ptr(np.ascontiguousarray(states[-1]), f64p), h, 1, ptr(row, f64p), 1,
                                   ptr(nxt, f64p), C.byref(st))
            if n == 0:
                reason = st.reason
                break
            states.append(nxt)
            rows.append(row)
            h = st.final_h
            if st.final_relative_residual <= stop_below:
                reason = 5
                break
        return np.array(states), np.array(rows).reshape(-1, 5), reason

    def step_residual(self, x, x_hat):
        """|residual| of the step system at x with prediction x_hat (the current time step)."""
        return float(np.linalg.norm(self.assemble(x, x_hat)[0]))

    def inner_solve(self, x, h):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(3 * self.V)
        it = C.c_int32(0)
        status = lib().or_ptc_inner_solve(self.h, ptr(x, f64p), h, ptr(out, f64p), C.byref(it))
        return status, it.value, out

    def time_newton_iteration(self, x, h):
        x = np.ascontiguousarray(x, np.float64)
        ok = C.c_int32(0)
        secs = lib().or_time_newton_iteration(self.h, ptr(x, f64p), h, C.byref(ok))
        return secs, bool(ok.value)

    def simulate(self, steps, tolerance, max_iterations, regularized, floor_factor):
        iters = np.zeros(steps, np.int32)
        hist = np.zeros(steps * (max_iterations + 2))
        offsets = np.zeros(steps + 1, np.int64)
        final = np.zeros(3 * self.V)
        lib().or_simulate(self.h, steps, tolerance, max_iterations, int(regularized), floor_factor, ptr(iters, i32p),
                          ptr(hist, f64p), hist.size, ptr(offsets, i64p), ptr(final, f64p))
        histories = [hist[offsets[n]:offsets[n + 1]].copy() for n in range(steps)]
        return iters, histories, final


    def simulate_states(self, steps, keep=(), x0=None, v0=None, kappa=0.0, tolerance=1e-10, max_iterations=200,
                        regularized=True, floor_factor=2.0, adaptive=False, kappa_distance=0.1,
                        kappa_max_factor=100.0):
        """Backward-Euler trajectory from (x0, v0) (rest, zero when None) keeping the states in `keep`:
        per-step Newton iterations, barrier stiffness and smallest collider distance, and the kept positions
        and velocities [len(keep), 3V]; `adaptive` applies the IPC stiffness update after each step."""
        keep = np.ascontiguousarray(keep, np.int32)
        it = np.zeros(steps, np.int32)
        kap = np.zeros(steps)
        dmin = np.zeros(steps)
        pos = np.zeros((max(len(keep), 1), 3 * self.V))
        vel = np.zeros((max(len(keep), 1), 3 * self.V))
        xp = None if x0 is None else ptr(np.ascontiguousarray(x0, np.float64), f64p)
        vp = None if v0 is None else ptr(np.ascontiguousarray(v0, np.float64), f64p)
        lib().or_simulate_states(self.h, steps, xp, vp, kappa, tolerance, max_iterations, int(regularized),
                                 floor_factor, int(adaptive), kappa_distance, kappa_max_factor, ptr(it, i32p),
                                 ptr(kap, f64p), ptr(dmin, f64p), ptr(keep, i32p) if len(keep) else None,
                                 len(keep), ptr(pos, f64p), ptr(vel, f64p))
        return it, kap, dmin, pos[:len(keep)], vel[:len(keep)]


def derivative_check(columns):
    out = DerivativeCheck()
    lib().or_drape_derivative_check(columns, C.byref(out))
    return out


class Case:
    """E1 linear system (cases.rs:64-161): kind 0 tension, 1 slack, 2 dynamic."""

    KINDS = {"tension": 0, "slack": 1, "dynamic": 2}

    def __init__(self, kind: str, columns: int):
        h = lib().or_case_create(self.KINDS[kind], columns)
        rows, blocks = C.c_int64(), C.c_int64()
        lib().or_case_sizes(h, C.byref(rows), C.byref(blocks))
        self.rows, self.blocks = rows.value, blocks.value
        self.row_ptr = np.zeros(self.rows + 1, np.int32)
        self.col_idx = np.zeros(self.blocks, np.int32)
        self.vals = np.zeros(9 * self.blocks)
        self.rhs = np.zeros(3 * self.rows)
        lib().or_case_arrays(h, ptr(self.row_ptr, i32p), ptr(self.col_idx, i32p), ptr(self.vals, f64p),
                             ptr(self.rhs, f64p))
        lib().or_case_destroy(h)


def pcg(row_ptr, col_idx, vals, rhs, tolerance, max_iterations=20000, replace_every=0, replace_above=0.0,
        single=False, breakdown_check=False, history_stride=16):
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    rows, blocks = len(row_ptr) - 1, len(col_idx)
    x = np.zeros(3 * rows)
    hist = np.zeros(2 * (max_iterations // max(history_stride, 1) + 4))
    hlen = C.c_int64(0)
    stats = PcgStats()
    lib().or_pcg(rows, blocks, ptr(row_ptr, i32p), ptr(col_idx, i32p), ptr(vals, f64p), ptr(rhs, f64p), tolerance,
                 max_iterations, replace_every, replace_above, int(single), int(breakdown_check), history_stride,
                 ptr(x, f64p), ptr(hist, f64p), hist.size, C.byref(hlen), C.byref(stats))
    history = [(int(hist[2 * i]), float(hist[2 * i + 1])) for i in range(hlen.value)]
    return x, stats, history


def band_solve(row_ptr, col_idx, vals, rhs):
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    rows, blocks = len(row_ptr) - 1, len(col_idx)
    x = np.zeros(3 * rows)
    bw = C.c_int32(0)
    secs = C.c_double(0)
    ok = lib().or_band_solve(rows, blocks, ptr(row_ptr, i32p), ptr(col_idx, i32p), ptr(vals, f64p), ptr(rhs, f64p),
                             ptr(x, f64p), C.byref(bw), C.byref(secs))
    return (x if ok == 0 else None), bw.value, secs.value
