"""The reference's file contract around the drape path, so that a case directory written by the reference's
generator is solved here and the reference's scorer reads the results unchanged.

In: problem.json + problem.bin -- little-endian arrays on 8-byte boundaries described by the "arrays" manifest
(io.rs:5-55, written by e3_generate.rs:84-104 and e7_run.rs:41-50, read by drape.ts:26-63 and problem.ts:69-75).
Out, into <case>/gpu-<mode>/ (run-e7.ts:93-107): gpu-iterations.bin and gpu-robust.bin (f64 tables of the
IterationRecord / RobustRecord fields, outputs.ts:8-33), gpu-ptc.bin (7 f64 per accepted step), gpu-attempts.bin
(6 f64 per attempt), gpu-states.bin (the kept f32 states), gpu-equilibrium.{bin,json} (state-files.ts:12-23) and
gpu-result.json (schema e7-gpu/1); with checks, gpu-kernel.bin, gpu-contact-kernel.bin and gpu-inertia-h.bin
(run-e7.ts:86-91, 140-145).

The files are the reference's; the solve behind them is this library's (drape_ptc_solve and the kernel-check
switches of the C ABI).
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys

import numpy as np

from . import (DrapeError, Drape, PHASES, Problem, default_rules, device_name, f32p, prolong_state, reference_scene,
               scene_colliders, u32p)

DTYPES = {"f32": np.dtype("<f4"), "u32": np.dtype("<u4"), "f64": np.dtype("<f8")}

# drape.ts:25-27: the arrays every drape problem carries
F32_ARRAYS = ["rest_positions", "elem_geom", "hinge_stiffness", "row_mass", "kernel_state", "probe_positions",
              "probe_velocities"]
U32_ARRAYS = ["elem_vertices", "hinge_vertices", "row_vertex", "dof_mask", "row_ptr", "col_idx", "diag_idx",
              "block_contrib_ptr", "block_contrib_idx", "force_contrib_ptr", "force_contrib_idx", "fixed_vertex",
              "probe_steps"]
SIZES = ["grid_columns", "vertices", "triangles", "hinges", "rows", "unknowns", "blocks", "steps", "probes",
         "time_step_s", "gravity_m_per_s2"]
MATERIAL_KEYS = ["wale_N_per_m", "course_N_per_m", "coupling_N_per_m", "shear_N_per_m"]
NEWTON_KEYS = ["tolerance", "max_iterations", "armijo", "max_halvings"]
PCG_KEYS = ["tolerance", "max_iterations", "replace_every", "replace_above"]
ROBUST_KEYS = ["eta_start", "eta_min", "eta_max", "eta_gamma", "eta_safeguard", "update_stop", "mu_start", "mu_cap",
               "kernel_mu"]                                                      # robust-newton.ts:12
PTC_KEYS = ["h0", "h_max", "h_min_factor", "growth", "shrink", "inner_tolerance", "inner_max_iterations",
            "max_rejections", "equilibrium_f32", "stagnation_window", "stagnation_factor", "gravity_load"]  # ptc.ts:4-5

# outputs.ts:8-10 (the names gpu-result.json lists; the scorer reads the robust table by these names)
RECORD_FIELDS = ["step", "iteration", "relativeResidual", "pcgIterations", "pcgStatus", "linearResidual",
                 "stepLength", "halvings", "status"]
EXTRA_FIELDS = ["eta", "mu", "updateInf", "stepInf", "retries", "resolution", "alphaMax", "limitingVertex",
                "minDistance", "deltaSum", "deltaAbs", "noiseOnly", "slope", "slopeExact", "normSquared",
                "normSquaredExact"]
PTC_FIELDS = ["step", "h", "residual", "inner", "rejected", "pcgIterations", "minDistance"]   # run-e7.ts:96
ATTEMPT_FIELDS = ["step", "attempt", "h", "reason", "recordStart", "recordEnd"]                # run-e7.ts:97
# RobustSettings of run-e7.ts:22 (E6 rung D), which drape_ptc_solve implements
SETTINGS = {"floorStop": True, "forcing": True, "regularize": True, "noiseAware": True, "gpuReductions": True,
            "fewReadbacks": True}
MODES = ["cold", "warm", "prolong"]
CHECK_STEPS = [1.0 / 240.0, 0.1, 1.0]     # run-e7.ts:27
KEEP_STATES_COLUMNS = 65                  # run-e7.ts:29
CURVATURE_RETRY = 4                       # drape-e7.ts:11 REJECTION.curvature_retry


def _snake(name):
    return "".join("_" + ch.lower() if ch.isupper() else ch for ch in name)


class BinaryArrays:
    """Arrays concatenated little-endian, each starting on an 8-byte boundary, with a manifest of
    {"name", "dtype", "length" (elements), "offset" (bytes)} (io.rs:5-55). The file ends with the last array
    (no trailing padding)."""

    def __init__(self):
        self.parts, self.size, self.entries = [], 0, []

    def push(self, name, dtype, values):
        data = np.ascontiguousarray(np.asarray(values).ravel(), dtype=DTYPES[dtype]).tobytes()
        pad = (-self.size) % 8
        if pad:
            self.parts.append(b"\0" * pad)
            self.size += pad
        self.entries.append({"name": name, "dtype": dtype, "length": len(data) // DTYPES[dtype].itemsize,
                             "offset": self.size})
        self.parts.append(data)
        self.size += len(data)

    def tobytes(self):
        return b"".join(self.parts)

    def write(self, path):
        with open(path, "wb") as f:
            f.write(self.tobytes())


def read_array(entries, blob, name, dtype):
    """One array of a manifest (problem.ts:69-75 slice), checked against the file's size."""
    entry = next((e for e in entries if e.get("name") == name), None)
    if entry is None or entry.get("dtype") != dtype:
        raise DrapeError(f"problem.bin: array '{name}' ({dtype}) not found")
    n, off = int(entry["length"]), int(entry["offset"])
    if n < 0 or off < 0 or off + n * DTYPES[dtype].itemsize > len(blob):
        raise DrapeError(f"problem.bin: array '{name}' runs past the end of the file")
    return np.frombuffer(blob, DTYPES[dtype], n, off).copy()


def _field(source, key, kind, label):
    value = source.get(key) if isinstance(source, dict) else None
    ok = isinstance(value, (int, float)) and not isinstance(value, bool) if kind is float else isinstance(value, kind)
    if not ok:
        raise DrapeError(f"problem.json: field '{key}' of '{label}' is missing or has the wrong type")
    return value


def _numbers(source, keys, label):
    if not isinstance(source, dict):
        raise DrapeError(f"problem.json: '{label}' is not an object")
    return {k: _field(source, k, float, label) for k in keys}


def parse_manifest(text):
    """parseDrapeManifest (drape.ts:46-60): sizes, material, newton and pcg settings and the arrays."""
    try:
        value = json.loads(text)
    except json.JSONDecodeError as e:
        raise DrapeError(f"problem.json is not JSON: {e}") from None
    if not isinstance(value, dict):
        raise DrapeError("problem.json is not an object")
    out = dict(value)
    out.update(_numbers(value, SIZES, "root"))
    out["case"] = _field(value, "case", str, "root")
    out["tangent"] = _field(value, "tangent", str, "root")
    out["material"] = _numbers(value.get("material"), MATERIAL_KEYS, "material")
    out["newton"] = _numbers(value.get("newton"), NEWTON_KEYS, "newton")
    out["pcg"] = _numbers(value.get("pcg"), PCG_KEYS, "pcg")
    arrays = _field(value, "arrays", list, "root")
    for e in arrays:
        if not isinstance(e, dict):
            raise DrapeError("problem.json: array entry is not an object")
        _field(e, "name", str, "arrays")
        _field(e, "dtype", str, "arrays")
        _field(e, "length", float, "arrays")
        _field(e, "offset", float, "arrays")
    out["arrays"] = arrays
    return out


class ProblemFile:
    """A drape problem read from problem.json + problem.bin (parseDrapeProblem, drape.ts:62-66), usable where
    the library's scene arrays are (Drape(ProblemFile(...)))."""

    def __init__(self, case_dir):
        self.dir = case_dir
        with open(os.path.join(case_dir, "problem.json")) as f:
            self.text = f.read()
        self.manifest = m = parse_manifest(self.text)
        with open(os.path.join(case_dir, "problem.bin"), "rb") as f:
            blob = f.read()
        self.arrays = {}
        for name in F32_ARRAYS:
            self.arrays[name] = read_array(m["arrays"], blob, name, "f32")
        for name in U32_ARRAYS:
            self.arrays[name] = read_array(m["arrays"], blob, name, "u32")
        names = {e["name"] for e in m["arrays"]}
        self.arrays["colliders"] = (read_array(m["arrays"], blob, "colliders", "f32") if "colliders" in names
                                    else np.zeros(4, np.float32))
        self.arrays["contact_state"] = (read_array(m["arrays"], blob, "contact_state", "f32")
                                        if "colliders" in names else None)
        self.unit_mass = read_array(m["arrays"], blob, "row_unit_mass", "f32") if "row_unit_mass" in names else None
        self.V, self.T, self.H = int(m["vertices"]), int(m["triangles"]), int(m["hinges"])
        self.R, self.B = int(m["rows"]), int(m["blocks"])
        self._check_sizes()
        ptc = m.get("ptc")
        self.ptc = _numbers(ptc, PTC_KEYS, "ptc") if isinstance(ptc, dict) else None
        self.gravity_load = self.ptc["gravity_load"] if self.ptc else 0.0
        self._build_struct()

    def _check_sizes(self):
        a, V, T, H, R, B = self.arrays, self.V, self.T, self.H, self.R, self.B
        want = {"rest_positions": 3 * V, "elem_vertices": 3 * T, "elem_geom": 7 * T, "hinge_vertices": 4 * H,
                "hinge_stiffness": H, "row_vertex": R, "dof_mask": R, "row_mass": R, "row_ptr": R + 1, "col_idx": B,
                "diag_idx": R, "block_contrib_ptr": B + 1, "force_contrib_ptr": R + 1, "kernel_state": 3 * V}
        bad = [k for k, n in want.items() if a[k].size != n]
        if self.manifest["unknowns"] != 3 * R:
            bad.append("unknowns")
        if self.unit_mass is not None and self.unit_mass.size != R:
            bad.append("row_unit_mass")
        c = a["colliders"]
        if c.size < 4 or c.size != 4 + 8 * int(c[0]):
            bad.append("colliders")
        if a["row_ptr"].size and (a["row_ptr"][0] != 0 or a["row_ptr"][-1] != B):
            bad.append("row_ptr")
        if bad:
            raise DrapeError(f"problem.bin: array lengths disagree with problem.json ({', '.join(bad)})")

    def _build_struct(self):
        a, m = self.arrays, self.manifest
        # the library takes unit masses (M, not M/h^2); problems before the PTC format carry only M/h^2 at the
        # file's time step, from which M follows
        h = float(m["time_step_s"])
        self.unit = (self.unit_mass if self.unit_mass is not None
                     else (a["row_mass"].astype(np.float64) * h * h).astype(np.float32))
        p = Problem()
        p.vertices, p.triangles, p.hinges, p.rows, p.blocks = self.V, self.T, self.H, self.R, self.B
        for name, t in (("rest_positions", f32p), ("elem_vertices", u32p), ("elem_geom", f32p),
                        ("hinge_vertices", u32p), ("hinge_stiffness", f32p), ("row_vertex", u32p),
                        ("row_ptr", u32p), ("col_idx", u32p), ("colliders", f32p)):
            setattr(p, name, a[name].ctypes.data_as(t))
        p.row_unit_mass = self.unit.ctypes.data_as(f32p)
        p.collider_floats = int(a["colliders"].size)
        mat = m["material"]
        p.wale, p.course = mat["wale_N_per_m"], mat["course_N_per_m"]
        p.coupling, p.shear = mat["coupling_N_per_m"], mat["shear_N_per_m"]
        p.gravity, p.time_step = float(m["gravity_m_per_s2"]), h
        p.grid_columns = int(m["grid_columns"])
        self.problem = p
        self.problem_ptr = C.pointer(p)

    def array(self, name):
        if name == "row_unit_mass":
            return self.unit.copy()
        return self.arrays[name].copy()

    def rules(self, **overrides):
        """The rule constants of problem.json (parseRobustRules, robust-newton.ts:17-23; parsePtcRules,
        ptc.ts:10-14; the newton / pcg sections; the tangent kind) over the library's defaults."""
        m = self.manifest
        r = default_rules()
        n, p = m["newton"], m["pcg"]
        r.newton_tolerance, r.newton_max_iterations = n["tolerance"], int(n["max_iterations"])
        r.armijo, r.max_halvings = n["armijo"], int(n["max_halvings"])
        r.pcg_tolerance, r.pcg_max_iterations = p["tolerance"], int(p["max_iterations"])
        r.replace_every, r.replace_above = int(p["replace_every"]), p["replace_above"]
        robust = m.get("robust")
        if isinstance(robust, dict):
            rb = _numbers(robust, ROBUST_KEYS, "robust")
            for k in ROBUST_KEYS[:-1]:
                setattr(r, k, rb[k])
            factor = robust.get("resolution_factor")
            r.resolution_factor = factor if isinstance(factor, (int, float)) else 0.0
            self.kernel_mu = rb["kernel_mu"]
        else:
            self.kernel_mu = 0.37
        if self.ptc:
            for k in PTC_KEYS:
                v = self.ptc[k]
                setattr(r, k, int(v) if k in ("inner_max_iterations", "max_rejections", "stagnation_window") else v)
        r.exact_tangent = 1 if m["tangent"] == "exact" else 0
        for k, v in overrides.items():
            setattr(r, k, v)
        return r


def write_problem(case_dir, header, arrays):
    """problem.bin from `arrays` [(name, dtype, values)] in order and problem.json from `header` (a dict, written
    in its order) with the manifest appended last, as e3_generate.rs:184-229 and e7_run.rs:41-50 do."""
    os.makedirs(case_dir, exist_ok=True)
    ba = BinaryArrays()
    for name, dtype, values in arrays:
        ba.push(name, dtype, values)
    ba.write(os.path.join(case_dir, "problem.bin"))
    doc = dict(header)
    doc.pop("arrays", None)
    doc["arrays"] = ba.entries
    with open(os.path.join(case_dir, "problem.json"), "w") as f:
        json.dump(doc, f)
    return ba.entries


def write_state(prefix, positions, final_h):
    """<prefix>.bin (f32 x y z per vertex) + <prefix>.json (state-files.ts:12-15)."""
    x = np.ascontiguousarray(positions, "<f4").ravel()
    with open(prefix + ".bin", "wb") as f:
        f.write(x.tobytes())
    with open(prefix + ".json", "w") as f:
        f.write(json.dumps({"vertices": x.size // 3, "final_h": float(final_h), "dtype": "f32"},
                           separators=(",", ":")) + "\n")


def read_state(prefix):
    """An f32 state file (state-files.ts:18-23): (positions, final_h)."""
    with open(prefix + ".json") as f:
        meta = json.load(f)
    if not isinstance(meta, dict) or not isinstance(meta.get("final_h"), (int, float)) or meta.get("dtype") != "f32":
        raise DrapeError(f"{prefix}.json: not an f32 state file")
    x = np.fromfile(prefix + ".bin", "<f4")
    if "vertices" in meta and x.size != 3 * int(meta["vertices"]):
        raise DrapeError(f"{prefix}.bin: {x.size} floats for {meta['vertices']} vertices")
    return x, float(meta["final_h"])


def problem_colliders(problem):
    """The fine problem's colliders in f64 for the prolongation: the reference scene's own f64 geometry when
    problem.json names it (scene and scene_params, drape.rs:53-155), else problem.bin's f32 values widened."""
    m = problem.manifest
    sp = m.get("scene_params")
    if m.get("scene") == "sphere_on_floor" and isinstance(sp, dict):
        scene = reference_scene(int(m["grid_columns"]), sphere_on_floor=True,
                                material_factor=sp.get("material_factor", 1.0),
                                radius_factor=sp.get("radius_factor", 1.0), rest_factor=sp.get("rest_factor", 1.0),
                                placed=bool(sp.get("placed", False)))
        return scene_colliders(scene)
    return problem.arrays["colliders"].astype(np.float64)


def prolong_start(coarse_prefix, coarse_columns, fine_case_dir, out_prefix, log=print):
    """e7_prolong.rs:51-64 run: the coarse state prolonged to the fine case's mesh, projected out of its
    colliders and written as the fine start with the coarse solve's final step size; an infeasible start is
    refused."""
    problem = ProblemFile(fine_case_dir)
    x, h = read_state(coarse_prefix)
    fine, moved, bad = prolong_state(x, coarse_columns, problem_colliders(problem))
    if fine.size != 3 * problem.V:
        raise DrapeError(f"prolonged {fine.size // 3} vertices, the fine mesh has {problem.V}")
    log(f"prolongation: {moved} vertices projected to d = dhat, {bad} infeasible")
    if bad:
        raise DrapeError(f"infeasible start after projection: {bad} vertices")
    write_state(out_prefix, fine, h)
    return fine, h


def table(rows, fields):
    """Rows as one f64 table in field order (outputs.ts:15-16)."""
    names = [_snake(f) for f in fields]
    return np.array([[float(getattr(r, n)) for n in names] for r in rows], "<f8").reshape(len(rows), len(fields))


def write_run(gpu_dir, prefix, records):
    """gpu-<prefix>iterations.bin and gpu-<prefix>robust.bin (outputs.ts:28-31)."""
    table(records, RECORD_FIELDS).tofile(os.path.join(gpu_dir, f"gpu-{prefix}iterations.bin"))
    table(records, EXTRA_FIELDS).tofile(os.path.join(gpu_dir, f"gpu-{prefix}robust.bin"))


def write_probes_input(case_dir, steps, h, positions):
    """probes.json / probes.bin as the oracle writes them for a cold base solve (e7_run.rs:93-103): the 1-based
    probe steps (u32), the f32 step size and the f32 start state of each."""
    arrays = BinaryArrays()
    arrays.push("probe_steps", "u32", np.asarray(steps, np.uint32))
    arrays.push("probe_h", "f32", np.asarray(h, np.float32))
    arrays.push("probe_positions", "f32", np.asarray(positions, np.float32))
    arrays.write(os.path.join(case_dir, "probes.bin"))
    with open(os.path.join(case_dir, "probes.json"), "w") as f:
        json.dump({"probes": len(steps), "steps": [float(v) for v in steps], "arrays": arrays.entries}, f)


def load_probes(case_dir, vertices):
    """loadProbes (run-e7.ts:65-81): (steps, h, start states [n, 3V]) from probes.json / probes.bin, or None when
    the case has none."""
    try:
        with open(os.path.join(case_dir, "probes.json")) as f:
            manifest = json.load(f)
        with open(os.path.join(case_dir, "probes.bin"), "rb") as f:
            blob = f.read()
    except OSError:
        return None
    entries = manifest.get("arrays", [])
    steps = read_array(entries, blob, "probe_steps", "u32")
    h = read_array(entries, blob, "probe_h", "f32")
    positions = read_array(entries, blob, "probe_positions", "f32")
    if h.size != steps.size or positions.size != 3 * vertices * steps.size:
        raise DrapeError("probes.bin: probe arrays do not match the problem's vertices")
    return steps, h, positions.reshape(steps.size, 3 * vertices)


def run_probes(ctx, probes):
    """runPtcProbes (drape-e7.ts:183-196): one PTC attempt from each probe state at its step size, keeping every
    record and the positions at each record. Returns (records, iterates [n, 3V])."""
    records, iterates = [], []
    for step, h, start in zip(*probes):
        out = ctx.ptc_probe(start, float(h), int(step), cap=256, iterates=True)
        records += out["records"]
        iterates.append(out["iterates"])
    return records, np.concatenate(iterates) if iterates else np.zeros((0, 0), np.float32)


def write_probes(gpu_dir, records, iterates):
    """writeProbes (run-e7.ts:147-150): gpu-probe-iterations.bin, gpu-probe-robust.bin and
    gpu-probe-iterates.bin (the positions at each record, f32)."""
    write_run(gpu_dir, "probe-", records)
    np.ascontiguousarray(iterates, "<f4").tofile(os.path.join(gpu_dir, "gpu-probe-iterates.bin"))


def cap_for(columns):
    """capFor (ptc.ts:20): 1000 accepted steps up to n33, 400 above."""
    return 1000 if columns <= 33 else 400


def host_info():
    """hostInfo (hardware.ts): the host's OS and GPU."""
    return {"host_os": platform.platform(), "host_gpus": device_name()}


def write_solve(gpu_dir, mode, problem, out, adapter):
    """writeSolve (run-e7.ts:93-107) from a Drape.solve_equilibrium result."""
    os.makedirs(gpu_dir, exist_ok=True)
    res, steps, attempts, records = out["result"], out["steps"], out["attempts"], out["records"]
    write_run(gpu_dir, "", records)
    table(steps, PTC_FIELDS).tofile(os.path.join(gpu_dir, "gpu-ptc.bin"))
    table(attempts, ATTEMPT_FIELDS).tofile(os.path.join(gpu_dir, "gpu-attempts.bin"))
    np.ascontiguousarray(out["states"], "<f4").tofile(os.path.join(gpu_dir, "gpu-states.bin"))
    write_state(os.path.join(gpu_dir, "gpu-equilibrium"), out["states"][-1], res.final_h)
    updates = [r for r in records if np.isfinite(r.pcg_iterations)]
    columns = int(problem.manifest["grid_columns"])
    n = max(1, len(steps))
    ms = 1000.0 * res.seconds
    report = {
        "schema": "e7-gpu/1", "backend": "hip-drape", "adapter": adapter, "mode": mode, "settings": SETTINGS,
        "record_fields": RECORD_FIELDS, "robust_fields": EXTRA_FIELDS, "grid_columns": columns,
        "stop": out["stop"], "cap": cap_for(columns), "accepted_steps": len(steps),
        "rejected_attempts": sum(1 for a in attempts if a.reason != 0 and a.reason != CURVATURE_RETRY),
        "curvature_retry_attempts": sum(1 for a in attempts if a.reason == CURVATURE_RETRY),
        "curvature_dropped_steps": sum(1 for s in steps if s.curvature_dropped),
        "newton_iterations_total": len(updates),
        "pcg_iterations_total": int(sum(r.pcg_iterations for r in updates)),
        "final_h": res.final_h, "final_residual": res.final_residual,
        "final_relative_residual": res.final_residual / problem.gravity_load,
        "trajectory_ms": ms, "ms_per_step": ms / n,
        "phase_ms_per_step": {k: v / n for k, v in zip(PHASES, res.phase_ms)},
        "read_bytes_total": int(res.read_bytes),
        **host_info(),
    }
    with open(os.path.join(gpu_dir, "gpu-result.json"), "w") as f:
        f.write(json.dumps(report, indent=2) + "\n")
    return report


def run_checks(ctx, problem):
    """runChecks (run-e7.ts:140-145): the element kernels at the kernel state (exact, then projected) plus the
    inertia kernel without colliders (drape-e4.ts:108-120), the barrier at the contact state
    (contact-host.ts:76-90), and the inertia kernel at the check step sizes (drape-e7.ts:203-217)."""
    T, H, R = problem.T, problem.H, problem.R
    ne, nh = 9 * T + 12 * H, 81 * T + 144 * H
    state, rest = problem.array("kernel_state"), problem.array("rest_positions")
    row_mass = problem.array("row_mass")
    parts = []
    for exact in (True, False):
        ctx.set_tangent(exact)
        g, hs = ctx.eval_elements(state)
        parts += [g[:ne], hs[:nh]]
    ctx.set_tangent(problem.manifest["tangent"] == "exact")
    ctx.set_state(state, rest, 0.0, problem.kernel_mu)
    ctx.set_row_mass(row_mass)
    ctx.set_colliders(False)
    g, hs = ctx.eval_elements(state)
    ctx.set_colliders(True)
    parts += [g[ne:], hs[nh:]]
    kernel = np.concatenate(parts).astype(np.float32)
    contact = problem.arrays["contact_state"]
    if contact is None:
        raise DrapeError("problem.bin: colliders without contact_state")
    ctx.set_state(contact, contact, 0.0, 0.0)
    ctx.set_row_mass(np.zeros(R, np.float32))
    g, hs = ctx.eval_elements(contact)
    ctx.set_row_mass(row_mass)
    contact_kernel = np.concatenate([g[ne:], hs[nh:]]).astype(np.float32)
    inertia = []
    for h in CHECK_STEPS:
        ctx.set_state(state, rest, 0.0, 0.0)
        ctx.set_row_mass(np.array([m / (h * h) for m in problem.unit.astype(np.float64)], np.float32))
        ctx.set_colliders(False)
        g, hs = ctx.eval_elements(state)
        ctx.set_colliders(True)
        inertia += [g[ne:], hs[nh:]]
    return kernel, contact_kernel, np.concatenate(inertia).astype(np.float32)


def write_checks(gpu_dir, checks):
    """writeChecks (run-e7.ts:86-91)."""
    os.makedirs(gpu_dir, exist_ok=True)
    for name, values in zip(("gpu-kernel.bin", "gpu-contact-kernel.bin", "gpu-inertia-h.bin"), checks):
        np.ascontiguousarray(values, "<f4").tofile(os.path.join(gpu_dir, name))


def run_e7(case_dir, mode="cold", start=None, checks=False, preconditioner="block_jacobi", budget_s=0.0,
           keep=200000, log=print):
    """main of run-e7.ts: load the case, optionally run the kernel checks, solve the equilibrium by PTC from the
    flat release (cold) or from the state file `start` (warm / prolong) and write <case>/gpu-<mode>/."""
    if mode not in MODES:
        raise DrapeError(f"unknown mode '{mode}' ({'|'.join(MODES)})")
    if mode != "cold" and start is None:
        raise DrapeError(f"mode {mode} needs a start state prefix")
    problem = ProblemFile(case_dir)
    if int(problem.arrays["colliders"][0]) == 0 or problem.unit_mass is None or problem.ptc is None:
        raise DrapeError(f"{case_dir}: not a PTC problem (colliders, row_unit_mass and the ptc section are required)")
    ctx = Drape(problem, problem.rules())
    try:
        gpu_dir = os.path.join(case_dir, f"gpu-{mode}")
        os.makedirs(gpu_dir, exist_ok=True)
        if preconditioner != "block_jacobi":
            ctx.set_preconditioner(preconditioner)
        if checks:
            write_checks(gpu_dir, run_checks(ctx, problem))
            ctx.set_state(problem.array("rest_positions"), None, 0.0, 0.0)
            probes = load_probes(case_dir, problem.V)
            if probes is not None:
                write_probes(gpu_dir, *run_probes(ctx, probes))
        if start is None:
            x0, h0 = problem.array("rest_positions"), problem.ptc["h0"]
        else:
            x0, h0 = read_state(start)
            if x0.size != 3 * problem.V:
                raise DrapeError(f"{start}.bin: {x0.size // 3} vertices, the problem has {problem.V}")
        columns = int(problem.manifest["grid_columns"])
        out = ctx.solve_equilibrium(x0, h0, cap=cap_for(columns), budget_s=budget_s, keep=keep,
                                    keep_states=columns <= KEEP_STATES_COLUMNS)
        report = write_solve(gpu_dir, mode, problem, out, device_name())
    finally:
        ctx.close()
    phases = ", ".join(f"{k} {v * max(1, report['accepted_steps']) / 1000:.2f}s"
                       for k, v in report["phase_ms_per_step"].items())
    log(f"{problem.manifest['case']} {mode}: {report['accepted_steps']} steps (+{report['rejected_attempts']} "
        f"rejected), stop {report['stop']}, r/L {report['final_relative_residual']:.2e}, "
        f"{report['trajectory_ms'] / 1000:.2f} s [{phases}], h next {report['final_h']:.2e}")
    return report


def main(argv=None):
    """`--case DIR [--mode M] [--start PREFIX] [--checks]` runs run-e7.ts; `--coarse PREFIX --coarse-n N
    --fine-dir DIR --out PREFIX` writes a prolonged start as the oracle's e7-prolong does (main.rs:75-78)."""
    ap = argparse.ArgumentParser(description="PTC drape on a case directory (run-e7.ts) or a prolonged start "
                                             "(e7-prolong)")
    ap.add_argument("--case")
    ap.add_argument("--mode", default="cold", choices=MODES)
    ap.add_argument("--start")
    ap.add_argument("--checks", action="store_true")
    ap.add_argument("--preconditioner", default="block_jacobi", choices=["block_jacobi", "two_level"])
    ap.add_argument("--budget-s", type=float, default=float(os.environ.get("E7_BUDGET_S", "0") or 0))
    ap.add_argument("--coarse")
    ap.add_argument("--coarse-n", type=int, default=33)
    ap.add_argument("--fine-dir")
    ap.add_argument("--out")
    a = ap.parse_args(argv)
    try:
        if a.coarse is not None:
            if a.fine_dir is None or a.out is None:
                raise DrapeError("usage: --coarse <state prefix> --coarse-n <columns> --fine-dir <case dir> "
                                 "--out <state prefix>")
            prolong_start(a.coarse, a.coarse_n, a.fine_dir, a.out)
        elif a.case is not None:
            run_e7(a.case, a.mode, a.start, a.checks, a.preconditioner, a.budget_s)
        else:
            raise DrapeError("usage: --case <case dir> [--mode cold|warm|prolong] [--start <state prefix>] "
                             "[--checks]  |  --coarse <state prefix> --coarse-n N --fine-dir <dir> --out <prefix>")
    except DrapeError as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
