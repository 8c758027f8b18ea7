"""Test infrastructure, not product code: a restatement of the reference scorer's rule replays over the GPU
runner's output files (e7_compare.rs:112-136, 240-250; e4_compare.rs:45-84, 86-106, 258-260; e4_rules.rs:15-50),
so that the tests score this library's gpu-<mode>/ directories as the reference's `oracle e7-compare` would.
Only tests may import it.

Pinned by tests/test_file_contract.py: the forcing rule by closed forms of e4_rules.rs (eta_0, the clamp at
both ends, the squared ratio), and every replay by tables with a planted mistake that it must reject (a step
size off by one part in 1e9, an attempt with an update, an early update-rule stop, an eta off by 1e-4).
"""
import json
import os

import numpy as np

RECORD_FIELDS = 9            # e4_compare.rs:20
STATUS_UPDATE = 4.0          # e4_compare.rs:22
ETA_START, ETA_MIN, ETA_MAX, GAMMA, SAFEGUARD = 1e-2, 1e-7, 1e-2, 0.9, 0.1   # e4_rules.rs:6-11
UPDATE_STOP, RESOLUTION_FACTOR = 1e-4, 2.0                                    # e4_rules.rs:13, 34
CONTROLLER_TOLERANCE = 1e-12                                                  # e7_compare.rs:22


def read_f64_rows(path, width):
    """e4_compare.rs:45-49."""
    v = np.fromfile(path, "<f8")
    n = v.size // width
    return v[:n * width].reshape(n, width)


def read_iterates(gpu_dir, prefix=""):
    """e4_compare.rs:61-73: records joined with the robust fields named in gpu-result.json (NaN when absent)."""
    with open(os.path.join(gpu_dir, "gpu-result.json")) as f:
        fields = json.load(f)["robust_fields"]
    records = read_f64_rows(os.path.join(gpu_dir, f"gpu-{prefix}iterations.bin"), RECORD_FIELDS)
    extras = read_f64_rows(os.path.join(gpu_dir, f"gpu-{prefix}robust.bin"), len(fields))
    out = []
    for r, e in zip(records, extras):
        it = {"step": int(r[0]), "relative": r[2], "pcg_iterations": r[3], "status": r[8]}
        for name in ("eta", "mu", "updateInf", "stepInf", "retries", "resolution", "alphaMax", "minDistance",
                     "deltaAbs", "noiseOnly", "slope", "slopeExact", "normSquared", "normSquaredExact"):
            it[name] = e[fields.index(name)] if name in fields else float("nan")
        out.append(it)
    return out


def next_forcing(previous, ratio):
    """e4_rules.rs:15-21."""
    eta = GAMMA * ratio * ratio
    safeguard = GAMMA * previous * previous
    guarded = max(eta, safeguard) if safeguard > SAFEGUARD else eta
    return min(max(guarded, ETA_MIN), ETA_MAX)


def forcing_sequence(norms):
    """e4_rules.rs:24-31."""
    out, eta = [], ETA_START
    for k, norm in enumerate(norms):
        if k > 0:
            eta = next_forcing(eta, norm / norms[k - 1])
        out.append(eta)
    return out


def update_is_small(update_inf, step_inf, resolution):
    """e4_rules.rs:46-48."""
    return update_inf <= max(UPDATE_STOP * step_inf, RESOLUTION_FACTOR * resolution)


def resolution_of(it):
    """e4_compare.rs:258-260."""
    return it["resolution"] if np.isfinite(it["resolution"]) else 0.0


def forcing_replay(groups, forcing=True, fixed=float("nan")):
    """e4_compare.rs:87-94: the largest relative gap between a recorded eta and the f64 rule on the recorded
    residuals of the group's updates."""
    worst = 0.0
    for group in groups:
        updates = [it for it in group if np.isfinite(it["pcg_iterations"])]
        norms = [it["relative"] for it in updates]
        expected = forcing_sequence(norms) if forcing else [fixed] * len(norms)
        for it, eta in zip(updates, expected):
            worst = max(worst, abs((it["eta"] - eta) / eta))
    return worst


def termination_replay(groups, floor_stop=True):
    """e4_compare.rs:97-106: an update-rule stop is justified by its last update, and with the floor stop no
    earlier update satisfied the rule."""
    for group in groups:
        updates = [it for it in group if np.isfinite(it["pcg_iterations"]) and np.isfinite(it["updateInf"])]
        stopped = bool(group) and group[-1]["status"] == STATUS_UPDATE
        small = [update_is_small(it["updateInf"], it["stepInf"], resolution_of(it)) for it in updates]
        last_ok = not stopped or (bool(updates) and small[-1])
        earlier = small[:-1] if stopped else small
        if not last_ok or (floor_stop and any(earlier)):
            return False
    return True


def controller_replay(problem, rows):
    """e7_compare.rs:111-121: h_{k+1} = clamp(h_k growth (inner < cap) or h_k shrink) shrunk once per rejected
    attempt of step k+1, with rows of gpu-ptc.bin [step, h, residual, inner, rejected, pcg, min distance]."""
    p = problem["ptc"]
    growth, shrink, cap = p["growth"], p["shrink"], p["inner_max_iterations"]
    floor, top = p["h0"] * p["h_min_factor"], p["h_max"]
    for a, b in zip(rows[:-1], rows[1:]):
        factor = growth if a[3] < cap else shrink
        base = min(max(a[1] * factor, floor), top)
        expected = max(base * shrink ** int(b[4]), floor)
        if abs((b[1] - expected) / expected) > CONTROLLER_TOLERANCE:
            return False
    return True


def attempt_groups(iterates, attempts):
    """e7_compare.rs:124-126: the records of every attempt, from gpu-attempts.bin."""
    return [iterates[int(a[4]):int(a[5])] for a in attempts]


def rejection_replay(iterates, attempts):
    """e7_compare.rs:129-136: a rejection is a regularisation-cap (1, last status 5) or line-search (2, last
    status 3) failure before any update, a non-finite state (3) or a run superseded by the curvature trigger (4)."""
    for a in attempts:
        if a[3] == 0.0:
            continue
        records = iterates[int(a[4]):int(a[5])]
        updates = sum(1 for r in records if r["status"] == 0.0 and np.isfinite(r["pcg_iterations"]))
        last = records[-1]["status"] if records else float("nan")
        ok = a[3] in (3.0, 4.0) or (updates == 0 and ((a[3] == 1.0 and last == 5.0) or (a[3] == 2.0 and last == 3.0)))
        if not ok:
            return False
    return True


def replays(problem, gpu_dir):
    """The replays of e7_compare.rs:240-250 without the probe replays (no probe outputs): a dict of the
    scorer's numbers and flags."""
    steps = read_f64_rows(os.path.join(gpu_dir, "gpu-ptc.bin"), 7)
    attempts = read_f64_rows(os.path.join(gpu_dir, "gpu-attempts.bin"), 6)
    iterates = read_iterates(gpu_dir)
    groups = attempt_groups(iterates, attempts)
    forcing = forcing_replay(groups, True)
    out = {"forcing_replay_error": forcing, "termination_replay_ok": termination_replay(groups, True),
           "controller_replay_ok": controller_replay(problem, steps),
           "rejection_replay_ok": rejection_replay(iterates, attempts)}
    out["rules_replayed_ok"] = (forcing <= 1e-5 and out["termination_replay_ok"] and out["controller_replay_ok"]
                                and out["rejection_replay_ok"])
    return out


ONE_STEP_LIMIT = 1e-3        # e7_compare.rs:16
FLOOR_FACTOR = 2.0           # e7_compare.rs:17
STATUS_RESIDUAL = 1.0        # e7_compare.rs:24


def probe_results(gpu_dir, vertices, previous, following):
    """e7_compare.rs:170-190 probes: per probed step (records grouped by step), the GPU's last iterate against the
    oracle's next state, relative to the oracle's step: |x_gpu - x_s| / |x_s - x_{s-1}|; the f32 floor (the least
    relative residual of the step's records) and the stop status of its last record. `previous` / `following`
    hold the oracle's states x_{s-1} and x_s of each probed step, in the order of the probe files."""
    iterates = read_iterates(gpu_dir, "probe-")
    states = np.fromfile(os.path.join(gpu_dir, "gpu-probe-iterates.bin"), "<f4").reshape(-1, 3 * vertices)
    if len(states) != len(iterates):
        raise ValueError("gpu-probe-iterates.bin does not hold one state per record")
    groups = []
    for i, it in enumerate(iterates):
        if groups and groups[-1][0] == it["step"]:
            groups[-1][1].append(i)
        else:
            groups.append((it["step"], [i]))
    out = []
    for k, (step, idx) in enumerate(groups):
        x = states[idx[-1]].astype(np.float64)
        prev, nxt = np.asarray(previous[k], np.float64), np.asarray(following[k], np.float64)
        out.append({"step": step, "error": float(np.linalg.norm(x - nxt) / np.linalg.norm(nxt - prev)),
                    "f32_floor": min(iterates[i]["relative"] for i in idx), "status": iterates[idx[-1]]["status"],
                    "last": x})
    return out


def probe_stops_ok(status, f32_floor, representation_floor):
    """e7_compare.rs:160-163 Probe::stops_ok: the step stopped on its residual, or by the update rule at a floor
    within twice the representation floor."""
    return status == STATUS_RESIDUAL or (status == STATUS_UPDATE and f32_floor <= FLOOR_FACTOR * representation_floor)
