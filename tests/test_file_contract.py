"""CPU tests of the file contract (drape_amd/filecontract.py): the BinaryArrays layout reproduces the manifests of
the reference's committed problem.json files byte offset for byte offset, a case directory reads back into the
same arrays and into a problem the library accepts (its GPU-path arrays equal the library's own scene builder
bit for bit), malformed files are refused, the state and record files have the reference's layout, and the
scorer replays restated in oracle/scorer.py are pinned."""
import json
import os

import numpy as np
import pytest

import drape_amd as D
from drape_amd import filecontract as FC
from contract_case import CONTRACT, problem_arrays, write_case
from oracle import scorer as S

CASES = sorted(CONTRACT["problems"])


@pytest.mark.parametrize("rel", CASES)
def test_manifest_matches_the_reference_byte_for_byte(rel, tmp_path):
    """Names, dtypes, lengths and byte offsets of every array equal the reference's problem.json manifest
    (io.rs:5-55), and the header numbers the generator computes (kappa, gravity load) agree."""
    od, entries = write_case(rel, str(tmp_path))
    ref = CONTRACT["problems"][rel]
    assert entries == ref["arrays"]
    size = os.path.getsize(tmp_path / "problem.bin")
    last = ref["arrays"][-1]
    assert size == last["offset"] + 4 * last["length"]
    assert od.kappa == pytest.approx(ref["contact"]["kappa"], rel=1e-12)
    assert od.gravity_load == pytest.approx(ref["ptc"]["gravity_load"], rel=1e-12)
    doc = json.loads((tmp_path / "problem.json").read_text())
    assert list(doc) == list(ref)


@pytest.mark.parametrize("rel", ["drape-n17/base", "drape-n33/w2"])
def test_case_reads_back_into_the_library_problem(rel, tmp_path):
    od, _ = write_case(rel, str(tmp_path))
    pf = FC.ProblemFile(str(tmp_path))
    for name, dtype, values in problem_arrays(od):
        got = pf.unit_mass if name == "row_unit_mass" else pf.arrays[name]
        assert got.dtype == FC.DTYPES[dtype] and np.array_equal(got, values), name
    assert (pf.V, pf.T, pf.H, pf.R, pf.B) == (od.V, od.T, od.H, od.R, od.B)
    # the library's own generator gives the same GPU-path arrays bit for bit
    sp = CONTRACT["problems"][rel]["scene_params"]
    sa = D.SceneArrays(D.reference_scene(pf.manifest["grid_columns"], sphere_on_floor=True,
                                         material_factor=sp["material_factor"], radius_factor=sp["radius_factor"],
                                         rest_factor=sp["rest_factor"], placed=sp["placed"]))
    for name in ("rest_positions", "elem_vertices", "elem_geom", "hinge_vertices", "hinge_stiffness", "row_vertex",
                 "row_unit_mass", "row_ptr", "col_idx", "colliders"):
        a, b = pf.array(name), sa.array(name)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    p, q = pf.problem, sa.problem
    for k in ("wale", "course", "coupling", "shear", "gravity", "time_step", "grid_columns", "collider_floats"):
        assert getattr(p, k) == getattr(q, k), k
    assert pf.gravity_load == pytest.approx(sa.gravity_load, rel=1e-12)


def test_rules_follow_problem_json(tmp_path):
    write_case("drape-n17/base", str(tmp_path))
    pf = FC.ProblemFile(str(tmp_path))
    r, d = pf.rules(), D.default_rules()
    for k in ("newton_tolerance", "newton_max_iterations", "armijo", "max_halvings", "pcg_tolerance",
              "pcg_max_iterations", "replace_every", "replace_above", "eta_start", "eta_min", "eta_max", "eta_gamma",
              "eta_safeguard", "update_stop", "resolution_factor", "mu_start", "mu_cap", "h0", "h_max",
              "h_min_factor", "growth", "shrink", "inner_tolerance", "inner_max_iterations", "max_rejections",
              "equilibrium_f32", "stagnation_window", "stagnation_factor"):
        assert getattr(r, k) == getattr(d, k), k
    assert r.gravity_load == CONTRACT["problems"]["drape-n17/base"]["ptc"]["gravity_load"]
    assert r.exact_tangent == 1 and pf.kernel_mu == 0.37
    # a changed constant in the file changes the rule
    doc = json.loads((tmp_path / "problem.json").read_text())
    doc["ptc"]["growth"] = 3.0
    doc["tangent"] = "projected"
    (tmp_path / "problem.json").write_text(json.dumps(doc))
    r2 = FC.ProblemFile(str(tmp_path)).rules()
    assert r2.growth == 3.0 and r2.exact_tangent == 0


@pytest.mark.parametrize("damage", ["missing_field", "wrong_type", "missing_array", "wrong_dtype", "truncated",
                                    "inconsistent"])
def test_malformed_case_is_refused(damage, tmp_path):
    write_case("drape-n17/base", str(tmp_path))
    doc = json.loads((tmp_path / "problem.json").read_text())
    if damage == "missing_field":
        del doc["triangles"]
    elif damage == "wrong_type":
        doc["material"]["wale_N_per_m"] = "stiff"
    elif damage == "missing_array":
        doc["arrays"] = [e for e in doc["arrays"] if e["name"] != "col_idx"]
    elif damage == "wrong_dtype":
        next(e for e in doc["arrays"] if e["name"] == "row_ptr")["dtype"] = "f32"
    elif damage == "truncated":
        blob = (tmp_path / "problem.bin").read_bytes()
        (tmp_path / "problem.bin").write_bytes(blob[:len(blob) // 2])
    elif damage == "inconsistent":
        doc["blocks"] += 1
    (tmp_path / "problem.json").write_text(json.dumps(doc))
    with pytest.raises(D.DrapeError):
        FC.ProblemFile(str(tmp_path))


def test_binary_arrays_alignment():
    ba = FC.BinaryArrays()
    ba.push("a", "f32", [1.0, 2.0, 3.0])
    ba.push("b", "f64", [4.0])
    ba.push("c", "u32", [5])
    ba.push("d", "f32", [])
    ba.push("e", "u32", [6, 7])
    assert [(e["name"], e["length"], e["offset"]) for e in ba.entries] == \
        [("a", 3, 0), ("b", 1, 16), ("c", 1, 24), ("d", 0, 32), ("e", 2, 32)]
    blob = ba.tobytes()
    assert len(blob) == 40
    assert FC.read_array(ba.entries, blob, "b", "f64")[0] == 4.0
    assert list(FC.read_array(ba.entries, blob, "e", "u32")) == [6, 7]


def test_state_files_have_the_reference_format(tmp_path):
    ref = CONTRACT["state_files"]["drape-n17/base/gpu-cold/gpu-equilibrium.json"]
    x = np.arange(3 * 289, dtype=np.float32)
    FC.write_state(str(tmp_path / "s"), x, 1.0 / 15.0)
    assert (tmp_path / "s.json").read_text() == ref
    y, h = FC.read_state(str(tmp_path / "s"))
    assert np.array_equal(x, y) and h == 1.0 / 15.0
    assert (tmp_path / "s.bin").stat().st_size == 4 * x.size
    # the oracle's f64 state files are not GPU starts (state-files.ts:18-20)
    (tmp_path / "t.json").write_text('{"vertices":1,"final_h":0.25,"dtype":"f64"}')
    (tmp_path / "t.bin").write_bytes(b"\0" * 24)
    with pytest.raises(D.DrapeError):
        FC.read_state(str(tmp_path / "t"))


def test_record_tables_are_f64_rows_in_field_order(tmp_path):
    recs = (D.Record * 3)()
    for i in range(3):
        for j, name in enumerate(D.RECORD_FIELDS + D.EXTRA_FIELDS):
            setattr(recs[i], name, 100 * i + j)
    FC.write_run(str(tmp_path), "", list(recs))
    it = S.read_f64_rows(str(tmp_path / "gpu-iterations.bin"), 9)
    rb = S.read_f64_rows(str(tmp_path / "gpu-robust.bin"), 16)
    assert it.shape == (3, 9) and rb.shape == (3, 16)
    assert np.array_equal(it[1], 100 + np.arange(9)) and np.array_equal(rb[2], 209 + np.arange(16))
    assert FC.RECORD_FIELDS == CONTRACT["gpu_results"]["drape-n17/base/gpu-cold"]["record_fields"]
    assert FC.EXTRA_FIELDS == CONTRACT["gpu_results"]["drape-n17/base/gpu-cold"]["robust_fields"]
    assert FC.SETTINGS == CONTRACT["gpu_results"]["drape-n17/base/gpu-cold"]["settings"]
    assert D.PHASES == CONTRACT["gpu_results"]["drape-n17/base/gpu-cold"]["phase_keys"]
    assert [FC.cap_for(r["grid_columns"]) for r in CONTRACT["gpu_results"].values()] == \
        [r["cap"] for r in CONTRACT["gpu_results"].values()]


# ---- the scorer restatement -----------------------------------------------------------------------------
def test_forcing_rule_closed_forms():
    assert S.forcing_sequence([1.0]) == [1e-2]
    # ratio 0.5: 0.9 * 0.25 = 0.225 -> clamped to eta_max; ratio 1e-2: 9e-5 inside; ratio 1e-5: clamped to 1e-7
    assert S.forcing_sequence([1.0, 0.5]) == [1e-2, 1e-2]
    assert S.forcing_sequence([1.0, 1e-2])[1] == pytest.approx(9e-5, rel=1e-15)
    assert S.forcing_sequence([1.0, 1e-5])[1] == 1e-7
    # the safeguard 0.9 eta_{k-1}^2 acts once it exceeds 0.1 (eta_{k-1} > 1/3), never from inside the clamp
    assert S.next_forcing(0.5, 1e-3) == 1e-2
    assert S.next_forcing(1e-2, 1e-3) == pytest.approx(9e-7, rel=1e-15)
    assert S.update_is_small(1e-9, 1e-4, 0.0) and not S.update_is_small(2e-8, 1e-4, 0.0)
    assert S.update_is_small(2e-8, 1e-4, 1e-8)


def _synthetic_run(tmp_path):
    """A consistent two-step run: step 1 accepted after one line-search rejection, step 2 accepted."""
    problem = CONTRACT["problems"]["drape-n17/base"]
    h0 = problem["ptc"]["h0"]
    recs = (D.Record * 5)()
    rows = [  # step, iteration, relative, pcg, pcg status, lin, alpha, halvings, status; eta, update, step_inf
        (1, 0, 1.0, float("nan"), 0, 0, 0, 30, 3, float("nan"), float("nan"), float("nan")),
        (1, 0, 1.0, 12, 1, 1e-3, 1, 0, 0, 1e-2, 1e-3, 1e-3),
        (1, 1, 0.05, float("nan"), 0, 0, 0, 0, 1, float("nan"), float("nan"), float("nan")),
        (2, 0, 1.0, 9, 1, 1e-3, 1, 0, 0, 1e-2, 1e-3, 1e-3),
        (2, 1, 0.01, 7, 1, 1e-3, 1, 0, 4, 9e-5, 1e-9, 1e-3),
    ]
    for r, (step, it, rel, pcg, ps, lin, a, hv, st, eta, up, si) in zip(recs, rows):
        r.step, r.iteration, r.relative_residual, r.pcg_iterations, r.pcg_status = step, it, rel, pcg, ps
        r.linear_residual, r.step_length, r.halvings, r.status = lin, a, hv, st
        r.eta, r.update_inf, r.step_inf, r.resolution = eta, up, si, 0.0
    FC.write_run(str(tmp_path), "", list(recs))
    h1 = h0 * 0.5
    h2 = min(max(h1 * 2.0, h0 * 0.1), 1.0)
    steps = np.array([[1, h1, 1.0, 1, 1, 12, 1e-3], [2, h2, 0.5, 2, 0, 16, 1e-3]], "<f8")
    attempts = np.array([[1, 0, h0, 2, 0, 1], [1, 1, h1, 0, 1, 3], [2, 0, h2, 0, 3, 5]], "<f8")
    steps.tofile(tmp_path / "gpu-ptc.bin")
    attempts.tofile(tmp_path / "gpu-attempts.bin")
    (tmp_path / "gpu-result.json").write_text(json.dumps({"robust_fields": FC.EXTRA_FIELDS}))
    return problem, steps, attempts


def test_scorer_replays_accept_a_consistent_run_and_reject_planted_mistakes(tmp_path):
    problem, steps, attempts = _synthetic_run(tmp_path)
    ok = S.replays(problem, str(tmp_path))
    assert ok["rules_replayed_ok"], ok
    # a step size off by one part in 1e9
    bad = steps.copy()
    bad[1, 1] *= 1 + 1e-9
    assert not S.controller_replay(problem, bad)
    # a rejected attempt that holds an update
    its = S.read_iterates(str(tmp_path))
    bad_att = attempts.copy()
    bad_att[1, 3] = 2
    assert not S.rejection_replay(its, bad_att)
    # an eta off by 1e-4 relative
    groups = S.attempt_groups(its, attempts)
    groups[2][1]["eta"] *= 1 + 1e-4
    assert S.forcing_replay(groups) > 1e-5
    # an update that met the stop rule but was followed by another
    its = S.read_iterates(str(tmp_path))
    groups = S.attempt_groups(its, attempts)
    groups[2][0]["updateInf"] = 1e-9
    assert not S.termination_replay(groups, True)


def test_command_line_refuses_incomplete_arguments(tmp_path, capsys):
    assert FC.main([]) == 1
    assert FC.main(["--coarse", "x"]) == 1
    assert FC.main(["--case", str(tmp_path), "--mode", "warm"]) == 1    # warm needs a start
    assert "error:" in capsys.readouterr().err
