"""Test helper: a case directory in the reference's Experiment 7 layout (problem.json + problem.bin), written from
the oracle's restatement of the reference generator (e3_generate.rs:84-104 problem_arrays, e7_run.rs:41-50
generate) with this library's BinaryArrays writer. The reference's own .bin files are not committed upstream;
the header sections come from the reference's committed problem.json of the same case."""
import json
import os

import numpy as np

from drape_amd import filecontract as FC
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "file_contract.json")) as f:
    CONTRACT = json.load(f)


def oracle_drape(rel):
    """The oracle drape of a drape7 case from its columns and scene_params."""
    p = CONTRACT["problems"][rel]
    sp = p["scene_params"]
    scene = O.reference_scene(p["grid_columns"], sphere=p["scene"] == "sphere_on_floor",
                              material_factor=sp["material_factor"], radius_factor=sp["radius_factor"],
                              rest_factor=sp["rest_factor"], placed=sp["placed"])
    return O.Drape(scene)


def problem_arrays(od):
    """[(name, dtype, values)] in the reference's order (e3_generate.rs:84-104, then row_unit_mass,
    e7_run.rs:46)."""
    a = od.problem_arrays()
    fixed = np.setdiff1d(np.arange(od.V, dtype=np.uint32), a["row_vertex"]).astype(np.uint32)
    empty_u, empty_f = np.zeros(0, np.uint32), np.zeros(0, np.float32)
    out = [("rest_positions", "f32", a["rest_positions"]), ("elem_vertices", "u32", a["elem_vertices"]),
           ("elem_geom", "f32", a["elem_geom"]), ("hinge_vertices", "u32", a["hinge_vertices"]),
           ("hinge_stiffness", "f32", a["hinge_stiffness"]), ("row_vertex", "u32", a["row_vertex"]),
           ("dof_mask", "u32", np.full(od.R, 7, np.uint32)), ("row_mass", "f32", a["row_mass"]),
           ("row_ptr", "u32", a["row_ptr"]), ("col_idx", "u32", a["col_idx"]), ("diag_idx", "u32", a["diag_idx"]),
           ("block_contrib_ptr", "u32", a["block_contrib_ptr"]), ("block_contrib_idx", "u32", a["block_contrib_idx"]),
           ("force_contrib_ptr", "u32", a["force_contrib_ptr"]), ("force_contrib_idx", "u32", a["force_contrib_idx"]),
           ("fixed_vertex", "u32", fixed), ("kernel_state", "f32", a["kernel_state"]),
           ("probe_steps", "u32", empty_u), ("probe_positions", "f32", empty_f), ("probe_velocities", "f32", empty_f)]
    if a["colliders"].size:
        out += [("colliders", "f32", a["colliders"]), ("contact_state", "f32", a["contact_state"])]
    out.append(("row_unit_mass", "f32", a["row_unit_mass"]))
    return out


def write_case(rel, case_dir):
    """Writes the case; returns (oracle drape, manifest entries)."""
    od = oracle_drape(rel)
    header = {k: v for k, v in CONTRACT["problems"][rel].items() if k != "arrays"}
    entries = FC.write_problem(case_dir, header, problem_arrays(od))
    return od, entries
