"""Writes tests/golden/file_contract.json from the reference's committed experiment files: the problem.json of
every Experiment 7 case (header sections and the problem.bin manifest; the .bin files themselves are not
committed upstream), the key set and field lists of the GPU runner's gpu-result.json, a state file header, and
the probe scores the reference's scorer gave its own GPU runs.

Run once in a container that holds the reference: python tests/golden/make_contract_fixtures.py [/root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["drape-n17/base", "drape-n33/base", "drape-n33/w1", "drape-n33/w2", "drape-n33/w3", "drape-n65/base",
         "drape-n129/base"]


def main(ref):
    exp = os.path.join(ref, "experiments", "drape7")
    out = {"problems": {}, "gpu_results": {}, "state_files": {}}
    for rel in CASES:
        with open(os.path.join(exp, rel, "problem.json")) as f:
            out["problems"][rel] = json.load(f)
    for rel in ("drape-n17/base/gpu-cold", "drape-n33/base/gpu-cold", "drape-n33/w1/gpu-warm"):
        with open(os.path.join(exp, rel, "gpu-result.json")) as f:
            r = json.load(f)
        out["gpu_results"][rel] = {"keys": list(r), "schema": r["schema"], "settings": r["settings"],
                                   "record_fields": r["record_fields"], "robust_fields": r["robust_fields"],
                                   "phase_keys": list(r["phase_ms_per_step"]), "cap": r["cap"],
                                   "grid_columns": r["grid_columns"], "mode": r["mode"]}
        with open(os.path.join(exp, rel, "gpu-equilibrium.json")) as f:
            out["state_files"][rel + "/gpu-equilibrium.json"] = f.read()
    with open(os.path.join(exp, "drape-n33/base/start-prolong-gpu.json")) as f:
        out["state_files"]["drape-n33/base/start-prolong-gpu.json"] = f.read()
    # the oracle's PTC from the n17 equilibrium prolonged to n33 (e7_prolong.rs; e7_run.rs solve, mode prolong)
    base = os.path.join(exp, "drape-n33", "base")
    with open(os.path.join(base, "prolong.log")) as f:
        lines = [line.rstrip("\n") for line in f if line.startswith("ptc step")]
    with open(os.path.join(base, "reference-prolong.json")) as f:
        r = json.load(f)
    with open(os.path.join(base, "start-prolong-oracle.json")) as f:
        start = json.load(f)
    out["prolong_n33"] = {"log": lines, "start": start,
                          "summary": {k: r[k] for k in ("reason", "ptc_steps", "rejected_attempts", "newton_iterations",
                                                        "final_relative_residual", "static_energy", "final_h")}}
    # the reference runner's own probe scores (e7_compare.rs:320-327 on its WebGPU run): step, one-step error,
    # stop status, from the scored result.json of the cold base runs
    out["gpu_probes"] = {}
    for rel in ("drape-n17/base/gpu-cold", "drape-n33/base/gpu-cold"):
        with open(os.path.join(exp, rel, "result.json")) as f:
            r = json.load(f)
        out["gpu_probes"][rel] = {"probes": [{k: p[k] for k in ("step", "one_step_displacement_error", "f32_floor",
                                                                  "representation_floor", "stop_status", "stop_ok")}
                                             for p in r["probes"]],
                                  "verdict": {k: r["verdict"][k] for k in ("one_step_ok", "f32_stops_ok_at_probes")}}
    with open(os.path.join(HERE, "file_contract.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "file_contract.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
