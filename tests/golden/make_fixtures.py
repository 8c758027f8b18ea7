"""Extracts the golden vectors this test-suite checks the oracle against from the reference's committed
experiment data (JSON results and progress logs under experiments/). Only numbers are copied: no
source file of the reference is kept. Run once in a container that has the reference mounted:

    python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load(path):
    with open(path) as f:
        return json.load(f)


def manifest_lengths(problem):
    return {a["name"]: a["length"] for a in problem["arrays"]}


def main(ref):
    exp = os.path.join(ref, "experiments")
    out = {}
    # E1 linear PCG known answers: experiments/cases/<kind>-n<n>/{problem,reference}.json
    e1 = {}
    for kind in ("tension", "dynamic", "slack"):
        for n in (33, 65, 129, 257):
            d = os.path.join(exp, "cases", f"{kind}-n{n}")
            if not os.path.exists(os.path.join(d, "reference.json")):
                continue
            p, r = load(os.path.join(d, "problem.json")), load(os.path.join(d, "reference.json"))
            e1[f"{kind}-n{n}"] = {
                "rows": p["rows"], "blocks": p["blocks"], "replace_every": p["replace_every"],
                "replace_above": p["replace_above"],
                "bandwidth": r["cholesky_f64"]["bandwidth"],
                "f64_iterations": r["pcg_f64_cpu"]["iterations"], "f64_history": r["pcg_f64_cpu"]["history"],
                "f32_iterations": r["pcg_f32_cpu"]["iterations"], "f32_history": r["pcg_f32_cpu"]["history"],
                "f32_converged": r["pcg_f32_cpu"]["converged"],
                "condition": r["spectrum_jacobi_preconditioned"]["condition"],
            }
    out["e1"] = e1
    # Problem structure of the drape problems (array lengths of problem.bin manifests).
    structure = {}
    for rel in ("drape7/drape-n17/base", "drape7/drape-n33/base", "drape7/drape-n65/base", "drape7/drape-n129/base",
                "drape4-final/drape-n33", "drape5/drape-n33", "drape5/drape-n65"):
        p = load(os.path.join(exp, rel, "problem.json"))
        structure[rel] = {"grid_columns": p["grid_columns"], "scene": p.get("scene", "pinned_corners"),
                          "placed": bool(p.get("scene_params", {}).get("placed", False)),
                          "lengths": manifest_lengths(p),
                          "kappa": p.get("contact", {}).get("kappa"),
                          "gravity_load": p.get("ptc", {}).get("gravity_load")}
    out["structure"] = structure
    # PTC f64 traces (e7_ptc.rs:183 progress lines) and summaries.
    ptc = {}
    for n in (17, 33):
        d = os.path.join(exp, "drape7", f"drape-n{n}", "base")
        with open(os.path.join(d, "cold.log")) as f:
            lines = [line.rstrip("\n") for line in f if line.startswith("ptc step")]
        r = load(os.path.join(d, "reference-cold.json"))
        ptc[f"n{n}"] = {"log": lines, "summary": {k: r[k] for k in ("reason", "ptc_steps", "newton_iterations",
                                                                   "final_relative_residual", "gravity_load",
                                                                   "static_energy", "max_mu")}}
    out["ptc"] = ptc
    # E4-final n33: f64 per-step Newton residual histories and the derivative check.
    r = load(os.path.join(exp, "drape4-final", "drape-n33", "reference.json"))
    out["e4_final_n33"] = {"derivative_check": r["derivative_check"],
                           "steps": [{k: s[k] for k in ("step", "status", "newton_iterations", "relative_residuals",
                                                         "scale", "total_energy_J")} for s in r["steps"][:12]]}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_vectors.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
