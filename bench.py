"""Benchmark of the implicit Newton step at n129: ms per Newton iteration and PCG iterations per Newton
iteration of the quasi-static sphere drape solved by pseudo-transient continuation (129 x 129 grid,
16641 vertices, 32768 triangles; isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm, 0.2 kg/m^2, hinge 1e-5 N m,
sphere r = 0.15 m, dhat = 1 mm), the learned two-level preconditioner against block-Jacobi FP32 PCG.

A bench step is one accepted PTC step (one inexact Newton solve of the step's energy, several Newton
iterations). `--warmup W` steps run untimed from the flat start; the next `--steps K` are timed, once with
the selected preconditioner (the reported value; by default the two-level preconditioner with the learned
prolongation shipped with the package), then from the same state with the bilinear two-level preconditioner
and with block-Jacobi (the comparisons in `config`). Prints one JSON line.

The roofline object is for the SpMV of PCG, the dominant kernel of the block-Jacobi solve (92% of the
reference's n129 run) and the largest of the two-level iteration: algorithmic bytes 40 B per block + 28 B
per row + 4 (vals 36 B + column 4 B per block; row pointer 4 B, x gather 12 B and y 12 B per row), divided
by the kernel's average duration from HIP events on the context's stream. The CPU baseline is the oracle
(f64 Newton with band Cholesky, as the reference's Rust oracle does) timed for one Newton iteration of the
same first PTC step on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import drape_amd as D  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of k_pcg_spmv, from rocprofv3 FETCH_SIZE and WRITE_SIZE counters scaled by the same
# counters of a known-size copy (FETCH_SIZE x2.0, WRITE_SIZE x1.0 on this part); keyed by grid columns.
# (At n129 the bytes are L2 misses served mostly by the Infinity Cache.)
SPMV_TRAFFIC = {129: 8_574_432, 1025: 523_831_470}


def spmv_bytes(rows: int, blocks: int) -> int:
    return 40 * blocks + 28 * rows + 4


def large_grid_spmv(columns: int, launches: int = 50) -> dict:
    """The plain SpMV of PCG on a grid whose matrix exceeds the 256 MB Infinity Cache (n1025: 11.5 M blocks,
    414 MB of values), timed with HIP events: the HBM-bound roofline of the kernel."""
    scene = D.drape_config_scene(columns)
    arrays = D.SceneArrays(scene)
    ctx = D.Drape(arrays)
    rest = arrays.array("rest_positions")
    ctx.set_state(rest, rest, 1.0 / 60.0, 0.0)
    ctx.residual()
    ctx.solve_once(1e-2, 2)   # fills the direction vector the SpMV reads
    spmv_ms, _ = ctx.spmv_timing(launches)
    nbytes = spmv_bytes(arrays.R, arrays.B)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": SPMV_TRAFFIC.get(columns), "kernel": "k_pcg_spmv",
           "avg_kernel_us": round(1000.0 * spmv_ms, 3), "algorithmic_bytes": nbytes,
           "workload": f"n{columns} ({arrays.R} rows, {arrays.B} blocks)"}
    ctx.close()
    return out


def copy_calibration(mbytes: int = 512) -> dict:
    """A plain device copy of known size (k_copy: read + write), for scaling the PMC byte counters."""
    nbytes = mbytes << 20
    ms = D.copy_timing(nbytes, 10)
    return {"kernel": "k_copy", "bytes_read": nbytes, "bytes_written": nbytes, "avg_kernel_us": round(1000.0 * ms, 3),
            "achieved_gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def cpu_baseline(columns: int) -> dict:
    """One f64 Newton iteration (regularised band-Cholesky direction + Armijo, e7_ptc.rs:129-134) of the first
    PTC step of the same scene, on one core of this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    s = O.Scene()
    ps = D.drape_config_scene(columns)
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(ps, name))
    od = O.Drape(s)
    x = od.rest_positions()
    seconds, ok = od.time_newton_iteration(x, 1.0 / 240.0)
    iterations = 1
    try:
        cpu = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in cpu.splitlines() if l.startswith("Model name")), "?")
    except OSError:
        model = "?"
    return {"value": 1000.0 * seconds / iterations, "unit": "ms/Newton iteration", "cores": 1, "kind": "port",
            "sample": f"one f64 Newton iteration (tangent, band-Cholesky direction, Armijo search) of the first "
                      f"PTC step of the n{columns} drape, {seconds:.1f} s on 1 core ({model})"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--columns", type=int, default=129)
    ap.add_argument("--preconditioner", choices=["learned", "two_level", "block_jacobi"], default="learned",
                    help="learned: two-level with the learned prolongation; two_level: with bilinear prolongation")
    ap.add_argument("--no-compare", action="store_true", help="skip the comparison runs")
    ap.add_argument("--roofline-columns", type=int, default=1025,
                    help="grid for the HBM-bound SpMV roofline (0 skips it)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the large-grid SpMV and a copy of known size (for PMC byte counters)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--self-contact", action="store_true",
                    help="turn on the sheet's self-contact (point-triangle / edge-edge barriers) in every run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: multi-GPU runs (domain decomposition with one "
                                   "halo exchange per PCG iteration) are not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.spmv_only:
        print(json.dumps({"spmv": large_grid_spmv(args.roofline_columns or 1025), "copy": copy_calibration()}),
              flush=True)
        return

    scene = D.drape_config_scene(args.columns)
    arrays = D.SceneArrays(scene)
    kind = 0 if args.preconditioner == "block_jacobi" else 1

    def context(pre):
        c = D.Drape(arrays, D.default_rules(preconditioner=0 if pre == "block_jacobi" else 1,
                                            self_contact=1 if args.self_contact else 0))
        if pre == "learned":
            from drape_amd.learned import load_model
            c.set_prolongation_model(load_model())
        return c

    ctx = context(args.preconditioner)
    print(f"[bench] n{args.columns}: V {arrays.V} T {arrays.T} H {arrays.H} rows {arrays.R} blocks {arrays.B}",
          file=sys.stderr, flush=True)

    start = arrays.array("rest_positions")
    h = ctx.rules.h0
    if args.warmup > 0:
        w = ctx.solve_equilibrium(start=start, h_start=h, cap=args.warmup)
        start, h = w["positions"], w["result"].final_h
        print(f"[bench] warmup: {w['result'].accepted_steps} steps, stop {w['stop']}", file=sys.stderr, flush=True)

    def timed(context, label):
        t0 = time.perf_counter()
        out = context.solve_equilibrium(start=start, h_start=h, cap=args.steps)
        wall = time.perf_counter() - t0
        res = out["result"]
        newton = max(res.newton_iterations, 1)
        print(f"[bench] {label}: {res.accepted_steps} steps, {res.newton_iterations} Newton, {res.pcg_iterations} "
              f"PCG, {res.seconds:.3f} s, fallbacks {res.fallbacks}, stop {out['stop']}, "
              f"r/L {res.final_residual / res.gravity_load:.3e}", file=sys.stderr, flush=True)
        return out, res, newton, wall

    out, res, newton, wall = timed(ctx, args.preconditioner)
    ms_per_newton = 1000.0 * res.seconds / newton
    pcg_per_newton = res.pcg_iterations / newton
    comparisons = {}
    if not args.no_compare:
        # the same steps from the same state with the other preconditioners
        for pre in ("learned", "two_level", "block_jacobi"):
            if pre == args.preconditioner:
                continue
            other = context(pre)
            _, ores, onewton, _ = timed(other, pre)
            comparisons[pre] = {"ms_per_newton": round(1000.0 * ores.seconds / onewton, 4),
                                "pcg_per_newton": round(ores.pcg_iterations / onewton, 2),
                                "newton_iterations": ores.newton_iterations, "pcg_iterations": ores.pcg_iterations,
                                "fallbacks": ores.fallbacks, "accepted_steps": ores.accepted_steps}
            other.close()
    layout = ctx.two_level_layout() if kind != 0 else None

    # dominant kernel: SpMV of the PCG on the last assembled system, timed with HIP events
    spmv_ms, _ = ctx.spmv_timing(400)
    nbytes = spmv_bytes(arrays.R, arrays.B)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": SPMV_TRAFFIC.get(args.columns),
                "kernel": "k_pcg_spmv", "avg_kernel_us": round(1000.0 * spmv_ms, 3),
                "algorithmic_bytes": nbytes, "note": "n129 working set (~10 MB) is cache-resident"}
    line = {
        "metric": "ms/Newton step + PCG iters at n129: learned 2-level vs block-Jacobi PCG",
        "value": round(ms_per_newton, 4),
        "unit": "ms/Newton iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * res.seconds / max(res.accepted_steps, 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"quasi-static sphere drape by PTC, n{args.columns} "
                               f"({arrays.V} vertices, {arrays.T} triangles), isotropic E=1e6 nu=0.3 t=0.5mm, "
                               "sphere r=0.15 m, dhat=1 mm",
                   "preconditioner": args.preconditioner, "self_contact": args.self_contact,
                   "pcg_per_newton": round(pcg_per_newton, 2),
                   "newton_iterations": res.newton_iterations, "pcg_iterations": res.pcg_iterations,
                   "accepted_steps": res.accepted_steps, "phase_ms": {k: round(v, 2) for k, v in zip(D.PHASES, res.phase_ms)},
                   "fallbacks": res.fallbacks, "wall_s": round(wall, 3),
                   "coarse_rows": layout[0] if layout else None,
                   "factorisations": layout[3][0] if layout else None,
                   "factorisation_ms": round(layout[3][3] / 1000.0, 2) if layout else None,
                   "two_level_learned": comparisons.get("learned"),
                   "two_level_bilinear": comparisons.get("two_level"),
                   "block_jacobi": comparisons.get("block_jacobi")},
        "roofline": roofline,
    }
    if args.roofline_columns > 0:
        line["roofline_large_grid"] = large_grid_spmv(args.roofline_columns)
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(args.columns)
        except Exception as e:  # the baseline is reported, never required
            line["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
