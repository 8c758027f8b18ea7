"""Benchmark of the implicit Newton step at n129: ms per Newton iteration and PCG iterations per Newton
iteration of the quasi-static sphere drape solved by pseudo-transient continuation (129 x 129 grid,
16641 vertices, 32768 triangles; isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm, 0.2 kg/m^2, hinge 1e-5 N m,
sphere r = 0.15 m, dhat = 1 mm), the learned two-level preconditioner against block-Jacobi FP32 PCG.

A bench step is one accepted PTC step (one inexact Newton solve of the step's energy, several Newton
iterations). `--warmup W` steps (learned two-level) run untimed from the flat start; from the state and step
size they reach, every leg runs the next `--steps K` steps: the selected preconditioner (the reported value;
by default the two-level preconditioner with the learned prolongation shipped with the package), the bilinear
two-level preconditioner and block-Jacobi, all under one negative-curvature policy (`--policy`), and then the
same three under the other policy. Each leg reports its stop reason, its final and least r/L (static residual
over the gravity load) and its truncation, breakdown and block-Jacobi re-solve counts. With the defaults the
window is PTC steps 6-25, which end near the FP32 representation floor of the n129 drape (r/L about 0.46).
Prints one JSON line.

`roofline` is the fused direction update + SpMV (`k_pcg_dir_spmv`), the most-launched SpMV of the measured
run; `rooflines` adds the plain SpMV, the coarse solve (the fused one-pass `k_gram` + `k_gram_sum`, or the two
triangular GEMVs where the layout keeps them) and the whole coarse factorisation. Each is timed live with HIP
events on the context's stream over back-to-back launches on the last assembled system, against its algorithmic
bytes per launch (SpMV 40 B per block + 28 B per row + 4; the fused kernel 24 B per row more; the coarse solve
the stored triangle of B in FP32 plus 8 B per coarse unknown) or, for the factorisation, the FP64 flops of its
tile Cholesky factorisation and inverse sweep against the 78.6 TFLOP/s FP64 matrix peak. `roofline_large_grid` is the plain SpMV at n1025,
whose matrix exceeds the Infinity Cache. The CPU baseline is the oracle (f64 Newton with band Cholesky, as
the reference's Rust oracle does) timed for one Newton iteration of the same first PTC step on one core.
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
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
# HBM bytes per launch, from rocprofv3 FETCH_SIZE and WRITE_SIZE counters scaled by the same counters of a
# known-size copy (FETCH_SIZE x2.0, WRITE_SIZE x1.0 on this part), keyed by (kernel, grid columns, coarse
# unknowns; 0 for the SpMVs); the summaries are under profiles/. (At n129 the bytes are L2 misses served mostly
# by the Infinity Cache.)
TRAFFIC = {("k_pcg_spmv", 1025, 0): 523_831_470, ("k_pcg_spmv", 129, 0): 8_638_625,
           ("k_pcg_dir_spmv", 129, 0): 9_762_528,
           ("k_tri_gemv_lower", 129, 3267): 12_372_435, ("k_tri_gemv_upper", 129, 3267): 13_493_029,
           ("k_tri_gemv_lower", 129, 6534): 48_591_794, ("k_tri_gemv_upper", 129, 6534): 50_698_324,
           ("k_gram", 129, 6534): 58_299_872}


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
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get(("k_pcg_spmv", columns, 0)), "kernel": "k_pcg_spmv",
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


LEG_NAMES = {"learned": "two-level, learned prolongation", "two_level": "two-level, bilinear prolongation",
             "rotations": "two-level, bilinear prolongation with rotation unknowns", "block_jacobi": "block-Jacobi"}
# the coarse space with three rotation unknowns per node beside the translations, its Cholesky factorisation
# with the pivot floor
ROTATIONS = {"coarse_rotations": 1, "coarse_pivot_floor": 1e-10}
# negative-curvature policies applied to every leg alike: "truncate" returns the last iterate before
# p^T A p <= 0 when it reduced the residual (two-level fallback = 2, block_jacobi_truncate = 1) and otherwise
# reports the breakdown; "regularize" always reports it (fallback = 0, block_jacobi_truncate = 0), the
# reference's rule. Either way the Newton loop raises mu on a reported breakdown.
# Negative-curvature policies, applied to every leg alike. "truncate": a solve that meets p^T A p <= 0 returns
# its last iterate when that reduced the residual, and a step starts at the mu its predecessor's first update
# needed (rules.mu_carry); "regularize": the reference's rule, every breakdown raises mu from 0 at each step.
POLICIES = {"truncate": {"fallback": 2, "block_jacobi_truncate": 1, "mu_carry": 1},
            "regularize": {"fallback": 0, "block_jacobi_truncate": 0, "mu_carry": 0}}


def kernel_roofline(ctx, kernel: str, launches: int, note: str) -> dict:
    ms, nbytes = ctx.kernel_timing(kernel, launches)
    achieved = nbytes / (ms * 1e-3) / 1e9
    unknowns = ctx.coarse_unknowns() if kernel.startswith(("k_tri_gemv", "k_gram")) else 0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get((kernel, ctx.columns, unknowns)),
            "kernel": kernel, "avg_kernel_us": round(1000.0 * ms, 3), "algorithmic_bytes": int(nbytes),
            "launches": launches, "note": note}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--columns", type=int, default=129)
    ap.add_argument("--preconditioner", choices=list(LEG_NAMES), default="learned",
                    help="learned: two-level with the learned prolongation; two_level: with bilinear prolongation")
    ap.add_argument("--policy", choices=list(POLICIES), default="truncate",
                    help="negative-curvature policy of every leg")
    ap.add_argument("--no-compare", action="store_true", help="skip the comparison legs")
    ap.add_argument("--roofline-columns", type=int, default=1025,
                    help="grid for the HBM-bound SpMV roofline (0 skips it)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the large-grid SpMV and a copy of known size (for PMC byte counters)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the n129 kernel timings after the warm-up (for PMC byte counters)")
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

    def context(pre, policy):
        if pre == "learned":
            from drape_amd.learned import model_rules
            extra = model_rules()
        else:
            extra = ROTATIONS if pre == "rotations" else {}
        c = D.Drape(arrays, D.default_rules(preconditioner=0 if pre == "block_jacobi" else 1,
                                            self_contact=1 if args.self_contact else 0, **POLICIES[policy], **extra))
        if pre == "learned":
            from drape_amd.learned import load_model
            c.set_prolongation_model(load_model())
        c.columns = args.columns
        return c

    print(f"[bench] n{args.columns}: V {arrays.V} T {arrays.T} H {arrays.H} rows {arrays.R} blocks {arrays.B}",
          file=sys.stderr, flush=True)

    # the warm-up steps (learned two-level) give every leg the same start state and step size
    start = arrays.array("rest_positions")
    warm = context("learned", args.policy)
    h = warm.rules.h0
    if args.warmup > 0:
        w = warm.solve_equilibrium(start=start, h_start=h, cap=args.warmup)
        start, h = w["positions"], w["result"].final_h
        print(f"[bench] warmup: {w['result'].accepted_steps} steps, stop {w['stop']}", file=sys.stderr, flush=True)
    warm.close()

    def leg(pre, policy):
        c = context(pre, policy)
        t0 = time.perf_counter()
        out = c.solve_equilibrium(start=start, h_start=h, cap=args.steps)
        wall = time.perf_counter() - t0
        res = out["result"]
        newton = max(res.newton_iterations, 1)
        load = res.gravity_load
        rl = [s.residual / load for s in out["steps"]] + [res.final_residual / load]
        info = {"preconditioner": LEG_NAMES[pre], "policy": policy,
                "ms_per_newton": round(1000.0 * res.seconds / newton, 4),
                "pcg_per_newton": round(res.pcg_iterations / newton, 2),
                "newton_iterations": res.newton_iterations, "pcg_iterations": res.pcg_iterations,
                "accepted_steps": res.accepted_steps, "seconds": round(res.seconds, 4), "wall_s": round(wall, 3),
                "stop": out["stop"], "final_relative_residual": round(rl[-1], 5),
                "min_relative_residual": round(min(rl), 5),
                "truncations": res.truncations, "breakdowns": res.breakdowns,
                "block_jacobi_resolves": res.fallbacks,
                "phase_ms": {k: round(v, 2) for k, v in zip(D.PHASES, res.phase_ms)}}
        if pre != "block_jacobi":
            layout = c.two_level_layout()
            info["factorisations"] = int(layout[3][0])
            info["factorisation_ms"] = round(layout[3][3] / 1000.0, 2)
            info["factorisation_share"] = round(layout[3][3] / 1000.0 / (1000.0 * res.seconds), 4)
        print(f"[bench] {policy} {pre}: {res.accepted_steps} steps, {res.newton_iterations} Newton, "
              f"{res.pcg_iterations} PCG, {res.seconds:.3f} s, truncations {res.truncations}, breakdowns "
              f"{res.breakdowns}, stop {out['stop']}, r/L final {rl[-1]:.3e} min {min(rl):.3e}",
              file=sys.stderr, flush=True)
        return c, info

    if args.kernels_only:
        ctx, _ = leg(args.preconditioner, args.policy)
        for k in ("k_pcg_spmv", "k_pcg_dir_spmv", "k_tri_gemv_lower", "k_tri_gemv_upper", "k_gram"):
            try:
                ms, nbytes = ctx.kernel_timing(k, 400)
            except D.DrapeError as e:
                print(json.dumps({"kernel": k, "skipped": str(e)}), flush=True)
                continue
            print(json.dumps({"kernel": k, "avg_kernel_us": round(1000 * ms, 3), "algorithmic_bytes": nbytes}),
                  flush=True)
        print(json.dumps({"copy": copy_calibration()}), flush=True)
        return

    ctx, main_leg = leg(args.preconditioner, args.policy)
    legs = {args.policy: {args.preconditioner: main_leg}}
    if not args.no_compare:
        for policy in [args.policy] + [p for p in POLICIES if p != args.policy]:
            for pre in LEG_NAMES:
                if policy == args.policy and pre == args.preconditioner:
                    continue
                other, info = leg(pre, policy)
                other.close()
                legs.setdefault(policy, {})[pre] = info
    same = legs[args.policy]
    ratios = None
    if "block_jacobi" in same and args.preconditioner != "block_jacobi":
        bj = same["block_jacobi"]
        ratios = {"time_per_newton": round(bj["ms_per_newton"] / main_leg["ms_per_newton"], 2),
                  "pcg_per_newton": round(bj["pcg_per_newton"] / main_leg["pcg_per_newton"], 2)}

    # rooflines of the PCG iteration's kernels on the last assembled system of the measured leg, timed with
    # HIP events on the context's stream
    if args.preconditioner != "block_jacobi":
        n = ctx.coarse_unknowns()
        gemv_mb = ctx.kernel_timing("k_tri_gemv_lower", 1)[1] / 1e6   # the stored triangle of B, one GEMV's bytes
        try:
            ctx.kernel_timing("k_gram", 1)
            fused = True
        except D.DrapeError:
            fused = False
        cache = (f"n{args.columns} working set (matrix 7.7 MB, coarse inverse factor B over {n} coarse unknowns, "
                 f"{gemv_mb:.1f} MB in the split ordering{'' if fused else ', and B^T'}) is resident in the 256 MB "
                 "Infinity Cache")
    else:
        cache = f"n{args.columns} working set (matrix 7.7 MB) is resident in the 256 MB Infinity Cache"
    kernels = [kernel_roofline(ctx, "k_pcg_dir_spmv", 400, "direction update fused into the SpMV, every "
                               "iteration after a chunk's first; " + cache),
               kernel_roofline(ctx, "k_pcg_spmv", 400, "plain SpMV, the first iteration of a chunk; " + cache)]
    if args.preconditioner != "block_jacobi" and fused:
        kernels += [kernel_roofline(ctx, "k_gram", 400, "fused coarse solve y_c = B^T (B r_c), one pass over B "
                                    "(k_gram and k_gram_sum, timed together); " + cache)]
    elif args.preconditioner != "block_jacobi":
        kernels += [kernel_roofline(ctx, "k_tri_gemv_lower", 400, "coarse solve t = B r_c; " + cache),
                    kernel_roofline(ctx, "k_tri_gemv_upper", 400, "coarse solve y_c = B^T t; " + cache)]
    if args.preconditioner != "block_jacobi":
        # the coarse factorisation, the other part of the measured time: FP64 MFMA work, latency-bound
        ms, flops = ctx.kernel_timing("coarse_factorisation", 10)
        achieved = flops / (ms * 1e-3) / 1e12
        kernels += [{"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "coarse factorisation (k_galerkin_band, the k_chol_step chain, k_inverse_row and "
                               "k_separator_border on the inverse stream)",
                     "avg_kernel_us": round(1000.0 * ms, 1), "algorithmic_flops": flops, "launches": 10,
                     "note": f"one factorisation of the last assembled system over {n} coarse unknowns, HIP events "
                             "around the whole of it; FP64 flops of the tile Cholesky factorisation and the inverse "
                             "sweep (the Galerkin product's terms not counted); its 59 dependent chain steps make it "
                             "latency-bound"}]
    ctx.close()
    line = {
        "metric": "ms/Newton step + PCG iters at n129: learned 2-level vs block-Jacobi PCG",
        "value": main_leg["ms_per_newton"],
        "unit": "ms/Newton iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * main_leg["seconds"] / max(main_leg["accepted_steps"], 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"quasi-static sphere drape by PTC, n{args.columns} "
                               f"({arrays.V} vertices, {arrays.T} triangles), isotropic E=1e6 nu=0.3 t=0.5mm, "
                               "sphere r=0.15 m, dhat=1 mm; PTC steps warmup+1..warmup+steps from the flat start, "
                               "every leg from the same state and step size",
                   "preconditioner": args.preconditioner, "policy": args.policy, "self_contact": args.self_contact,
                   "pcg_per_newton": main_leg["pcg_per_newton"],
                   "block_jacobi_over_measured": ratios, "legs": legs},
        "roofline": kernels[0],
        "rooflines": kernels[1:],
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
