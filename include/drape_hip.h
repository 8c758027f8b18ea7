/* drape_hip.h -- C ABI of the MI355X (gfx950) implicit Newton step of drape's cloth solver.
 *
 * The reference has no FFI: its GPU runner (engine/gpu, TypeScript over WebGPU) and its Rust oracle talk
 * through problem.json/problem.bin and record files. This ABI exposes the same step behind plain pointers
 * and sizes: mesh, material and state go in; positions and Newton/PCG statistics come out. Each entry
 * point names the reference interface it replaces (file:line relative to the reference repository).
 *
 * Conventions
 *  - Every call returns DRAPE_OK (0) or a negative DRAPE_E_* code; drape_last_error() then describes it.
 *    Numeric outcomes (PCG status 0-3, Newton step status 0-5, PTC rejection 0-4) are data, not errors.
 *  - The caller owns every host array; arrays are copied on the call. A context owns its device memory
 *    and one HIP stream; calls on a context are synchronous at return and a context is not thread-safe.
 *  - Positions are vertex-major float triples [3V]; "rows" are the free vertices (row_vertex order).
 */
#ifndef DRAPE_HIP_H
#define DRAPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRAPE_OK 0
#define DRAPE_E_INVALID (-1) /* malformed problem or argument (problem.ts:28-34,78-86 throw) */
#define DRAPE_E_DEVICE (-2)  /* HIP runtime or kernel failure (webgpu-backend.ts:23-34 throw) */
#define DRAPE_E_STATE (-3)   /* the solve never raised a stop status (solver.ts:161 throw) */

typedef struct drape_ctx drape_ctx;

/* The problem.bin arrays the reference's runner loads (e3_generate.rs:84-104, e7_run.rs:46,
 * e5_checks.rs:80-86; parsed by drape.ts:26-63). The library derives its own row-sorted scatter map
 * and sliced-BSR layout from them, so block_contrib_* / force_contrib_* / diag_idx / dof_mask are not
 * needed (all dofs of a row are free: pinned vertices have no row). */
typedef struct {
  int32_t vertices, triangles, hinges, rows, blocks;
  const float* rest_positions;     /* [3V] */
  const uint32_t* elem_vertices;   /* [3T] */
  const float* elem_geom;          /* [7T] g0x g0y g1x g1y g2x g2y area */
  const uint32_t* hinge_vertices;  /* [4H] opposite1 opposite2 edge0 edge1 */
  const float* hinge_stiffness;    /* [H] */
  const uint32_t* row_vertex;      /* [R] */
  const float* row_unit_mass;      /* [R] vertex mass in kg (M/h^2 is formed for each step size) */
  const uint32_t* row_ptr;         /* [R+1] BSR pattern (bsr.rs:12-29) */
  const uint32_t* col_idx;         /* [B] sorted per row */
  const float* colliders;          /* [4 + 8K]: count, dhat, kappa, drop-curvature, then kind a b c w 0 0 0; or NULL */
  int32_t collider_floats;
  double wale, course, coupling, shear; /* membrane constants, N/m (material.rs:4-26) */
  double gravity;                  /* m/s^2 */
  double time_step;                /* h of M/h^2 for dynamic steps (PTC rewrites it per attempt) */
  int32_t grid_columns;            /* vertices per side of a sheet numbered as a grid (vertex = row * columns +
                                      column; positions may be jittered), 0 otherwise; the two-level
                                      preconditioner needs it */
} drape_problem;

/* The rule constants problem.json carries (e3_generate.rs:195-229, e4_rules.rs:391-396, e7_run.rs:28-33)
 * plus the run switches of RobustSettings (robust-newton.ts:27-34). drape_default_rules() fills the
 * reference's values (E6 rung D / E7 switches). */
typedef struct {
  double newton_tolerance;
  int32_t newton_max_iterations;
  double armijo;
  int32_t max_halvings;
  double pcg_tolerance;
  int32_t pcg_max_iterations;
  int32_t replace_every;
  double replace_above;
  double eta_start, eta_min, eta_max, eta_gamma, eta_safeguard, update_stop, resolution_factor, mu_start, mu_cap;
  double h0, h_max, h_min_factor, growth, shrink, inner_tolerance;
  int32_t inner_max_iterations, max_rejections;
  double equilibrium_f32;
  int32_t stagnation_window;
  double stagnation_factor, gravity_load;
  int32_t floor_stop, forcing, regularize, noise_aware;
  int32_t exact_tangent;  /* 1 exact membrane+hinge tangent, 0 projected / Gauss-Newton */
  int32_t refine_fp64;    /* 1: FP32 PCG inside an FP64 true-residual refinement loop */
  int32_t preconditioner; /* DRAPE_PRECOND_* */
  int32_t coarse_reuse;   /* two-level: 0 refactors the coarse operator for every system; 1 keeps it while mu
                             and the step size are unchanged and PCG needs at most twice the iterations per
                             decade of the first solve after it; 2 keeps it across mu and step-size changes too
                             (refactoring only on such a slowdown or after a fallback); 3 keeps it across mu
                             changes but refactors when the step size changes; 4 is 3 that also keeps the
                             factorisation of the previous step size and returns to it when the step size
                             does (PTC alternates between two step sizes) */
  int32_t fallback;       /* two-level breakdown (p^T A p <= 0): 1 re-solves with block-Jacobi, 0 reports the
                             breakdown to the Newton loop, which raises mu; 2 returns the last iterate before
                             the negative curvature (status 4) when there is one, else reports the breakdown.
                             A coarse operator that is not positive definite always falls back to
                             block-Jacobi. Every such event counts as a fallback */
  double coarse_shift;    /* two-level: the coarse operator's diagonal is scaled by (1 + coarse_shift) before
                             the factorisation, capping the coarse correction of modes whose curvature is
                             below what FP32 resolves of the stiff membrane */
  double smoother;        /* two-level: 0 applies D^-1 + P A_c^-1 P^T (additive); w > 0 the symmetric
                             multiplicative cycle z1 = w D^-1 r, z2 = z1 + P A_c^-1 P^T (r - A z1),
                             z = z2 + w D^-1 (r - A z2), which is positive definite while
                             w lambda_max(D^-1 A) < 2 */
  int32_t coarse_half;    /* two-level: 1 applies the coarse inverse from FP16 storage, B~ = fp16(S^-1 B) with
                             S = diag(B) (the reciprocal factor diagonal), as (S B~)^T (S B~): half the bytes of
                             the FP32 copy and still exactly a Gram form, so symmetric positive definite;
                             0 applies the FP32 B = L^-1; 2 (default) chooses FP16 when the FP32 B and B^T
                             together exceed 128 MiB (3 Nc above about 4,000: n257 and larger) */
  double refresh_rate;    /* two-level, coarse_reuse >= 1: a kept factorisation is refactored at the next system
                             when a solve needs more than refresh_rate x (+4) the iterations per decade of the
                             first solve after the factorisation (default 2; 0 never) */
  int32_t refresh_on_breakdown; /* two-level: 1 (default) also refactors after a breakdown */
  int32_t self_contact;   /* 1 adds the sheet's self-contact: the IPC barrier b(d) of every point-triangle and
                             edge-edge pair closer than dhat (the colliders' dhat and kappa), found by a
                             spatial-hash broad phase, with its Gauss-Newton Hessian b'' grad d grad d^T; the
                             step bound then also keeps every such pair above 0.1 of its distance. No reference
                             counterpart (the reference's contact is vertex-collider only). Default 0 */
  int32_t coarse_async;   /* two-level, coarse_reuse 1-3: 1 runs a refactorisation that a kept factorisation
                             would otherwise wait for on a second stream beside the solve, which keeps the
                             kept one; the new factorisation serves from the next system on. 2 does so only
                             for a refresh after a slowdown or breakdown at an unchanged step size (a new
                             step size is factored before its first solve). 0 keeps every refactorisation
                             synchronous. The first factorisation and any after a failure are always
                             synchronous. Default 0 */
  int32_t block_jacobi_truncate; /* 1: a block-Jacobi solve that meets p^T A p <= 0 returns the last iterate
                             before it (status 4) when that reduced the residual, the rule rules.fallback = 2
                             applies to two-level solves, so both preconditioners can run under one
                             negative-curvature policy; 0 (default) reports the breakdown to the Newton loop,
                             which raises mu (the reference's rule, robust-newton.ts:121-133) */
  int32_t adaptive_kappa; /* 1: the barrier stiffness kappa adapts once per time step of drape_simulate by the
                             update of Incremental Potential Contact (Li et al. 2020, section 4.3): after a step
                             whose smallest vertex-collider distance is below kappa_distance * dhat and smaller
                             than the previous step's, kappa doubles, up to kappa_max_factor times its start
                             value (the reference's resting_kappa, contact.rs:103-107); kappa never decreases.
                             0 (default) keeps the start value (the reference's fixed kappa) */
  double kappa_distance;   /* default 0.1 */
  double kappa_max_factor; /* default 100 */
  int32_t coarse_factor;   /* two-level: the coarse vertices are every coarse_factor-th grid vertex each way
                              ((grid_columns - 1) divisible by it); 0 (default) picks the smallest power of two
                              from 4 up that keeps the coarse grid at most 33 vertices a side (n129: 4, n257: 8) */
  int32_t coarse_split;    /* two-level: 1 (default) orders the coarse unknowns for the factorisation as two
                              domains (the coarse grid rows above and below a middle separator as wide as the
                              operator's coupling) and the separator, so that the domains factor side by side
                              and the inverse factor is zero between them; 0 keeps row order (one band) */
  int32_t mu_carry;        /* 1: a step of the Newton loop starts at the mu its predecessor's first update was
                              solved at, instead of 0 (robust-newton.ts:232), so that a run whose steps keep
                              needing regularisation does not meet the same breakdowns again at every step;
                              0 (default) is the reference's rule. No reference counterpart */
  int32_t coarse_rotations; /* two-level: 1 gives every coarse node three rotation unknowns beside its three
                              translations: slot k of fine row i adds the columns w_ik (e_a x v_ik), a = x, y, z,
                              with the lever v_ik = sigma_ik (x_i - X_I) / H from the positions at the
                              factorisation (X_I the coarse node's vertex, H the coarse grid spacing, sigma 1 or
                              the learned prolongation's rotation scale), so P reproduces every rigid motion of
                              the current shape; 6 unknowns per coarse node. 0 (default): translations only.
                              No reference counterpart */
  double coarse_pivot_floor; /* two-level: a pivot of the coarse Cholesky factorisation at most coarse_pivot_floor
                              times its diagonal entry of A_c (near-dependent or negative-curvature coarse
                              directions) is replaced by that entry (a modified Cholesky factorisation of
                              A_c + E, E >= 0 diagonal), so the coarse operator stays positive definite; 0 keeps
                              the rule that such a pivot fails the factorisation (block-Jacobi fallback).
                              Default 0; 1e-10 with coarse_rotations */
} drape_rules;

#define DRAPE_PRECOND_BLOCK_JACOBI 0
#define DRAPE_PRECOND_TWO_LEVEL 1

/* PCG outcome of one solve (solver.ts:33-43): status 0 running, 1 converged, 2 iteration limit,
 * 3 breakdown (state.wgsl:2-3); 4 (two-level only, rules.fallback = 2) truncated at negative curvature:
 * the solution is the last iterate before it. */
typedef struct {
  int32_t status;
  int32_t iterations;
  double recursive_relative_residual; /* sqrt(rr/rr0) */
  double true_relative_residual;      /* sqrt(|b - Ax|^2 / rr0), f32 arithmetic on the device */
  int32_t chunks;                     /* host readbacks */
  int32_t refinements;                /* FP64 refinement passes (0 without refine_fp64) */
  double fp64_relative_residual;      /* |b - Ax|/|b| in FP64 (refine_fp64 only, else NaN) */
} drape_pcg_stats;

/* One Newton iterate: the 9 RECORD_FIELDS and 16 EXTRA_FIELDS of node/outputs.ts:8-11, in that order. */
typedef struct {
  double step, iteration, relative_residual, pcg_iterations, pcg_status, linear_residual, step_length, halvings,
      status;
  double eta, mu, update_inf, step_inf, retries, resolution, alpha_max, limiting_vertex, min_distance, delta_sum,
      delta_abs, noise_only, slope, slope_exact, norm_squared, norm_squared_exact;
} drape_record;

/* One accepted PTC step (drape-e7.ts:14-23, written as the 7-field ptc rows of run-e7.ts:96). */
typedef struct {
  double step, h, residual, inner, rejected, pcg_iterations, min_distance, curvature_dropped;
} drape_ptc_step;

/* One attempt (drape-e7.ts:26-33): reason 0 accepted, 1 regularization cap, 2 line search failed,
 * 3 non-finite, 4 curvature retry. */
typedef struct {
  double step, attempt, h, reason, record_start, record_end;
} drape_ptc_attempt;

typedef struct {
  int32_t stop;            /* 0 tolerance, 1 stagnation, 2 cap, 3 stalled, 4 time budget (ptc.ts:260) */
  int32_t accepted_steps;
  int32_t newton_iterations; /* accepted updates over all records */
  int64_t pcg_iterations;
  int32_t records, attempts;
  double final_h, final_residual, gravity_load;
  double seconds;          /* wall clock of the solve loop */
  double phase_ms[7];      /* assembly, pcg, line_search, ccd, reduce, io, other (timed-backend.ts:3-53) */
  int64_t fallbacks;       /* two-level solves redone with block-Jacobi: the coarse operator was not positive
                              definite, or a breakdown under rules.fallback = 1 */
  int64_t read_bytes;      /* device-to-host bytes read back by the loop (EquilibriumRun.readBytes) */
  int64_t truncations;     /* solves of either preconditioner truncated at negative curvature (status 4) */
  int64_t breakdowns;      /* solves that reported a breakdown (status 3) to the Newton loop */
} drape_ptc_result;

/* A sheet scene: the reference's Scene / SceneParams and the material and step constants its Rust
 * generator holds as consts (drape.rs:53-155, cases.rs:9-12, material.rs:15-26), made inputs.
 * jitter / flip_fraction > 0 select an irregular mesh (no reference counterpart). */
typedef struct {
  int32_t columns;
  double sheet_size, rest_factor;
  double wale, course, coupling, shear, bend_wale, bend_course, density;
  double gravity, time_step;
  int32_t pinned;            /* pin the two top corners (drape.rs:148) */
  int32_t n_colliders;       /* <= 4 */
  double colliders[4 * 8];   /* kind a b c w 0 0 0: 1 plane (n, offset), 2 sphere (centre, radius) */
  double dhat, kappa, lift;  /* kappa <= 0: resting stiffness of the heaviest vertex (contact.rs:104-107) */
  int32_t exact_tangent;     /* unused by the builder (kept for layout parity with the oracle's scene) */
  double jitter, flip_fraction;
  uint64_t mesh_seed;
} drape_scene;

typedef struct drape_scene_arrays drape_scene_arrays;

/* Builds the problem arrays of a scene (what the reference's generator writes to problem.bin,
 * e3_generate.rs:84-104). The returned problem points into `arrays`; gravity_load = |m g| over rows
 * (e7_ptc.rs:38-40). */
int drape_scene_build(const drape_scene* scene, drape_scene_arrays** arrays);
/* The same arrays with the hinges, pattern, element geometry, masses and the contribution lists built on the
 * device by stable radix sorts (the reference builds them in its Rust generator: mesh.rs:83-107,
 * bsr.rs:12-29, e3_generate.rs:84-161); entry for entry equal to drape_scene_build. */
int drape_scene_build_device(const drape_scene* scene, drape_scene_arrays** arrays);
/* problem.bin's diag_idx [R] and contribution lists of a device-built scene (e3_generate.rs:111-161):
 * block_contrib_ptr [B + 1] / block_contrib_idx [sizes[0]] (local 3x3 blocks in units of 9 floats),
 * force_contrib_ptr [R + 1] / force_contrib_idx [sizes[1]] (local gradients in units of 3 floats). Any
 * output but `sizes` may be NULL (query the sizes first). */
int drape_scene_contributions(const drape_scene_arrays* arrays, int64_t* sizes, uint32_t* diag_idx,
                              uint32_t* block_contrib_ptr, uint32_t* block_contrib_idx, uint32_t* force_contrib_ptr,
                              uint32_t* force_contrib_idx);
const drape_problem* drape_scene_problem(const drape_scene_arrays* arrays, double* gravity_load);
void drape_scene_free(drape_scene_arrays* arrays);

void drape_default_rules(drape_rules* rules);
const char* drape_last_error(void);

/* Uploads the problem and builds the device layout (replaces prepareDrapeGpu, drape-gpu.ts:141-162). */
int drape_create(const drape_problem* problem, const drape_rules* rules, drape_ctx** out);
void drape_destroy(drape_ctx* ctx);

/* Writes positions [3V], x_hat [3V] (NULL keeps it) and mu; `h` > 0 rewrites M/h^2 from the unit masses
 * (drape-e7.ts:64-72 prepare; robust-newton.ts:109-111 writeMu). */
int drape_set_state(drape_ctx* ctx, const float* positions, const float* x_hat, double h, float mu);
int drape_get_positions(drape_ctx* ctx, float* positions);
/* Drops (1) or keeps (0) the b'grad^2 d term of the barrier Hessian (contact-host.ts:232-236). */
int drape_set_barrier_curvature(drape_ctx* ctx, int32_t drop);

/* Evaluate + assemble at the current state and reduce |rhs|^2 (newton-ops.ts:77-84 gpuResidual):
 * norm = f32(sqrt(norm_squared)). */
int drape_residual(drape_ctx* ctx, float* norm, double* norm_squared);
/* One PCG solve of the assembled system from x = 0 (solver.ts:179-191 solveOnce); solution [3R] may be NULL. */
int drape_pcg(drape_ctx* ctx, double tolerance, int32_t max_iterations, drape_pcg_stats* stats, float* solution);
/* g^T d of the assembled rhs and the last solution: slope = -sum(rhs * x) (newton-ops.ts:86-90). */
int drape_slope(drape_ctx* ctx, double* slope);
/* Scatters the last PCG solution to the vertices as the line-search direction (robust-newton.ts:169-175);
 * with `rows` non-NULL the direction is taken from it instead ([3R]). */
int drape_set_direction(drape_ctx* ctx, const float* rows);
/* Sum and absolute sum of the energy-change terms at alpha (newton-ops.ts:92-97), summed in FP64. */
int drape_energy_change(drape_ctx* ctx, float alpha, double* sum, double* abs_sum);
/* alpha_max of the colliders and its limiting vertex (-1 when 1) (contact-host.ts:175-180). */
int drape_max_step(drape_ctx* ctx, float* alpha, int32_t* vertex, float* min_distance);

/* Robust Newton from x_inout (updated in place) with origin x_n and residual scale
 * (robust-newton.ts:231-250 solveStepRobust). Records are appended to `records` (capacity `cap`). */
int drape_newton_step(drape_ctx* ctx, float* x_inout, float scale, int32_t step, const float* origin,
                      drape_record* records, int32_t cap, int32_t* count);

/* Dynamic trajectory by backward Euler (drape.rs:426-439 simulate; time_step with the prediction
 * x_hat = x + h v + h^2 g of drape.rs prediction; one robust Newton solve per step as solveStepRobust,
 * robust-newton.ts:231-250): `steps` steps of problem.time_step from x0 / v0 (NULL: rest, zero). Outputs
 * (any may be NULL): final positions and velocities [3V], per-step Newton updates and final relative
 * residual [steps], the records (capacity `cap`). */
int drape_simulate(drape_ctx* ctx, int32_t steps, const float* x0, const float* v0, float* x_out, float* v_out,
                   int32_t* updates, float* final_relative, drape_record* records, int32_t cap, int32_t* count);

/* Quasi-static drape by pseudo-transient continuation from `start` with first step size h_start
 * (drape-e7.ts:156-181 solveEquilibrium). `cap` accepted steps, `budget_s` > 0 a wall-clock budget.
 * Any output array may be NULL; final positions go to `final_positions` [3V]. */
int drape_ptc_solve(drape_ctx* ctx, const float* start, double h_start, int32_t cap, double budget_s,
                    drape_ptc_result* result, drape_ptc_step* steps, int32_t steps_cap, drape_ptc_attempt* attempts,
                    int32_t attempts_cap, drape_record* records, int32_t records_cap, float* final_positions);
/* States kept by later drape_ptc_solve calls (solveEquilibrium's keepStates, drape-e7.ts:159-178):
 * keep != 0 keeps the start and every accepted state, keep == 0 the start and the last only. */
int drape_ptc_keep_states(drape_ctx* ctx, int32_t keep);
/* The states kept by the last drape_ptc_solve, `count` of them, [count][3V] into `states` (up to `cap`
 * states; NULL to query the count) -- the contents of gpu-states.bin (run-e7.ts:98). */
int drape_ptc_states(drape_ctx* ctx, float* states, int32_t cap, int32_t* count);

/* One PTC attempt from `start` [3V] at step size h, as the reference repeats each oracle probe step
 * (drape-e7.ts:183-196 runPtcProbes): x_hat and M/h^2 of `start`, the residual scale at `start`, the
 * CCD-cut Newton start and one robust Newton solve with the PTC inner settings; no rejection loop. Records
 * (step index `step`) go to `records` (capacity `cap`, count in *count), the positions after each accepted
 * update to `iterates` [cap][3V] (NULL: not kept) and the last iterate to `final_positions` [3V]. */
int drape_ptc_probe(drape_ctx* ctx, const float* start, double h, int32_t step, drape_record* records, int32_t cap,
                    int32_t* count, float* iterates, float* final_positions);

/* The barrier stiffness kappa the kernels use (the colliders' kappa, problem.colliders[2]): drape_set_kappa
 * writes it (kappa > 0), drape_kappa reads it. drape_kappa_history gives, for each step of the last
 * drape_simulate ([count], up to cap; NULL arrays query the count), the kappa the step used (the values
 * rules.adaptive_kappa produced) and the smallest vertex-collider distance after it (FP32, +inf without
 * colliders). The reference's kappa is a constant of the problem (drape.rs:81); these serve the update. */
int drape_set_kappa(drape_ctx* ctx, double kappa);
int drape_kappa(drape_ctx* ctx, double* kappa);
int drape_kappa_history(drape_ctx* ctx, double* kappa, float* min_distance, int32_t cap, int32_t* count);

/* Prolongation of a coarse state to the next grid as a PTC start (e7_prolong.rs:9-47, the oracle's
 * e7-prolong): `coarse` [3 coarse_columns^2] f32 -> `fine` [3 (2 coarse_columns - 1)^2] f32, computed in f64.
 * `colliders` [collider_doubles = 4 + 8K] = [K, dhat, kappa, flag, (kind, a, b, c, w, 0, 0, 0) x K] in f64
 * (NULL: no projection); `moved` counts vertices projected to d = dhat, `infeasible` those left at or inside a
 * collider (the caller refuses such a start, e7_prolong.rs:60-62). Context-free; coarse_columns <= 8193. */
int drape_prolong_state(int32_t coarse_columns, const float* coarse, const double* colliders,
                        int32_t collider_doubles, float* fine, int32_t* moved, int32_t* infeasible);

/* drape_prolong_state without any allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * `coarse` [3 coarse_columns^2] and `fine` [3 (2 coarse_columns - 1)^2] are DEVICE arrays of the current device
 * (checked; anything else is DRAPE_E_INVALID); `colliders` is a host array as in drape_prolong_state, with at
 * most 16 colliders. Work goes on `stream` (a hipStream_t; NULL: the default stream) and the call returns once
 * moved / infeasible are on the host. The only device memory it uses beyond the caller's arrays is a 9.5 KB
 * array of the library (collider staging and per-workgroup counts), one per GPU for the life of the process:
 * calls that overlap on one GPU race. Same positions, bit for bit, and counts as drape_prolong_state. */
int drape_prolong_state_noalloc(int32_t coarse_columns, const float* coarse, const double* colliders,
                                int32_t collider_doubles, float* fine, int32_t* moved, int32_t* infeasible,
                                void* stream);

/* The context's assembly scatter maps, built on the device at drape_create: problem.bin's
 * block_contrib_ptr [B + 1] / block_contrib_idx [sizes[0]] and force_contrib_ptr [R + 1] /
 * force_contrib_idx [sizes[1]] (e3_generate.rs:111-161). Any output but `sizes` may be NULL. */
int drape_get_contributions(drape_ctx* ctx, int64_t* sizes, uint32_t* block_contrib_ptr, uint32_t* block_contrib_idx,
                            uint32_t* force_contrib_ptr, uint32_t* force_contrib_idx);

/* Kernel-check switches (drape-e4.ts:108-120, drape-e7.ts:203-217, contact-host.ts:64-90):
 * drape_set_colliders(0) zeroes the collider count the kernels read (withoutColliders), 1 restores it;
 * drape_set_row_mass writes the row masses M/h^2 [R] directly (gpu.rowMass); drape_set_tangent selects the
 * exact (1) or projected (0) element tangent of later evaluations (exactElements / projectedElements). */
int drape_set_colliders(drape_ctx* ctx, int32_t enabled);
int drape_set_row_mass(drape_ctx* ctx, const float* row_mass);
int drape_set_tangent(drape_ctx* ctx, int32_t exact);
/* Name, architecture and compute units of the current device (the adapter of run-e7.ts:101). */
int drape_device_name(char* name, int32_t cap);

/* Kernel check (drape-step.ts evaluateDrapeKernels / drape-e4.ts:376-388): element and row terms at
 * `positions` with the current x_hat, mu and masses: grad [9T + 12H + 3R], hess [81T + 144H + 9R]
 * in the reference's local layout (block (k,l) of element e at 81e + 9(3k+l), row-major). */
int drape_eval_elements(drape_ctx* ctx, const float* positions, float* grad, float* hess);
/* The assembled system: rhs [3R] and vals [9B] in CSR block order. */
int drape_get_system(drape_ctx* ctx, float* rhs, float* vals);

/* A context over a plain BSR system (the E1 cases, generate.rs:330-339): drape_pcg then solves it. */
int drape_linear_create(int32_t rows, int32_t blocks, const uint32_t* row_ptr, const uint32_t* col_idx,
                        const float* vals, const float* rhs, const drape_rules* rules, drape_ctx** out);

/* Selects the preconditioner of later solves (DRAPE_PRECOND_*). DRAPE_PRECOND_TWO_LEVEL adds a Galerkin
 * coarse correction on every f-th grid vertex (f = rules.coarse_factor) to block-Jacobi; it needs
 * problem.grid_columns with (grid_columns - 1) divisible by f. `weights` [4R] are the prolongation weights of the four parent slots
 * of each row (NULL: bilinear); slots without a parent ignore theirs. No reference counterpart: the
 * reference's solver is block-Jacobi only (solver.ts:109-191). */
int drape_set_preconditioner(drape_ctx* ctx, int32_t kind, const float* weights);
/* The coarse layout of the two-level preconditioner: coarse rows, parent [4R] (coarse row of each parent
 * slot, -1 if none) and the bilinear weights [4R]; counters [4] = factorisations, failed factorisations
 * (A_c not positive definite), two-level solves redone with block-Jacobi, microseconds spent factorising
 * (device time of the factorisation chains, side-stream chains included).
 * Any output may be NULL. */
int drape_two_level_layout(drape_ctx* ctx, int32_t* coarse_rows, int32_t* parent, float* bilinear,
                           int64_t* counters);

/* The learned prolongation: `count` parameters of the graph network (13 input features, width 32, three
 * message-passing layers, 4 parent-slot outputs: W0 [32x13], b0 [32], then per layer Ws [32x32], Wn [32x32],
 * b [32], then Wo [4x32], bo [4]). From the next factorisation on, the two-level preconditioner takes its
 * weights from the network evaluated on each assembled system; NULL returns to the bilinear weights.
 * No reference counterpart (the learned two-level preconditioner is new). */
int drape_set_prolongation_model(drape_ctx* ctx, const float* params, int32_t count);
/* The parent-slot weights [4R] of the current coarse factorisation. */
int drape_two_level_weights(drape_ctx* ctx, float* weights);

/* Training of the learned prolongation on the device (no reference counterpart; the reference has no
 * two-level preconditioner). The error vectors the loss is taken over, kept on the device:
 * drape_training_vectors_set copies `count` vectors [count][3R] (row-major per vector);
 * drape_training_vectors_make makes them from uniform random starts in [-1, 1) (seeded) by `sweeps` damped
 * block-Jacobi sweeps e <- e - omega D^-1 A e on the assembled system, and copies them to `out`
 * [count][3R] unless it is NULL. */
int drape_training_vectors_set(drape_ctx* ctx, const float* vectors, int32_t count);
int drape_training_vectors_make(drape_ctx* ctx, int32_t count, int32_t sweeps, float omega, uint64_t seed,
                                float* out);
/* The projection loss L = mean_e |e - P c_e|_A^2 / |e|_A^2 (c_e = A_c^-1 P^T A e, A_c = P^T A P, P the
 * prolongation the network at `params` gives on the assembled system) into *loss and dL/dparams [count]
 * into `grad`, by the network's forward and backward kernels; the coarse solve is the preconditioner's own
 * factorisation. `params` becomes the context's prolongation model (as drape_set_prolongation_model).
 * DRAPE_E_INVALID when no vectors are set, `count` is not the model's or A_c is not positive definite. */
int drape_prolongation_gradient(drape_ctx* ctx, const float* params, int32_t count, double* loss, float* grad);

/* The self-contact pairs of the last evaluation under rules.self_contact (no reference counterpart): the
 * numbers of point-triangle and edge-edge pairs closer than dhat at the evaluated positions, their vertices
 * [4 (n_pt + n_ee)] (the point, then the triangle's three; or the first edge's two, then the second's) and
 * their FP64 distances. Any output may be NULL (a first call with NULL arrays gives the counts). */
int drape_self_contact_pairs(drape_ctx* ctx, int32_t* n_pt, int32_t* n_ee, int32_t* vertices, double* distances);

/* The rotation unknowns of the coarse space (rules.coarse_rotations): *dof = unknowns per coarse node (3, or 6
 * with rotations), `levers` [4R x 4] the rotation lever (x, y, z, 0) of every parent slot of the current
 * factorisation (zero without rotations), *replaced the pivots the pivot floor (rules.coarse_pivot_floor) has
 * replaced over all factorisations. Any output may be NULL. No reference counterpart (the two-level
 * preconditioner is new). */
int drape_two_level_rotations(drape_ctx* ctx, int32_t* dof, float* levers, int64_t* replaced);

/* The Galerkin coarse operator A_c = P^T A P of the assembled system, row-major [dof Nc x dof Nc] doubles (Nc
 * from drape_two_level_layout, dof from drape_two_level_rotations; with rotations, P's levers are formed from
 * the current weights and positions), for checking; the preconditioner's factorisation is left as it is. */
int drape_two_level_coarse(drape_ctx* ctx, double* coarse);

/* B = L^-1 (A_c = L L^T) of the current coarse factorisation as FP32, row-major [dof Nc x dof Nc] (zero above the
 * diagonal), in the factor's ordering of the coarse unknowns (drape_two_level_ordering), for checking; the
 * coarse solve applies B^T B. No reference counterpart (the two-level preconditioner is new). */
int drape_two_level_inverse_factor(drape_ctx* ctx, float* factor);

/* The factor's ordering of the coarse unknowns: order[i] = dof I + q (coarse node I, component q) at position i
 * of drape_two_level_inverse_factor ([dof Nc]), so that B^T B = (Q A_c Q^T)^-1 with Q the permutation that picks
 * the rows order[0], order[1], ... No reference counterpart. */
int drape_two_level_ordering(drape_ctx* ctx, int32_t* order);

/* The FP64 iterate [3R] of the last drape_pcg under rules.refine_fp64 (drape_pcg's FP32 solution rounds it).
 * No reference counterpart (FP64 refinement is new). */
int drape_solution_fp64(drape_ctx* ctx, double* solution);

/* Device timing of the last drape_pcg (HIP events on the context's stream): average SpMV kernel time
 * and the whole solve, in milliseconds. */
int drape_pcg_timing(drape_ctx* ctx, double* spmv_ms, double* solve_ms, int32_t* spmv_launches);

/* Device timing of one kernel of the PCG iteration on the last assembled system (HIP events on the context's
 * stream, `launches` back-to-back launches): kernel 0 the plain SpMV (k_pcg_spmv), 1 the direction update fused
 * into the SpMV (k_pcg_dir_spmv), 2 / 3 the lower / upper triangular GEMV of the two-level coarse solve
 * (k_tri_gemv, current factorisation), 4 the fused coarse solve y_c = B^T B r_c (k_gram and k_gram_sum, one pass
 * over B; the layouts whose coarse solve runs as the two GEMVs reject it), 5 one whole coarse factorisation of
 * the current system (Galerkin product, Cholesky chain and inverse sweep; it is formed again from the same
 * inputs). Average ms per launch and the kernel's algorithmic bytes per launch (kernel 5: the FP64 flops of the
 * Cholesky factorisation and the inverse sweep of the tile algorithm).
 * No reference counterpart. */
int drape_kernel_timing(drape_ctx* ctx, int32_t kernel, int32_t launches, double* avg_ms, double* bytes);

/* A plain device-to-device copy of `bytes` (16-byte units), `launches` times between HIP events on a stream of
 * its own: the average kernel time in ms. No reference counterpart: the known-size kernel by which hardware
 * byte counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) are scaled. */
int drape_copy_timing(int64_t bytes, int32_t launches, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* DRAPE_HIP_H */
