/* C entry points of the oracle library (liboracle.so), for the Python test-suite and the CPU
 * baseline leg of the benchmark. Test infrastructure only: the product library never links it.
 * Positions are vertex-major doubles [3V]; rows follow the drape's free-vertex order. */
#ifndef DRAPE_ORACLE_CAPI_H
#define DRAPE_ORACLE_CAPI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t columns;
  double sheet_size, rest_factor;
  double wale, course, coupling, shear, bend_wale, bend_course, density;
  double gravity, time_step;
  int32_t pinned;
  int32_t n_colliders;
  double colliders[4 * 8]; /* kind, a, b, c, w, 0, 0, 0 per collider (contact.rs:57-62) */
  double dhat, kappa, lift;
  int32_t exact_tangent;
  double jitter, flip_fraction;
  uint64_t mesh_seed;
} or_scene;

typedef struct or_drape or_drape;

/* Scene::PinnedCorners / Scene::SphereOnFloor with SceneParams (drape.rs:53-155). */
void or_reference_scene(int32_t columns, int32_t sphere_on_floor, double material_factor, double radius_factor,
                        double rest_factor, int32_t placed, or_scene* out);
or_drape* or_drape_create(const or_scene* scene);
void or_drape_destroy(or_drape* d);
/* sizes[0..6] = V, T, H, R, B, block contributions, force contributions */
void or_drape_sizes(const or_drape* d, int64_t* sizes);
void or_drape_set_time_step(or_drape* d, double h);
double or_drape_kappa(const or_drape* d);
double or_drape_gravity_load(const or_drape* d);
void or_drape_rest_positions(const or_drape* d, double* x);

/* problem.bin arrays (e3_generate.rs:84-104, e7_run.rs:46, e5_checks.rs:80-86). Any pointer may be NULL. */
void or_problem_arrays(const or_drape* d, float* rest_positions, uint32_t* elem_vertices, float* elem_geom,
                       uint32_t* hinge_vertices, float* hinge_stiffness, uint32_t* row_vertex, float* row_mass,
                       float* row_unit_mass, uint32_t* row_ptr, uint32_t* col_idx, uint32_t* diag_idx,
                       uint32_t* block_contrib_ptr, uint32_t* block_contrib_idx, uint32_t* force_contrib_ptr,
                       uint32_t* force_contrib_idx, float* kernel_state, float* colliders, float* contact_state);

/* Element kernels at x: membrane grad [9T] and Hessian [81T] (exact = tangent, else projected);
 * hinge grad [12H] and Hessian [144H] (exact or Gauss-Newton); row terms grad [3R] and block [9R]. */
void or_membrane_elements(const or_drape* d, const double* x, int32_t exact, double* grad, double* hess);
void or_hinge_elements(const or_drape* d, const double* x, int32_t exact, double* grad, double* hess);
void or_row_terms(or_drape* d, const double* x, const double* x_hat, double mu, int32_t curvature, double* grad,
                  double* hess);
/* The assembled system at x: rhs = -residual [3R] and vals [9B] of K + mu*M/h^2. */
void or_assemble(or_drape* d, const double* x, const double* x_hat, double mu, int32_t curvature, double* rhs,
                 double* vals);
double or_energy_change(const or_drape* d, const double* x, const double* x_hat, const double* direction, double alpha);
/* alpha_max and its limiting vertex (-1 when no collider limits). */
double or_max_step(const or_drape* d, const double* x, const double* direction, int32_t* vertex);
double or_static_residual(const or_drape* d, const double* x);
double or_static_energy(const or_drape* d, const double* x);

typedef struct {
  int32_t steps, newton_iterations, rejected_attempts, reason; /* 0 tolerance 1 floor 2 cap 3 stalled */
  double final_relative_residual, final_h, static_energy, max_mu, seconds;
} or_ptc_stats;
/* PTC from `start` (NULL: rest positions). Writes the e7_ptc.rs:183 progress lines into `log`. */
void or_ptc_solve(or_drape* d, const double* start, double h_start, int32_t max_steps, char* log, int64_t log_cap,
                  double* final_positions, or_ptc_stats* stats);

/* PTC from `start` (NULL: rest) at most `max_steps` accepted steps (0: the oracle's own cap): per accepted step
 * `rows` [cap][5] = (h, r/L at the step's start state, inner Newton iterations, rejected attempts, max mu); the
 * last state into `final_positions`. Returns the number of accepted steps. */
int32_t or_ptc_steps(or_drape* d, const double* start, double h_start, int32_t max_steps, double* rows, int32_t cap,
                     double* final_positions, or_ptc_stats* stats);
/* The representation floor of the static residual at x (e7_ptc.rs static_floor). */
double or_static_floor(or_drape* d, const double* x);

/* Dynamic trajectory from rest (drape.rs:426-439): per step newton iterations, final residual and the
 * residual history; `history` holds `history_cap` doubles, `offsets` steps+1 entries. */
int32_t or_simulate(or_drape* d, int32_t steps, double tolerance, int32_t max_iterations, int32_t regularized,
                    double floor_factor, int32_t* iterations, double* history, int64_t history_cap, int64_t* offsets,
                    double* final_positions);

/* Sets the barrier stiffness of every collider term (the adaptive update's value at a probe step). */
void or_drape_set_kappa(or_drape* d, double kappa);

/* A trajectory from x0 / v0 (NULL: rest, zero velocity) with the barrier stiffness kappa_start (<= 0: the scene's),
 * keeping the states listed in keep[keep_count] (state k after k steps, 0 the start): positions and velocities
 * [keep_count][3V]; per step the Newton iterations, the stiffness used and the smallest vertex-collider distance
 * after it. adaptive != 0 applies the IPC stiffness update after each step (kappa doubles when that distance
 * is below kappa_distance * dhat and below the previous step's, up to kappa_max_factor times the scene's
 * stiffness), as the library's rules.adaptive_kappa. The scene's stiffness is restored on return. */
int32_t or_simulate_states(or_drape* d, int32_t steps, const double* x0, const double* v0, double kappa_start,
                           double tolerance, int32_t max_iterations, int32_t regularized, double floor_factor,
                           int32_t adaptive, double kappa_distance, double kappa_max_factor, int32_t* iterations,
                           double* kappa, double* min_distance, const int32_t* keep, int32_t keep_count,
                           double* positions, double* velocities);

/* One Newton solve of one PTC attempt at step size h from x (e7_ptc.rs:129-134), for timing. */
int32_t or_ptc_inner_solve(or_drape* d, const double* x, double h, double* out_positions, int32_t* iterations);

/* Seconds of one f64 Newton iteration (mu = 0) of a PTC attempt at step size h from x. */
double or_time_newton_iteration(or_drape* d, const double* x, double h, int32_t* ok);

typedef struct {
  double gradient_error, hessian_error, hinge_energy_change_error, energy_change_error;
} or_derivative_check;
void or_drape_derivative_check(int32_t columns, or_derivative_check* out);

/* E1 linear cases (cases.rs:64-161). */
typedef struct or_case or_case;
or_case* or_case_create(int32_t kind, int32_t columns);
void or_case_destroy(or_case* c);
void or_case_sizes(const or_case* c, int64_t* rows, int64_t* blocks);
void or_case_arrays(const or_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals, double* rhs);

typedef struct {
  int32_t iterations, converged, status;
  double recursive_relative_residual, true_relative_residual, seconds;
} or_pcg_stats;
/* Block-Jacobi PCG (pcg.rs:60-281) on any BSR system; `single` selects the f32 emulation. History
 * holds (iteration, relative) pairs every `history_stride` iterations plus the final one. */
void or_pcg(int64_t rows, int64_t blocks, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
            const double* rhs, double tolerance, int32_t max_iterations, int32_t replace_every, double replace_above,
            int32_t single, int32_t breakdown_check, int32_t history_stride, double* solution, double* history,
            int64_t history_cap, int64_t* history_len, or_pcg_stats* stats);
/* f64 band Cholesky solve (band.rs); returns 0 or -1 when not positive definite. */
int32_t or_band_solve(int64_t rows, int64_t blocks, const int32_t* row_ptr, const int32_t* col_idx,
                      const double* vals, const double* rhs, double* x, int32_t* bandwidth, double* seconds);

#ifdef __cplusplus
}
#endif
#endif
