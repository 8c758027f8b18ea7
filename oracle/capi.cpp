// C entry points of the oracle (test infrastructure only).
#include <chrono>
#include <cstring>

#include "oracle.hpp"
#include "oracle_capi.h"

using namespace oracle;

struct or_drape {
  Drape drape;
  explicit or_drape(const SceneSpec& s) : drape(s) {}
};
struct or_case {
  CaseProblem p;
};

static SceneSpec spec_of(const or_scene* s) {
  SceneSpec o;
  o.columns = s->columns;
  o.sheet_size = s->sheet_size;
  o.rest_factor = s->rest_factor;
  o.material = {s->wale, s->course, s->coupling, s->shear, s->bend_wale, s->bend_course, s->density};
  o.gravity = s->gravity;
  o.time_step = s->time_step;
  o.pinned = s->pinned != 0;
  for (int k = 0; k < s->n_colliders; ++k) {
    const double* c = &s->colliders[8 * k];
    o.colliders.push_back({(int)c[0], {c[1], c[2], c[3]}, c[4]});
  }
  o.dhat = s->dhat;
  o.kappa = s->kappa;
  o.lift = s->lift;
  o.exact_tangent = s->exact_tangent != 0;
  o.jitter = s->jitter;
  o.flip_fraction = s->flip_fraction;
  o.mesh_seed = s->mesh_seed;
  return o;
}

static std::vector<Vec3> vec3s(const double* x, size_t n) {
  std::vector<Vec3> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  return v;
}

template <class T> static void copy_out(const std::vector<T>& v, T* out) {
  if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(T));
}

extern "C" {

void or_reference_scene(int32_t columns, int32_t sphere, double mf, double rf, double restf, int32_t placed,
                        or_scene* out) {
  SceneSpec s = reference_scene(columns, sphere != 0, mf, rf, restf, placed != 0);
  std::memset(out, 0, sizeof *out);
  out->columns = s.columns;
  out->sheet_size = s.sheet_size;
  out->rest_factor = s.rest_factor;
  out->wale = s.material.wale;
  out->course = s.material.course;
  out->coupling = s.material.coupling;
  out->shear = s.material.shear;
  out->bend_wale = s.material.bend_wale;
  out->bend_course = s.material.bend_course;
  out->density = s.material.density;
  out->gravity = s.gravity;
  out->time_step = s.time_step;
  out->pinned = s.pinned;
  out->n_colliders = (int32_t)s.colliders.size();
  for (size_t k = 0; k < s.colliders.size(); ++k) {
    auto p = s.colliders[k].packed();
    for (int j = 0; j < 8; ++j) out->colliders[8 * k + j] = p[j];
  }
  out->dhat = s.dhat;
  out->kappa = s.kappa;
  out->lift = s.lift;
  out->exact_tangent = 1;
}

or_drape* or_drape_create(const or_scene* scene) { return new or_drape(spec_of(scene)); }
void or_drape_destroy(or_drape* d) { delete d; }

void or_drape_sizes(const or_drape* d, int64_t* sizes) {
  const Drape& dr = d->drape;
  sizes[0] = (int64_t)dr.mesh.rest.size();
  sizes[1] = (int64_t)dr.membranes.size();
  sizes[2] = (int64_t)dr.hinges.size();
  sizes[3] = dr.rows();
  sizes[4] = dr.pattern.blocks();
  // contribution list lengths (e3_generate.rs:118-152)
  int64_t blocks = dr.rows(), forces = dr.rows();
  for (const auto& e : dr.membranes)
    for (int k = 0; k < 3; ++k) {
      if (dr.row_of[e.vertices[k]] < 0) continue;
      ++forces;
      for (int l = 0; l < 3; ++l) blocks += dr.row_of[e.vertices[l]] >= 0;
    }
  for (const auto& h : dr.hinges)
    for (int k = 0; k < 4; ++k) {
      if (dr.row_of[h.vertices[k]] < 0) continue;
      ++forces;
      for (int l = 0; l < 4; ++l) blocks += dr.row_of[h.vertices[l]] >= 0;
    }
  sizes[5] = blocks;
  sizes[6] = forces;
}

void or_drape_set_time_step(or_drape* d, double h) { d->drape.time_step = h; }
double or_drape_kappa(const or_drape* d) { return d->drape.contact.barrier.kappa; }
double or_drape_gravity_load(const or_drape* d) { return gravity_load(d->drape); }
void or_drape_rest_positions(const or_drape* d, double* x) {
  auto r = d->drape.rest_positions();
  for (size_t i = 0; i < r.size(); ++i)
    for (int k = 0; k < 3; ++k) x[3 * i + k] = r[i][k];
}


void or_problem_arrays(const or_drape* d, float* rest_positions, uint32_t* elem_vertices, float* elem_geom,
                       uint32_t* hinge_vertices, float* hinge_stiffness, uint32_t* row_vertex, float* row_mass,
                       float* row_unit_mass, uint32_t* row_ptr, uint32_t* col_idx, uint32_t* diag_idx,
                       uint32_t* bcp, uint32_t* bci, uint32_t* fcp, uint32_t* fci, float* kernel_state, float* colliders,
                       float* contact_state) {
  ProblemArrays a = problem_arrays(d->drape);
  copy_out(a.rest_positions, rest_positions);
  copy_out(a.elem_vertices, elem_vertices);
  copy_out(a.elem_geom, elem_geom);
  copy_out(a.hinge_vertices, hinge_vertices);
  copy_out(a.hinge_stiffness, hinge_stiffness);
  copy_out(a.row_vertex, row_vertex);
  copy_out(a.row_mass, row_mass);
  copy_out(a.row_unit_mass, row_unit_mass);
  copy_out(a.row_ptr, row_ptr);
  copy_out(a.col_idx, col_idx);
  copy_out(a.diag_idx, diag_idx);
  copy_out(a.block_contrib_ptr, bcp);
  copy_out(a.block_contrib_idx, bci);
  copy_out(a.force_contrib_ptr, fcp);
  copy_out(a.force_contrib_idx, fci);
  copy_out(a.kernel_state, kernel_state);
  copy_out(a.colliders, colliders);
  copy_out(a.contact_state, contact_state);
}

void or_membrane_elements(const or_drape* d, const double* xp, int32_t exact, double* grad, double* hess) {
  const Drape& dr = d->drape;
  auto x = vec3s(xp, dr.mesh.rest.size());
  for (size_t e = 0; e < dr.membranes.size(); ++e) {
    const auto& el = dr.membranes[e];
    auto g = el.gradient(x, dr.material);
    auto h = exact ? el.tangent(x, dr.material) : el.hessian(x, dr.material);
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) grad[9 * e + 3 * k + c] = g[k][c];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) hess[81 * e + 9 * (3 * k + l) + 3 * r + c] = h[k][l][r][c];
  }
}

void or_hinge_elements(const or_drape* d, const double* xp, int32_t exact, double* grad, double* hess) {
  const Drape& dr = d->drape;
  auto x = vec3s(xp, dr.mesh.rest.size());
  for (size_t i = 0; i < dr.hinges.size(); ++i) {
    const auto& h = dr.hinges[i];
    auto g = h.gradient(x);
    auto b = exact ? h.tangent(x) : h.hessian(x);
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) grad[12 * i + 3 * k + c] = g[k][c];
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) hess[144 * i + 9 * (4 * k + l) + 3 * r + c] = b[k][l][r][c];
  }
}

void or_row_terms(or_drape* d, const double* xp, const double* xhp, double mu, int32_t curvature, double* grad,
                  double* hess) {
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = vec3s(xp, nv), xh = vec3s(xhp, nv);
  for (int row = 0; row < dr.rows(); ++row) {
    int i = dr.free_[row];
    double m = dr.mass_over_h2(i);
    Vec3 g = {m * (x[i][0] - xh[i][0]), m * (x[i][1] - xh[i][1]), m * (x[i][2] - xh[i][2])};
    Mat3 b = identity_scaled(m * (1.0 + mu));
    if (dr.has_contact()) {
      Vec3 cg = dr.contact.gradient(x[i]);
      Mat3 ch = dr.contact.hessian_with(x[i], curvature != 0);
      for (int r = 0; r < 3; ++r) {
        g[r] += cg[r];
        for (int c = 0; c < 3; ++c) b[r][c] += ch[r][c];
      }
    }
    for (int r = 0; r < 3; ++r) {
      grad[3 * row + r] = g[r];
      for (int c = 0; c < 3; ++c) hess[9 * row + 3 * r + c] = b[r][c];
    }
  }
}

void or_assemble(or_drape* d, const double* xp, const double* xhp, double mu, int32_t curvature, double* rhs,
                 double* vals) {
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  bool saved = dr.barrier_curvature;
  dr.barrier_curvature = curvature != 0;
  DrapeStep step{&dr, vec3s(xhp, nv)};
  auto x = vec3s(xp, nv);
  auto g = step.residual(x);
  for (size_t i = 0; i < g.size(); ++i) rhs[i] = -g[i];
  Bsr m = step.regularized_tangent(x, mu);
  std::memcpy(vals, m.vals.data(), m.vals.size() * sizeof(double));
  dr.barrier_curvature = saved;
}

double or_energy_change(const or_drape* d, const double* xp, const double* xhp, const double* dp, double alpha) {
  const Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  DrapeStep step{&dr, vec3s(xhp, nv)};
  return step.energy_change(vec3s(xp, nv), vec3s(dp, nv), alpha);
}

double or_max_step(const or_drape* d, const double* xp, const double* dp, int32_t* vertex) {
  const Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = vec3s(xp, nv), p = vec3s(dp, nv);
  double alpha = 1.0;
  int at = -1;
  for (int i : dr.free_) {
    int k;
    double s = dr.contact.max_step_limited(x[i], p[i], &k);
    if (k >= 0 && s < alpha) { alpha = s; at = i; }
  }
  if (vertex) *vertex = at;
  return alpha;
}

double or_static_residual(const or_drape* d, const double* xp) {
  return vector_norm(static_gradient(d->drape, vec3s(xp, d->drape.mesh.rest.size())));
}
double or_static_energy(const or_drape* d, const double* xp) {
  return static_energy(d->drape, vec3s(xp, d->drape.mesh.rest.size()));
}

void or_ptc_solve(or_drape* d, const double* start, double h_start, int32_t max_steps, char* log, int64_t log_cap,
                  double* final_positions, or_ptc_stats* stats) {
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = start ? vec3s(start, nv) : dr.rest_positions();
  std::string text;
  PtcRules rules;
  auto t0 = std::chrono::steady_clock::now();
  PtcRun run = ptc_solve(dr, x, h_start, rules, max_steps, false, &text);
  double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (log && log_cap > 0) {
    size_t n = std::min((size_t)log_cap - 1, text.size());
    std::memcpy(log, text.data(), n);
    log[n] = 0;
  }
  if (final_positions)
    for (size_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) final_positions[3 * i + k] = run.states.back()[i][k];
  if (stats) {
    stats->steps = (int32_t)run.steps.size();
    int newton = 0, rejected = 0;
    double max_mu = 0.0;
    for (const auto& s : run.steps) {
      newton += s.inner_iterations;
      rejected += s.rejected_attempts;
      max_mu = std::fmax(max_mu, s.max_mu);
    }
    stats->newton_iterations = newton;
    stats->rejected_attempts = rejected;
    stats->reason = run.reason == "tolerance" ? 0 : run.reason == "floor" ? 1 : run.reason == "cap" ? 2 : 3;
    stats->final_relative_residual = run.final_residual / gravity_load(dr);
    stats->final_h = run.final_h;
    stats->static_energy = static_energy(dr, run.states.back());
    stats->max_mu = max_mu;
    stats->seconds = seconds;
  }
}

int32_t or_ptc_steps(or_drape* d, const double* start, double h_start, int32_t max_steps, double* rows, int32_t cap,
                     double* final_positions, or_ptc_stats* stats) {
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = start ? vec3s(start, nv) : dr.rest_positions();
  PtcRules rules;
  const double load = gravity_load(dr);
  PtcRun run = ptc_solve(dr, x, h_start, rules, max_steps, false, nullptr);
  const int32_t n = (int32_t)run.steps.size();
  for (int32_t i = 0; i < n && i < cap && rows; ++i) {
    const PtcStep& s = run.steps[i];
    double* o = rows + 5 * i;
    o[0] = s.h;
    o[1] = s.residual / load;
    o[2] = s.inner_iterations;
    o[3] = s.rejected_attempts;
    o[4] = s.max_mu;
  }
  if (final_positions)
    for (size_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) final_positions[3 * i + k] = run.states.back()[i][k];
  if (stats) {
    stats->steps = n;
    int newton = 0, rejected = 0;
    double max_mu = 0.0;
    for (const auto& s : run.steps) {
      newton += s.inner_iterations;
      rejected += s.rejected_attempts;
      max_mu = std::fmax(max_mu, s.max_mu);
    }
    stats->newton_iterations = newton;
    stats->rejected_attempts = rejected;
    stats->reason = run.reason == "tolerance" ? 0 : run.reason == "floor" ? 1 : run.reason == "cap" ? 2 : 3;
    stats->final_relative_residual = run.final_residual / load;
    stats->final_h = run.final_h;
    stats->static_energy = static_energy(dr, run.states.back());
    stats->max_mu = max_mu;
    stats->seconds = 0.0;
  }
  return n;
}

double or_static_floor(or_drape* d, const double* x) {
  Drape& dr = d->drape;
  return static_floor(dr, vec3s(x, dr.mesh.rest.size()));
}

void or_drape_set_kappa(or_drape* d, double kappa) { d->drape.contact.barrier.kappa = kappa; }

int32_t or_simulate_states(or_drape* d, int32_t steps, const double* x0, const double* v0, double kappa_start,
                           double tolerance, int32_t max_iterations, int32_t regularized, double floor_factor,
                           int32_t adaptive, double kappa_distance, double kappa_max_factor, int32_t* iterations,
                           double* kappa, double* min_distance, const int32_t* keep, int32_t keep_count,
                           double* positions, double* velocities) {
  Drape& dr = d->drape;
  NewtonSettings s{tolerance, max_iterations, 1e-4, 30};
  auto x = x0 ? vec3s(x0, dr.mesh.rest.size()) : dr.rest_positions();
  std::vector<Vec3> v = v0 ? vec3s(v0, dr.mesh.rest.size()) : std::vector<Vec3>(x.size(), Vec3{0, 0, 0});
  const double kappa_scene = dr.contact.barrier.kappa, dhat = dr.contact.barrier.dhat;
  const double kappa0 = kappa_scene;
  if (kappa_start > 0) dr.contact.barrier.kappa = kappa_start;
  double previous = INFINITY;
  const size_t nv = x.size();
  auto store = [&](int state) {  // state k (0 = the start) when listed in keep
    for (int32_t j = 0; j < keep_count; ++j)
      if (keep[j] == state)
        for (size_t i = 0; i < nv; ++i)
          for (int k = 0; k < 3; ++k) {
            if (positions) positions[(size_t)j * 3 * nv + 3 * i + k] = x[i][k];
            if (velocities) velocities[(size_t)j * 3 * nv + 3 * i + k] = v[i][k];
          }
  };
  store(0);
  for (int n = 0; n < steps; ++n) {
    if (kappa) kappa[n] = dr.contact.barrier.kappa;
    TimeStep t = time_step(dr, x, v, s, regularized != 0, floor_factor);
    if (iterations) iterations[n] = (int32_t)t.solution.searches.size();
    x = t.solution.positions;
    v = t.velocity;
    double dmin = INFINITY;
    for (int i : dr.free_) dmin = std::fmin(dmin, dr.contact.min_distance(x[i]));
    if (min_distance) min_distance[n] = dmin;
    // the IPC stiffness update (Li et al. 2020, section 4.3) once per step, as the library's rules.adaptive_kappa
    if (adaptive && dr.has_contact()) {
      const double cap = kappa_max_factor * kappa0;
      if (dmin < kappa_distance * dhat && dmin < previous && dr.contact.barrier.kappa < cap)
        dr.contact.barrier.kappa = std::fmin(2.0 * dr.contact.barrier.kappa, cap);
      previous = dmin;
    }
    store(n + 1);
  }
  dr.contact.barrier.kappa = kappa_scene;
  return 0;
}

int32_t or_simulate(or_drape* d, int32_t steps, double tolerance, int32_t max_iterations, int32_t regularized,
                    double floor_factor, int32_t* iterations, double* history, int64_t history_cap, int64_t* offsets,
                    double* final_positions) {
  Drape& dr = d->drape;
  NewtonSettings s{tolerance, max_iterations, 1e-4, 30};
  auto x = dr.rest_positions();
  std::vector<Vec3> v(x.size(), Vec3{0, 0, 0});
  int64_t at = 0;
  offsets[0] = 0;
  for (int n = 0; n < steps; ++n) {
    TimeStep t = time_step(dr, x, v, s, regularized != 0, floor_factor);
    iterations[n] = (int32_t)t.solution.searches.size();
    for (double r : t.solution.residuals)
      if (at < history_cap) history[at++] = r;
    offsets[n + 1] = at;
    x = t.solution.positions;
    v = t.velocity;
  }
  if (final_positions)
    for (size_t i = 0; i < x.size(); ++i)
      for (int k = 0; k < 3; ++k) final_positions[3 * i + k] = x[i][k];
  return 0;
}

int32_t or_ptc_inner_solve(or_drape* d, const double* xp, double h, double* out, int32_t* iterations) {
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = vec3s(xp, nv);
  double residual = vector_norm(static_gradient(dr, x));
  dr.time_step = h;
  DrapeStep system{&dr, ptc_prediction(dr, x, h)};
  PtcRules r;
  NewtonSettings settings{r.inner_tolerance, r.inner_max, 1e-4, 30};
  StepSolution sol = solve_step_regularized(system, feasible_start(system, x), residual, settings, 0.0);
  if (iterations) *iterations = (int32_t)sol.searches.size();
  if (out)
    for (size_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) out[3 * i + k] = sol.positions[i][k];
  return (int32_t)sol.status;
}

double or_time_newton_iteration(or_drape* d, const double* xp, double h, int32_t* ok) {
  // One f64 Newton iteration of a PTC attempt at step size h from x (newton.rs:208-258 with mu = 0):
  // tangent assembly, band Cholesky factorisation and solve, Armijo line search.
  Drape& dr = d->drape;
  size_t nv = dr.mesh.rest.size();
  auto x = vec3s(xp, nv);
  auto t0 = std::chrono::steady_clock::now();
  dr.time_step = h;
  DrapeStep system{&dr, ptc_prediction(dr, x, h)};
  auto start = feasible_start(system, x);
  auto g = system.residual(start);
  BandCholesky f;
  bool good = f.factor(system.regularized_tangent(start, 0.0));
  if (good) {
    std::vector<double> neg(g.size());
    for (size_t i = 0; i < g.size(); ++i) neg[i] = -g[i];
    auto direction = system.scatter(f.solve(neg));
    LineSearch ls;
    NewtonSettings settings{0.1, 10, 1e-4, 30};
    good = line_search(system, start, g, direction, settings, &ls);
  }
  if (ok) *ok = good ? 1 : 0;
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void or_drape_derivative_check(int32_t columns, or_derivative_check* out) {
  DerivativeCheck c = drape_derivative_check(columns);
  out->gradient_error = c.gradient_error;
  out->hessian_error = c.hessian_error;
  out->hinge_energy_change_error = c.hinge_energy_change_error;
  out->energy_change_error = c.energy_change_error;
}

or_case* or_case_create(int32_t kind, int32_t columns) { return new or_case{build_case((CaseKind)kind, columns)}; }
void or_case_destroy(or_case* c) { delete c; }
void or_case_sizes(const or_case* c, int64_t* rows, int64_t* blocks) {
  *rows = c->p.matrix.rows();
  *blocks = c->p.matrix.blocks();
}
void or_case_arrays(const or_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals, double* rhs) {
  const Bsr& m = c->p.matrix;
  std::memcpy(row_ptr, m.row_ptr.data(), m.row_ptr.size() * sizeof(int));
  std::memcpy(col_idx, m.col_idx.data(), m.col_idx.size() * sizeof(int));
  std::memcpy(vals, m.vals.data(), m.vals.size() * sizeof(double));
  std::memcpy(rhs, c->p.rhs.data(), c->p.rhs.size() * sizeof(double));
}

static Bsr bsr_of(int64_t rows, int64_t blocks, const int32_t* row_ptr, const int32_t* col_idx, const double* vals) {
  Bsr m;
  m.row_ptr.assign(row_ptr, row_ptr + rows + 1);
  m.col_idx.assign(col_idx, col_idx + blocks);
  m.vals.assign(vals, vals + 9 * blocks);
  return m;
}

void or_pcg(int64_t rows, int64_t blocks, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
            const double* rhs, double tolerance, int32_t max_iterations, int32_t replace_every, double replace_above,
            int32_t single, int32_t breakdown_check, int32_t history_stride, double* solution, double* history,
            int64_t history_cap, int64_t* history_len, or_pcg_stats* stats) {
  Bsr m = bsr_of(rows, blocks, row_ptr, col_idx, vals);
  std::vector<double> b(rhs, rhs + 3 * rows);
  PcgSettings s{tolerance, max_iterations, history_stride, replace_every, replace_above};
  PcgReport r = single ? pcg_solve<float>(m, b, s, breakdown_check != 0) : pcg_solve<double>(m, b, s, breakdown_check != 0);
  if (solution) std::memcpy(solution, r.solution.data(), r.solution.size() * sizeof(double));
  int64_t n = 0;
  for (const auto& h : r.history) {
    if (2 * n + 1 >= history_cap) break;
    history[2 * n] = h.first;
    history[2 * n + 1] = h.second;
    ++n;
  }
  if (history_len) *history_len = n;
  stats->iterations = r.iterations;
  stats->converged = r.converged;
  stats->status = r.status;
  stats->recursive_relative_residual = r.recursive_relative_residual;
  stats->true_relative_residual = r.true_relative_residual_f;
  stats->seconds = r.seconds;
}

int32_t or_band_solve(int64_t rows, int64_t blocks, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                      const double* rhs, double* x, int32_t* bandwidth, double* seconds) {
  Bsr m = bsr_of(rows, blocks, row_ptr, col_idx, vals);
  auto t0 = std::chrono::steady_clock::now();
  BandCholesky f;
  bool ok = f.factor(m);
  if (bandwidth) *bandwidth = f.bandwidth;
  if (!ok) return -1;
  auto y = f.solve(std::vector<double>(rhs, rhs + 3 * rows));
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::memcpy(x, y.data(), y.size() * sizeof(double));
  return 0;
}

}  // extern "C"
