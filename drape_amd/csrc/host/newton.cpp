// The robust Newton loop of one step and the pseudo-transient continuation loop around it.
//
// Host-side control restated from the reference's TypeScript runner (robust-newton.ts:94-256,
// newton-ops.ts:122-139, drape-e4.ts:322-330, drape-e7.ts:56-181, ptc.ts:234-268) with the same f32
// decisions (Math.fround), but with the state resident on the device: the host only reads back the
// scalars each decision rests on (norms, slopes, energy sums, step bounds).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../internal.h"

namespace drape {

namespace {

using Clock = std::chrono::steady_clock;
enum Phase { ASSEMBLY = 0, PCG = 1, LINE_SEARCH = 2, CCD = 3, REDUCE = 4, IO = 5, OTHER = 6 };
enum StepStatusCode { RUNNING = 0, CONVERGED = 1, ITERATION_LIMIT = 2, LINE_SEARCH_FAILED = 3, UPDATE_RULE = 4,
                      REGULARIZATION_CAP = 5 };

inline float f32(double v) { return (float)v; }

struct Run {
  Context& c;
  drape_rules rules;  // with the Newton tolerance / cap of the caller (PTC inner settings)
  std::vector<drape_record> records;
  double phase_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  int64_t pcg_iterations = 0;
  double carried_mu = 0.0;  // rules.mu_carry: the mu the last step's first update was solved at
  explicit Run(Context& ctx) : c(ctx), rules(ctx.rules) {}
};

struct Timer {
  double* slot;
  Clock::time_point t0;
  Timer(Run& r, Phase p) : slot(&r.phase_ms[p]), t0(Clock::now()) {}
  ~Timer() { *slot += std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

// newton.ts:108-111 record: appends the record and, when iterates are kept, the positions at that moment.
int record(Run& run, const drape_record& e) {
  run.records.push_back(e);
  Context& c = run.c;
  if (!c.keep_iterates) return DRAPE_OK;
  Timer t(run, IO);
  std::vector<float4> v(c.V);
  TRY_RC(read_back(c, v.data(), c.pos, v.size() * sizeof(float4)));
  const size_t at = c.iterates.size();
  c.iterates.resize(at + 3 * (size_t)c.V);
  for (int i = 0; i < c.V; ++i) {
    c.iterates[at + 3 * i] = v[i].x;
    c.iterates[at + 3 * i + 1] = v[i].y;
    c.iterates[at + 3 * i + 2] = v[i].z;
  }
  return DRAPE_OK;
}

struct StepState {
  double mu = 0.0, eta = 0.0, previous_norm = NAN;
};

int write_mu(Run& run, double mu) {
  if (!run.rules.regularize) return DRAPE_OK;
  note_shift(run.c, f32(mu), 0.0);
  return launch_set_float(run.c, run.c.mu, f32(mu));
}

struct Residual {
  float norm;
  double norm_squared;
};

int residual(Run& run, Residual* out) {
  Timer t(run, ASSEMBLY);
  Context& c = run.c;
  int rc = launch_evaluate(c, run.rules.exact_tangent != 0);
  if (rc == DRAPE_OK) rc = launch_assemble(c);
  double n2 = 0.0;
  if (rc == DRAPE_OK) rc = launch_rhs_norm(c, &n2);
  out->norm_squared = n2;
  out->norm = f32(std::sqrt(n2));
  return rc;
}

// robust-newton.ts:94-102 (f32 Eisenstat-Walker choice 2).
double next_forcing(const drape_rules& r, double previous, double ratio) {
  const double eta = f32(r.eta_gamma * f32(ratio * ratio));
  const double safeguard = f32(r.eta_gamma * f32(previous * previous));
  const double guarded = safeguard > r.eta_safeguard ? std::fmax(eta, safeguard) : eta;
  return std::fmin(r.eta_max, std::fmax(r.eta_min, guarded));
}
double raise_mu(const drape_rules& r, double mu) { return std::fmin(r.mu_cap, std::fmax(r.mu_start, 10.0 * mu)); }
double relax_mu(const drape_rules& r, double mu) { return mu / 10.0 < r.mu_start ? 0.0 : mu / 10.0; }

drape_record terminal(int step, int iteration, double relative, int status, double mu) {
  drape_record e;
  double* f = &e.step;
  for (int i = 0; i < 25; ++i) f[i] = NAN;
  e.step = step;
  e.iteration = iteration;
  e.relative_residual = relative;
  e.status = status;
  e.mu = mu;
  e.retries = 0;
  return e;
}

struct Solved {
  drape_pcg_stats stats;
  int retries;
  double slope;
};

// robust-newton.ts:121-133: PCG for the current mu; mu rises (with reassembly) on breakdown or non-descent.
int solve_direction(Run& run, StepState& state, Solved* out, bool* capped) {
  *capped = false;
  for (int retries = 0;; ++retries) {
    drape_pcg_stats st{};
    {
      Timer t(run, PCG);
      int rc = pcg_solve(run.c, state.eta, run.rules.pcg_max_iterations, &st);
      if (rc != DRAPE_OK) return rc;
    }
    run.pcg_iterations += st.iterations;
    double slope = 0.0;
    {
      Timer t(run, REDUCE);
      int rc = launch_slope(run.c, &slope);
      if (rc != DRAPE_OK) return rc;
    }
    const bool failed = st.status == 3 || !(slope < 0.0);
    if (!failed || !run.rules.regularize) {
      *out = {st, retries, slope};
      return DRAPE_OK;
    }
    if (state.mu >= run.rules.mu_cap) {
      *capped = true;
      return DRAPE_OK;
    }
    state.mu = raise_mu(run.rules, state.mu);
    int rc = write_mu(run, state.mu);
    if (rc != DRAPE_OK) return rc;
    Residual r;
    rc = residual(run, &r);
    if (rc != DRAPE_OK) return rc;
  }
}

struct Search {
  bool found = false;
  double step_length = NAN;
  int halvings = 0;
  bool noise_only = false;
  double sum = NAN, abs = NAN;
};

// newton-ops.ts:127-139 searchStep (Armijo from alpha_max, noise-aware acceptance).
int search_step(Run& run, double slope, double initial, Search* out) {
  Timer t(run, LINE_SEARCH);
  *out = Search{};
  if (!(slope < 0.0)) return DRAPE_OK;
  const double EPSILON32 = std::ldexp(1.0, -23);
  double factor = 1.0;
  for (int halvings = 0; halvings <= run.rules.max_halvings; ++halvings, factor *= 0.5) {
    const double alpha = initial * factor;
    double sum = 0.0, abs = 0.0;
    int rc = launch_energy_change(run.c, f32(alpha), &sum, &abs);
    if (rc != DRAPE_OK) return rc;
    const double threshold = f32(run.rules.armijo * alpha * slope);
    const bool plain = sum <= threshold;
    const bool accepted = plain || (run.rules.noise_aware && sum <= threshold + 4.0 * EPSILON32 * abs);
    if (accepted) {
      *out = {true, alpha, halvings, !plain, sum, abs};
      return DRAPE_OK;
    }
  }
  return DRAPE_OK;
}

// 2^(exponent - 150) of an f32 value (robust-newton.ts:146-150 resolutionOf).
double f32_spacing(float largest) {
  uint32_t bits;
  std::memcpy(&bits, &largest, 4);
  return std::ldexp(1.0, (int)((bits >> 23) & 0xff) - 150);
}

// robust-newton.ts:209-228: one update; returns false when the step ends.
int update(Run& run, StepState& state, int step, int iteration, float relative, const Residual& res, bool* cont,
           drape_record* extra_out) {
  Context& c = run.c;
  *cont = false;
  Solved solved{};
  bool capped = false;
  int rc = solve_direction(run, state, &solved, &capped);
  if (rc != DRAPE_OK) return rc;
  if (capped) return record(run, terminal(step, iteration, relative, REGULARIZATION_CAP, state.mu));
  float alpha_max = 1.0f, min_dist = INFINITY;
  int32_t limit_row = -1, min_row = -1;
  rc = launch_scatter_direction(c, c.x);
  if (rc != DRAPE_OK) return rc;
  if (c.K > 0) {
    Timer t(run, CCD);
    rc = launch_max_step(c, &alpha_max, &limit_row, &min_dist, &min_row);
    if (rc != DRAPE_OK) return rc;
  }
  Search search;
  rc = search_step(run, solved.slope, alpha_max, &search);
  if (rc != DRAPE_OK) return rc;
  float maxes[4];
  rc = launch_row_maxes(c, c.org, maxes);  // max|d| (before the update)
  if (rc != DRAPE_OK) return rc;
  const double update_inf = search.found ? search.step_length * (double)maxes[0] : NAN;
  drape_record e = terminal(step, iteration, relative, search.found ? RUNNING : LINE_SEARCH_FAILED, state.mu);
  e.pcg_iterations = solved.stats.iterations;
  e.pcg_status = solved.stats.status;
  e.linear_residual = solved.stats.true_relative_residual;
  e.step_length = search.found ? search.step_length : NAN;
  e.halvings = search.found ? search.halvings : NAN;
  if (search.found) {
    rc = launch_advance(c, f32(search.step_length));
    if (rc != DRAPE_OK) return rc;
  }
  rc = launch_row_maxes(c, c.org, maxes);  // step_inf and resolution at the new positions
  if (rc != DRAPE_OK) return rc;
  float dist_after = NAN;
  if (c.K > 0) {
    Timer t(run, CCD);
    float a;
    int32_t r1, r2;
    rc = launch_max_step(c, &a, &r1, &dist_after, &r2);
    if (rc != DRAPE_OK) return rc;
  }
  e.eta = state.eta;
  e.mu = state.mu;
  e.update_inf = update_inf;
  e.step_inf = maxes[1];
  e.retries = solved.retries;
  e.resolution = f32_spacing(maxes[2]);
  e.alpha_max = alpha_max;
  e.limiting_vertex = (alpha_max < 1.0f && limit_row >= 0) ? (double)c.row_vertex[limit_row] : -1.0;
  e.min_distance = c.K > 0 ? dist_after : NAN;
  e.delta_sum = search.sum;
  e.delta_abs = search.abs;
  e.noise_only = search.found ? (search.noise_only ? 1.0 : 0.0) : NAN;
  e.slope = solved.slope;
  e.slope_exact = NAN;
  e.norm_squared = res.norm_squared;
  e.norm_squared_exact = NAN;
  rc = record(run, e);
  if (rc != DRAPE_OK) return rc;
  if (search.found && run.rules.regularize) state.mu = search.halvings == 0 ? relax_mu(run.rules, state.mu) : state.mu;
  *extra_out = e;
  *cont = search.found;
  return DRAPE_OK;
}

// robust-newton.ts:231-256 solveStepRobust. Positions live in c.pos; the origin x_n in c.org.
int solve_step_robust(Run& run, float scale, int step) {
  StepState state;
  state.eta = run.rules.eta_start;
  // robust-newton.ts:232 starts every step at mu = 0; rules.mu_carry starts it where the last step's first
  // update needed it, so that a step whose predecessor needed regularisation does not break down again on
  // the way up
  if (run.rules.mu_carry && run.rules.regularize) state.mu = run.carried_mu;
  for (int iteration = 0;; ++iteration) {
    int rc = write_mu(run, state.mu);
    if (rc != DRAPE_OK) return rc;
    Residual res;
    rc = residual(run, &res);
    if (rc != DRAPE_OK) return rc;
    const float relative = f32((double)res.norm / (double)scale);
    if (relative <= run.rules.newton_tolerance || iteration == run.rules.newton_max_iterations)
      return record(run, terminal(step, iteration, relative,
                                  relative <= run.rules.newton_tolerance ? CONVERGED : ITERATION_LIMIT, state.mu));
    if (run.rules.forcing)
      state.eta = iteration == 0 ? run.rules.eta_start
                                 : next_forcing(run.rules, state.eta, f32((double)res.norm / state.previous_norm));
    else
      state.eta = run.rules.pcg_tolerance;
    state.previous_norm = res.norm;
    bool cont = false;
    drape_record extra;
    rc = update(run, state, step, iteration, relative, res, &cont, &extra);
    if (rc != DRAPE_OK) return rc;
    if (iteration == 0 && cont) run.carried_mu = extra.mu;
    if (!cont) return DRAPE_OK;
    const bool small =
        extra.update_inf <= std::fmax(run.rules.update_stop * extra.step_inf, run.rules.resolution_factor * extra.resolution);
    if (run.rules.floor_stop && small) {
      rc = write_mu(run, state.mu);
      if (rc != DRAPE_OK) return rc;
      rc = residual(run, &res);
      if (rc != DRAPE_OK) return rc;
      return record(run, terminal(step, iteration + 1, f32((double)res.norm / (double)scale), UPDATE_RULE, state.mu));
    }
  }
}

// drape-e4.ts:322-330 newtonStart: x_n + alpha_max (x_hat - x_n) with colliders, else x_hat. c.pos = x_n on entry.
int newton_start(Run& run) {
  Context& c = run.c;
  if (c.K == 0) {
    HIP_TRY(hipMemcpyAsync(c.pos, c.xhat, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    return DRAPE_OK;
  }
  int rc = launch_towards(c);
  if (rc != DRAPE_OK) return rc;
  float alpha = 1.0f, d;
  int32_t r1, r2;
  {
    Timer t(run, CCD);
    rc = launch_max_step(c, &alpha, &r1, &d, &r2);
    if (rc != DRAPE_OK) return rc;
  }
  return launch_step_all(c, alpha);
}

// drape-e7.ts:79-86 classify.
void classify(const std::vector<drape_record>& recs, size_t start, bool non_finite, int* updates, int* reason) {
  int n = 0;
  for (size_t i = start; i < recs.size(); ++i)
    if (recs[i].status == RUNNING && std::isfinite(recs[i].pcg_iterations)) ++n;
  *updates = n;
  const double last = recs.size() > start ? recs.back().status : NAN;
  if (non_finite) *reason = 3;
  else if (n == 0 && last == REGULARIZATION_CAP) *reason = 1;
  else if (n == 0 && last == LINE_SEARCH_FAILED) *reason = 2;
  else *reason = 0;
}

struct Attempt {
  int updates = 0, reason = 0;
  size_t start = 0, end = 0;
};

// drape-e7.ts:97-105 attempt: from the accepted state (c.org) at step size h.
int attempt(Run& run, double h, float scale, int step, bool ready, Attempt* out) {
  Context& c = run.c;
  HIP_TRY(hipMemcpyAsync(c.pos, c.org, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
  int rc = launch_ptc_prepare(c, c.org, h);
  if (rc != DRAPE_OK) return rc;
  note_shift(c, 0.0f, 0.0);
  rc = launch_set_float(c, c.mu, 0.0f);
  if (rc != DRAPE_OK) return rc;
  if (!ready) {
    Residual r;
    rc = residual(run, &r);
    if (rc != DRAPE_OK) return rc;
  }
  rc = newton_start(run);
  if (rc != DRAPE_OK) return rc;
  out->start = run.records.size();
  rc = solve_step_robust(run, scale, step);
  if (rc != DRAPE_OK) return rc;
  out->end = run.records.size();
  float maxes[4];
  rc = launch_row_maxes(c, c.org, maxes);
  if (rc != DRAPE_OK) return rc;
  classify(run.records, out->start, maxes[3] > 0.0f, &out->updates, &out->reason);
  return DRAPE_OK;
}

// Appends the positions at `src` [V float4] to the kept states (Float32Array.from(taken.next),
// drape-e7.ts:175), timed as readback.
int keep_state(Run& run, const float4* src) {
  Timer t(run, IO);
  Context& c = run.c;
  std::vector<float4> v(c.V);
  TRY_RC(read_back(c, v.data(), src, v.size() * sizeof(float4)));
  const size_t at = c.kept_states.size();
  c.kept_states.resize(at + 3 * (size_t)c.V);
  for (int i = 0; i < c.V; ++i) {
    c.kept_states[at + 3 * i] = v[i].x;
    c.kept_states[at + 3 * i + 1] = v[i].y;
    c.kept_states[at + 3 * i + 2] = v[i].z;
  }
  return DRAPE_OK;
}

}  // namespace

int ptc_solve(Context& c, double h_start, int cap, double budget_s, drape_ptc_result* result,
              std::vector<drape_ptc_step>& steps, std::vector<drape_ptc_attempt>& attempts,
              std::vector<drape_record>& records_out) {
  Run run(c);
  const int64_t fallbacks0 = c.fallbacks, read0 = c.read_bytes, trunc0 = c.truncations, breaks0 = c.breakdowns;
  c.kept_states.clear();
  TRY_RC(keep_state(run, c.org));
  run.rules.newton_tolerance = c.rules.inner_tolerance;
  run.rules.newton_max_iterations = c.rules.inner_max_iterations;
  const drape_rules& r = c.rules;
  std::vector<double> residuals;
  double h = h_start;
  int stop = -1;
  const auto started = Clock::now();
  const double floor = r.h0 * r.h_min_factor;
  for (;;) {
    // prepare + residual at the accepted state (drape-e7.ts:164-166)
    HIP_TRY(hipMemcpyAsync(c.pos, c.org, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    int rc = launch_ptc_prepare(c, c.org, h);
    if (rc != DRAPE_OK) return rc;
    rc = write_mu(run, 0.0);
    if (rc != DRAPE_OK) return rc;
    Residual res;
    rc = residual(run, &res);
    if (rc != DRAPE_OK) return rc;
    const float scale = res.norm;
    residuals.push_back(scale);
    // ptc.ts:263-268 stopRule
    const double load = r.gravity_load;
    if (scale / load <= r.equilibrium_f32) stop = 0;
    else {
      const int w = r.stagnation_window;
      if ((int)residuals.size() >= 2 * w) {
        double last_min = INFINITY, before_min = INFINITY;
        for (int i = 0; i < w; ++i) {
          last_min = std::fmin(last_min, residuals[residuals.size() - 1 - i]);
          before_min = std::fmin(before_min, residuals[residuals.size() - 1 - w - i]);
        }
        if (last_min >= r.stagnation_factor * before_min && residuals.back() / load <= 0.25) stop = 1;
      }
      if (stop < 0 && (int)steps.size() >= cap) stop = 2;
    }
    if (stop < 0 && budget_s > 0 && std::chrono::duration<double>(Clock::now() - started).count() > budget_s) stop = 4;
    if (stop >= 0) break;
    // takeStep (drape-e7.ts:133-147)
    const int step_index = (int)steps.size() + 1;
    bool accepted = false;
    for (int rejected = 0; rejected <= r.max_rejections; ++rejected) {
      const double hh = std::fmax(floor, h * std::pow(r.shrink, rejected));
      Attempt a;
      rc = attempt(run, hh, scale, step_index, rejected == 0, &a);
      if (rc != DRAPE_OK) return rc;
      bool dropped = false;
      bool hit_cap = false;
      if (a.reason == 0 && a.updates > 0)
        for (size_t i = a.start; i < a.end; ++i) hit_cap |= run.records[i].status == REGULARIZATION_CAP;
      if (hit_cap) {
        attempts.push_back({(double)step_index, (double)rejected, hh, 4.0, (double)a.start, (double)a.end});
        rc = drape_set_barrier_curvature(reinterpret_cast<drape_ctx*>(&c), 1);
        if (rc != DRAPE_OK) return rc;
        rc = attempt(run, hh, scale, step_index, false, &a);
        int rc2 = drape_set_barrier_curvature(reinterpret_cast<drape_ctx*>(&c), 0);
        if (rc != DRAPE_OK) return rc;
        if (rc2 != DRAPE_OK) return rc2;
        dropped = true;
      }
      attempts.push_back({(double)step_index, (double)rejected, hh, (double)a.reason, (double)a.start, (double)a.end});
      if (a.reason == 0) {
        int64_t pcg = 0;
        double min_distance = NAN;
        for (size_t i = a.start; i < a.end; ++i) {
          if (std::isfinite(run.records[i].pcg_iterations)) pcg += (int64_t)run.records[i].pcg_iterations;
          if (std::isfinite(run.records[i].min_distance)) min_distance = run.records[i].min_distance;
        }
        steps.push_back({(double)step_index, hh, (double)scale, (double)a.updates, (double)rejected, (double)pcg,
                         min_distance, dropped ? 1.0 : 0.0});
        HIP_TRY(hipMemcpyAsync(c.org, c.pos, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
        if (c.keep_states) TRY_RC(keep_state(run, c.org));
        // ptc.ts:241-244 nextStepSize
        const double factor = a.updates < r.inner_max_iterations ? r.growth : r.shrink;
        h = std::fmin(r.h_max, std::fmax(r.h0 * r.h_min_factor, hh * factor));
        accepted = true;
        break;
      }
    }
    if (!accepted) {
      stop = 3;
      break;
    }
  }
  HIP_TRY(hipMemcpyAsync(c.pos, c.org, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
  if (!c.keep_states) TRY_RC(keep_state(run, c.org));  // drape-e7.ts:178: the first and the last state
  HIP_TRY(hipStreamSynchronize(c.stream));
  const double seconds = std::chrono::duration<double>(Clock::now() - started).count();
  result->stop = stop;
  result->accepted_steps = (int32_t)steps.size();
  int newton = 0;
  for (const auto& e : run.records)
    if (std::isfinite(e.pcg_iterations)) ++newton;
  result->newton_iterations = newton;
  result->pcg_iterations = run.pcg_iterations;
  result->records = (int32_t)run.records.size();
  result->attempts = (int32_t)attempts.size();
  result->final_h = h;
  result->final_residual = residuals.back();
  result->gravity_load = r.gravity_load;
  result->seconds = seconds;
  double tracked = 0.0;
  for (int i = 0; i < OTHER; ++i) {
    result->phase_ms[i] = run.phase_ms[i];
    tracked += run.phase_ms[i];
  }
  result->phase_ms[OTHER] = 1000.0 * seconds - tracked;
  result->fallbacks = c.fallbacks - fallbacks0;
  result->truncations = c.truncations - trunc0;
  result->breakdowns = c.breakdowns - breaks0;
  result->read_bytes = c.read_bytes - read0;
  records_out = std::move(run.records);
  return DRAPE_OK;
}

// The barrier-stiffness update of Incremental Potential Contact (Li et al. 2020, section 4.3), applied after
// each time step: when the smallest vertex-collider distance is below kappa_distance * dhat and smaller than
// after the previous step, kappa doubles (exactly, in FP32), up to kappa_max_factor times its start value.
int adapt_kappa(Context& c, float d, float* previous_min) {
  const float dhat = c.colliders_host[1];
  const float kappa = c.colliders_host[2];
  const float cap = (float)(c.rules.kappa_max_factor * c.kappa0);
  if (d < (float)(c.rules.kappa_distance * dhat) && d < *previous_min && kappa < cap)
    TRY_RC(set_kappa(c, std::fmin(2.0f * kappa, cap)));
  *previous_min = d;
  return DRAPE_OK;
}

// drape.rs:426-439 simulate / time_step: per step x_hat = x + h v + h^2 g, the residual scale |g(x_n)|, the
// start (x_hat, cut by CCD when colliders exist), one robust Newton solve, v = (x_{n+1} - x_n) / h; then,
// under rules.adaptive_kappa, the stiffness update.
int simulate(Context& c, int steps, std::vector<int32_t>& updates, std::vector<float>& relative,
             std::vector<drape_record>& records_out) {
  Run run(c);
  const double h = c.problem.time_step;
  c.kappa_history.clear();
  c.step_distance.clear();
  float previous_min = INFINITY;
  for (int n = 0; n < steps; ++n) {
    c.kappa_history.push_back(c.K > 0 ? (double)c.colliders_host[2] : 0.0);
    HIP_TRY(hipMemcpyAsync(c.org, c.pos, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    int rc = launch_dyn_predict(c, c.vel, h);
    if (rc != DRAPE_OK) return rc;
    note_shift(c, 0.0f, 0.0);
    rc = launch_set_float(c, c.mu, 0.0f);
    if (rc != DRAPE_OK) return rc;
    Residual r;
    rc = residual(run, &r);
    if (rc != DRAPE_OK) return rc;
    const float scale = r.norm > 0.0f ? r.norm : 1.0f;
    rc = newton_start(run);
    if (rc != DRAPE_OK) return rc;
    const size_t first = run.records.size();
    rc = solve_step_robust(run, scale, n + 1);
    if (rc != DRAPE_OK) return rc;
    int n_up = 0;
    for (size_t i = first; i < run.records.size(); ++i)
      if (run.records[i].status == RUNNING && std::isfinite(run.records[i].pcg_iterations)) ++n_up;
    updates.push_back(n_up);
    relative.push_back(run.records.empty() ? NAN : (float)run.records.back().relative_residual);
    rc = launch_dyn_velocity(c, c.vel, h);
    if (rc != DRAPE_OK) return rc;
    if (c.K > 0) {
      float d = INFINITY;
      TRY_RC(launch_min_distance(c, &d));
      c.step_distance.push_back(d);
      if (c.rules.adaptive_kappa) TRY_RC(adapt_kappa(c, d, &previous_min));
    }
  }
  records_out = std::move(run.records);
  return DRAPE_OK;
}

// drape-e7.ts:183-196 runPtcProbes, one probe: prepare x_hat and M/h^2 at `start` (c.org), the residual scale
// there, the Newton start and one robust Newton solve with the PTC inner settings.
int ptc_probe(Context& c, double h, int step, std::vector<drape_record>& records_out) {
  Run run(c);
  run.rules.newton_tolerance = c.rules.inner_tolerance;
  run.rules.newton_max_iterations = c.rules.inner_max_iterations;
  HIP_TRY(hipMemcpyAsync(c.pos, c.org, (size_t)c.V * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
  TRY_RC(launch_ptc_prepare(c, c.org, h));
  TRY_RC(write_mu(run, 0.0));
  Residual res;
  TRY_RC(residual(run, &res));
  TRY_RC(newton_start(run));
  TRY_RC(solve_step_robust(run, res.norm, step));
  records_out = std::move(run.records);
  return DRAPE_OK;
}

int newton_step(Context& c, float scale, int step, std::vector<drape_record>& records_out) {
  Run run(c);
  int rc = solve_step_robust(run, scale, step);
  records_out = std::move(run.records);
  return rc;
}

}  // namespace drape

using namespace drape;

extern "C" {

int drape_newton_step(drape_ctx* p, float* x_inout, float scale, int32_t step, const float* origin,
                      drape_record* records, int32_t cap, int32_t* count) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear || !x_inout || !origin || !(scale > 0)) {
    set_error("invalid Newton arguments");
    return DRAPE_E_INVALID;
  }
  std::vector<float4> v(c->V);
  for (int i = 0; i < c->V; ++i) v[i] = make_float4(origin[3 * i], origin[3 * i + 1], origin[3 * i + 2], 0.0f);
  HIP_TRY(hipMemcpyAsync(c->org, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  for (int i = 0; i < c->V; ++i) v[i] = make_float4(x_inout[3 * i], x_inout[3 * i + 1], x_inout[3 * i + 2], 0.0f);
  HIP_TRY(hipMemcpyAsync(c->pos, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  std::vector<drape_record> recs;
  int rc = newton_step(*c, scale, step, recs);
  if (rc != DRAPE_OK) return rc;
  HIP_TRY(hipMemcpyAsync(v.data(), c->pos, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->V; ++i) { x_inout[3 * i] = v[i].x; x_inout[3 * i + 1] = v[i].y; x_inout[3 * i + 2] = v[i].z; }
  const int n = (int)recs.size();
  if (records)
    for (int i = 0; i < n && i < cap; ++i) records[i] = recs[i];
  if (count) *count = n;
  return DRAPE_OK;
}

int drape_simulate(drape_ctx* p, int32_t steps, const float* x0, const float* v0, float* x_out, float* v_out,
                   int32_t* updates, float* final_relative, drape_record* records, int32_t cap, int32_t* count) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear || steps < 0) {
    set_error("invalid simulation arguments");
    return DRAPE_E_INVALID;
  }
  if (!c->vel && hipMalloc((void**)&c->vel, (size_t)c->V * sizeof(float4)) != hipSuccess) {
    set_error("velocity buffer allocation failed");
    return DRAPE_E_DEVICE;
  }
  std::vector<float4> v(c->V, make_float4(0, 0, 0, 0));
  const float* src = x0 ? x0 : c->rest_host.data();
  for (int i = 0; i < c->V; ++i) v[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0f);
  HIP_TRY(hipMemcpyAsync(c->pos, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->V; ++i)
    v[i] = v0 ? make_float4(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
  HIP_TRY(hipMemcpyAsync(c->vel, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<int32_t> up;
  std::vector<float> rel;
  std::vector<drape_record> recs;
  int rc = simulate(*c, steps, up, rel, recs);
  if (rc != DRAPE_OK) return rc;
  auto out3 = [&](const float4* d, float* host) -> int {
    if (!host) return DRAPE_OK;
    HIP_TRY(hipMemcpyAsync(v.data(), d, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->V; ++i) { host[3 * i] = v[i].x; host[3 * i + 1] = v[i].y; host[3 * i + 2] = v[i].z; }
    return DRAPE_OK;
  };
  rc = out3(c->pos, x_out);
  if (rc == DRAPE_OK) rc = out3(c->vel, v_out);
  if (rc != DRAPE_OK) return rc;
  for (int n = 0; n < steps; ++n) {
    if (updates) updates[n] = up[n];
    if (final_relative) final_relative[n] = rel[n];
  }
  const int n = (int)recs.size();
  if (records)
    for (int i = 0; i < n && i < cap; ++i) records[i] = recs[i];
  if (count) *count = n;
  return DRAPE_OK;
}

int drape_ptc_solve(drape_ctx* p, const float* start, double h_start, int32_t cap, double budget_s,
                    drape_ptc_result* result, drape_ptc_step* steps, int32_t steps_cap, drape_ptc_attempt* attempts,
                    int32_t attempts_cap, drape_record* records, int32_t records_cap, float* final_positions) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear || !result || !(h_start > 0) || cap < 0) {
    set_error("invalid PTC arguments");
    return DRAPE_E_INVALID;
  }
  if (!(c->rules.gravity_load > 0)) {
    set_error("rules.gravity_load must be the load |m g| of the problem (ptc section of problem.json)");
    return DRAPE_E_INVALID;
  }
  std::vector<float4> v(c->V);
  if (start) {
    for (int i = 0; i < c->V; ++i) v[i] = make_float4(start[3 * i], start[3 * i + 1], start[3 * i + 2], 0.0f);
    HIP_TRY(hipMemcpyAsync(c->org, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(c->org, c->pos, v.size() * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
  }
  std::vector<drape_ptc_step> st;
  std::vector<drape_ptc_attempt> at;
  std::vector<drape_record> recs;
  int rc = ptc_solve(*c, h_start, cap, budget_s, result, st, at, recs);
  if (rc != DRAPE_OK) return rc;
  if (steps)
    for (int i = 0; i < (int)st.size() && i < steps_cap; ++i) steps[i] = st[i];
  if (attempts)
    for (int i = 0; i < (int)at.size() && i < attempts_cap; ++i) attempts[i] = at[i];
  if (records)
    for (int i = 0; i < (int)recs.size() && i < records_cap; ++i) records[i] = recs[i];
  if (final_positions) {
    HIP_TRY(hipMemcpyAsync(v.data(), c->pos, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->V; ++i) {
      final_positions[3 * i] = v[i].x; final_positions[3 * i + 1] = v[i].y; final_positions[3 * i + 2] = v[i].z;
    }
  }
  return DRAPE_OK;
}

int drape_ptc_probe(drape_ctx* p, const float* start, double h, int32_t step, drape_record* records, int32_t cap,
                    int32_t* count, float* iterates, float* final_positions) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear || !start || !(h > 0) || cap < 0) {
    set_error("invalid probe arguments");
    return DRAPE_E_INVALID;
  }
  std::vector<float4> v(c->V);
  for (int i = 0; i < c->V; ++i) v[i] = make_float4(start[3 * i], start[3 * i + 1], start[3 * i + 2], 0.0f);
  HIP_TRY(hipMemcpyAsync(c->org, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  c->keep_iterates = iterates != nullptr;
  c->iterates.clear();
  std::vector<drape_record> recs;
  int rc = ptc_probe(*c, h, step, recs);
  c->keep_iterates = false;
  if (rc != DRAPE_OK) return rc;
  const int n = (int)recs.size();
  if (records)
    for (int i = 0; i < n && i < cap; ++i) records[i] = recs[i];
  if (count) *count = n;
  const size_t per = 3 * (size_t)c->V;
  if (iterates)
    for (int i = 0; i < n && i < cap; ++i)
      std::copy(c->iterates.begin() + i * per, c->iterates.begin() + (i + 1) * per, iterates + i * per);
  c->iterates.clear();
  if (final_positions) {
    HIP_TRY(hipMemcpyAsync(v.data(), c->pos, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->V; ++i) {
      final_positions[3 * i] = v[i].x; final_positions[3 * i + 1] = v[i].y; final_positions[3 * i + 2] = v[i].z;
    }
  }
  return DRAPE_OK;
}

int drape_ptc_keep_states(drape_ctx* p, int32_t keep) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear) {
    set_error("invalid context");
    return DRAPE_E_INVALID;
  }
  c->keep_states = keep != 0;
  return DRAPE_OK;
}

int drape_ptc_states(drape_ctx* p, float* states, int32_t cap, int32_t* count) {
  Context* c = reinterpret_cast<Context*>(p);
  if (!c || c->linear || cap < 0) {
    set_error("invalid context or cap");
    return DRAPE_E_INVALID;
  }
  const size_t per = 3 * (size_t)c->V;
  const int32_t n = (int32_t)(c->kept_states.size() / per);
  if (count) *count = n;
  if (states)
    for (int32_t i = 0; i < n && i < cap; ++i)
      std::copy(c->kept_states.begin() + i * per, c->kept_states.begin() + (i + 1) * per, states + i * per);
  return DRAPE_OK;
}

}  // extern "C"
