// Oracle: a CPU f64 restatement of drape's Rust oracle (engine/oracle/src), used ONLY as test
// infrastructure. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker (or as the timed CPU baseline) -- never as the thing measured or shipped.
//
// Every function keeps the operation order of the Rust file:line it cites, so that f64 results
// (and the f32 PCG emulation) reproduce the reference's committed golden vectors. Build with
// -ffp-contract=off: the Rust oracle is compiled without FMA contraction.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace oracle {

using Vec2 = std::array<double, 2>;
using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

// ---- linalg.rs:1-60 -------------------------------------------------------------------------
inline Vec3 add(Vec3 a, Vec3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline Vec3 scale(Vec3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot(Vec3 a, Vec3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Mat3 outer(Vec3 a, Vec3 b) {
  Mat3 m{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = a[r] * b[c];
  return m;
}
inline void mat_add_scaled(Mat3& t, const Mat3& s, double f) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[r][c] += s[r][c] * f;
}
inline Mat3 identity_scaled(double s) {
  Mat3 m{};
  m[0][0] = s; m[1][1] = s; m[2][2] = s;
  return m;
}
double vector_norm(const std::vector<double>& v);                         // linalg.rs:54-56
double difference_norm(const std::vector<double>& a, const std::vector<double>& b);  // :58-60
double free_dot(const std::vector<double>& a, const std::vector<double>& b);  // coupon.rs:248-250

// ---- rng.rs:2-13 ----------------------------------------------------------------------------
struct Lcg {
  uint64_t state;
  explicit Lcg(uint64_t seed) : state(seed * 6364136223846793005ULL + 1442695040888963407ULL) {}
  double uniform() {
    state = state * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(state >> 11) / (double)(1ULL << 53) * 2.0 - 1.0;
  }
};

// ---- mesh.rs:3-107 (+ the net-new irregular variant) ----------------------------------------
struct Mesh {
  int columns = 0, rows = 0;
  double width = 0, height = 0;
  std::vector<Vec2> rest;
  std::vector<std::array<int, 3>> triangles;
  std::vector<std::array<int, 4>> hinges;

  static Mesh grid(int columns, int rows, double width, double height);
  // Net-new (no reference): interior vertices jittered by U(+-jitter*spacing) per axis and a
  // `flip_fraction` of the cells split along the other diagonal, from one seeded Lcg.
  static Mesh irregular(int columns, double size, double jitter, double flip_fraction, uint64_t seed);
  bool is_boundary(int v) const;
  std::vector<double> vertex_areas() const;
};
double rest_area(const std::vector<Vec2>& rest, const std::array<int, 3>& t);
std::vector<std::array<int, 4>> interior_hinges(const std::vector<Vec2>& rest,
                                                const std::vector<std::array<int, 3>>& tris);

// ---- material.rs:4-56 -----------------------------------------------------------------------
struct Material {
  double wale, course, coupling, shear, bend_wale, bend_course, density;
  static Material knit_baseline() { return {1000.0, 300.0, 100.0, 50.0, 2.0e-6, 1.0e-6, 0.13}; }
  // Isotropic plane-stress sheet (E, nu, thickness) in the reference's parameters:
  // wale = course = E t/(1-nu^2), coupling = nu E t/(1-nu^2), shear = E t/(2(1+nu)).
  static Material isotropic(double E, double nu, double t, double bend, double density);
  std::array<double, 3> stress(std::array<double, 3> e) const {
    return {wale * e[0] + coupling * e[1], course * e[1] + coupling * e[0], 2.0 * shear * e[2]};
  }
  double energy_density(std::array<double, 3> e) const {
    return 0.5 * wale * e[0] * e[0] + 0.5 * course * e[1] * e[1] + coupling * e[0] * e[1] +
           2.0 * shear * e[2] * e[2];
  }
  bool is_convex() const { return coupling * coupling < wale * course && shear > 0.0; }
  double hinge_bending(Vec2 d) const { return bend_wale * d[1] * d[1] + bend_course * d[0] * d[0]; }
};

// ---- membrane.rs:6-140 ----------------------------------------------------------------------
using ElementHessian = std::array<std::array<Mat3, 3>, 3>;
struct MembraneElement {
  std::array<int, 3> vertices;
  std::array<Vec2, 3> grad_n;
  double area;
  static MembraneElement make(std::array<int, 3> v, const std::vector<Vec2>& rest);
  void columns(const std::vector<Vec3>& x, Vec3& wale, Vec3& course) const;
  std::array<double, 3> strain(const std::vector<Vec3>& x) const;
  double energy(const std::vector<Vec3>& x, const Material& m) const;
  std::array<Vec3, 3> gradient(const std::vector<Vec3>& x, const Material& m) const;
  ElementHessian hessian(const std::vector<Vec3>& x, const Material& m) const;  // projected
  ElementHessian tangent(const std::vector<Vec3>& x, const Material& m) const;  // exact
  double energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha,
                       const Material& m) const;
  ElementHessian hessian_with_stress(const std::vector<Vec3>& x, const Material& m,
                                     std::array<double, 3> s) const;
};
std::array<double, 3> clamp_negative_stress(std::array<double, 3> s);

// ---- dual.rs:6-93 ---------------------------------------------------------------------------
struct Dual {
  double v, t;
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.t + b.t}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.t - b.t}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.t * b.v + a.v * b.t}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.t * b.v - a.v * b.t) / (b.v * b.v)}; }
inline Dual operator-(Dual a) { return {-a.v, -a.t}; }
inline Dual dsqrt(Dual a) {
  double s = std::sqrt(a.v);
  return {s, a.t / (2.0 * s)};
}
inline double dsqrt(double a) { return std::sqrt(a); }
template <class T> inline T constant_of(double v);
template <> inline double constant_of<double>(double v) { return v; }
template <> inline Dual constant_of<Dual>(double v) { return {v, 0.0}; }

// ---- hinge.rs:6-120 -------------------------------------------------------------------------
using HingeBlocks = std::array<std::array<Mat3, 4>, 4>;
struct HingeElement {
  std::array<int, 4> vertices;
  double stiffness;
  double angle(const std::vector<Vec3>& x) const;
  std::array<Vec3, 4> angle_gradient(const std::vector<Vec3>& x) const;
  double energy(const std::vector<Vec3>& x) const;
  std::array<Vec3, 4> gradient(const std::vector<Vec3>& x) const;
  HingeBlocks hessian(const std::vector<Vec3>& x) const;  // Gauss-Newton
  HingeBlocks angle_hessian(const std::vector<Vec3>& x) const;
  HingeBlocks tangent(const std::vector<Vec3>& x) const;  // exact
  double energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const;
};
double hinge_angle_of(const std::array<Vec3, 4>& p);
template <class T>
std::array<std::array<T, 3>, 4> angle_gradient_of(const std::array<std::array<T, 3>, 4>& p);
double hinge_stiffness(const std::vector<Vec2>& rest, std::array<int, 4> v, double bending);

// ---- contact.rs:5-176 -----------------------------------------------------------------------
constexpr double KEEP_FRACTION = 0.1;
struct Collider {
  int kind;  // 1 plane (n, offset), 2 sphere (centre, radius)
  Vec3 a;
  double w;
  double distance(Vec3 x) const;
  Vec3 normal(Vec3 x) const;
  double curvature(Vec3 x) const;
  double max_step(Vec3 x, Vec3 p) const;
  std::array<double, 8> packed() const { return {(double)kind, a[0], a[1], a[2], w, 0, 0, 0}; }
};
struct Barrier {
  double dhat = 0, kappa = 0;
  double value(double d) const;
  double first(double d) const;
  double second(double d) const;
  static double resting_kappa(double dhat, double mass, double gravity);
};
struct Contact {
  std::vector<Collider> colliders;
  Barrier barrier;
  double energy(Vec3 x) const;
  Vec3 gradient(Vec3 x) const;
  Mat3 hessian_with(Vec3 x, bool curvature) const;
  double max_step(Vec3 x, Vec3 p) const;
  double max_step_limited(Vec3 x, Vec3 p, int* at) const;
  double min_distance(Vec3 x) const;
};

// ---- bsr.rs:4-93 ----------------------------------------------------------------------------
struct Bsr {
  std::vector<int> row_ptr, col_idx;
  std::vector<double> vals;
  static Bsr with_pattern(int rows, const std::vector<std::vector<int>>& groups);
  int rows() const { return (int)row_ptr.size() - 1; }
  int blocks() const { return (int)col_idx.size(); }
  int position(int row, int column) const;
  void add(int row, int column, const Mat3& b);
  std::vector<double> multiply(const std::vector<double>& x) const;
  std::vector<int> diagonal_positions() const;
  double asymmetry() const;
  Bsr rounded_to_f32() const;
};

// ---- band.rs:4-100 --------------------------------------------------------------------------
struct BandCholesky {
  int size = 0, bandwidth = 0;
  std::vector<double> lower;
  // Returns false (and the reason) when a pivot is not positive.
  bool factor(const Bsr& m, std::string* error = nullptr);
  std::vector<double> solve(const std::vector<double>& rhs) const;
  size_t index(int i, int j) const { return (size_t)i * (bandwidth + 1) + (j + bandwidth - i); }
  int first(int i) const { return i > bandwidth ? i - bandwidth : 0; }
};

// ---- pcg.rs:60-281 (f32 and f64 instantiations) ---------------------------------------------
constexpr int WORKGROUP = 256;
struct PcgSettings {
  double tolerance;
  int max_iterations;
  int history_stride;
  int replace_every;
  double replace_above;
};
struct PcgReport {
  int iterations = 0;
  bool converged = false;
  int status = 0;  // GPU status code: 1 converged, 2 iteration limit, 3 breakdown
  double recursive_relative_residual = 1.0;
  double true_relative_residual_f = NAN;  // sqrt(true_rr/rr0) in the solver's precision
  std::vector<double> solution, alphas, betas;
  std::vector<std::pair<int, double>> history;
  double seconds = 0;
};
// `breakdown_check` adds the GPU's pq <= 0 / non-finite alpha stop (status 3), which the Rust
// emulation does not have; on an SPD matrix it never fires and results are identical.
template <class T>
PcgReport pcg_solve(const Bsr& m, const std::vector<double>& rhs, const PcgSettings& s,
                    bool breakdown_check = false);
template <class T> T reduce_rows_of(int rows, const std::vector<T>& values);  // pcg.rs:236-257
template <class T> std::array<T, 9> invert3(const T* m);                        // pcg.rs:268-281

// ---- newton.rs:38-264 + e4_rules.rs ---------------------------------------------------------
struct NewtonSettings {
  double tolerance;
  int max_iterations;
  double armijo;
  int max_halvings;
};
enum StepStatus { Converged = 0, IterationLimit, LineSearchFailed, RegularizationCap, ConvergedAtFloor };
const char* step_status_name(StepStatus s);
struct LineSearch {
  double step_length;
  int halvings;
};
struct Regularization {
  double max_mu = 0;
  int retries = 0;
  bool cap_hit = false;
};
struct StepSolution {
  std::vector<Vec3> positions;
  std::vector<double> residuals;
  std::vector<LineSearch> searches;
  StepStatus status = IterationLimit;
  Regularization regularization;
};
constexpr double MU_START = 1e-3, MU_CAP = 1e3;
double raise_mu(double mu);
double relax_mu(double mu);
double next_forcing(double previous, double ratio);  // e4_rules.rs:15-21
double f32_spacing(double max_abs);                  // e4_rules.rs:37-40
bool update_is_small(double update_inf, double step_inf, double resolution);

// ---- drape.rs:53-439 with every Rust `const` turned into an input ----------------------------
struct SceneSpec {
  int columns = 33;
  double sheet_size = 0.5;
  double rest_factor = 1.0;
  Material material = Material::knit_baseline();
  double gravity = 9.81;
  double time_step = 1.0 / 240.0;
  bool pinned = true;  // the two top corners (rows-1)*columns and rows*columns-1
  std::vector<Collider> colliders;
  double dhat = 0.0;
  double kappa = 0.0;  // <= 0: Barrier::resting_kappa(dhat, heaviest vertex, gravity)
  double lift = 0.0;
  bool exact_tangent = true;
  double jitter = 0.0, flip_fraction = 0.0;
  uint64_t mesh_seed = 0;
};
// The reference's own scenes: Scene::PinnedCorners and Scene::SphereOnFloor with SceneParams
// (drape.rs:53-84, 145-155), with the knit material scaled by `material_factor`.
SceneSpec reference_scene(int columns, bool sphere_on_floor, double material_factor,
                          double radius_factor, double rest_factor, bool placed);

struct Drape {
  SceneSpec spec;
  Mesh mesh;
  Material material;
  std::vector<MembraneElement> membranes;
  std::vector<HingeElement> hinges;
  std::vector<int> fixed, free_, row_of;  // row_of[v] = -1 for fixed vertices
  Bsr pattern;
  std::vector<double> mass;
  bool exact_tangent = true;
  Contact contact;
  double lift = 0;
  double time_step = 1.0 / 240.0;
  bool barrier_curvature = true;

  explicit Drape(const SceneSpec& spec);
  int rows() const { return (int)free_.size(); }
  bool has_contact() const { return !contact.colliders.empty(); }
  std::vector<Vec3> rest_positions() const;
  std::vector<Vec3> prediction(const std::vector<Vec3>& x, const std::vector<Vec3>& v) const;
  double potential(const std::vector<Vec3>& x) const;
  double barrier_energy(const std::vector<Vec3>& x) const;
  std::vector<Vec3> potential_gradient(const std::vector<Vec3>& x) const;
  double total_energy(const std::vector<Vec3>& x, const std::vector<Vec3>& v) const;
  double mass_over_h2(int v) const { return mass[v] / (time_step * time_step); }
};

struct DrapeStep {
  const Drape* drape;
  std::vector<Vec3> x_hat;
  double value(const std::vector<Vec3>& x) const;
  std::vector<double> residual(const std::vector<Vec3>& x) const;
  Bsr tangent(const std::vector<Vec3>& x) const;
  Bsr regularized_tangent(const std::vector<Vec3>& x, double mu) const;
  double energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const;
  double barrier_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const;
  std::vector<Vec3> scatter(const std::vector<double>& rows) const;
  std::vector<double> gather(const std::vector<Vec3>& v) const;
  double max_step(const std::vector<Vec3>& x, const std::vector<Vec3>& d) const;
  double residual_floor(const std::vector<Vec3>& x) const;
  std::vector<double> absolute_contributions(const std::vector<Vec3>& x) const;
  std::vector<double> stiffness_times_positions(const std::vector<Vec3>& x) const;
};

std::vector<Vec3> feasible_start(const DrapeStep& step, const std::vector<Vec3>& x);
bool line_search(const DrapeStep& s, const std::vector<Vec3>& x, const std::vector<double>& g,
                 const std::vector<Vec3>& d, const NewtonSettings& settings, LineSearch* out);
StepSolution solve_step(const DrapeStep& s, std::vector<Vec3> start, double scale,
                        const NewtonSettings& settings);
StepSolution solve_step_regularized(const DrapeStep& s, std::vector<Vec3> start, double scale,
                                    const NewtonSettings& settings, double floor_factor /*<=0: none*/);

struct TimeStep {
  StepSolution solution;
  double scale;
  std::vector<Vec3> velocity;
};
TimeStep time_step(const Drape& d, const std::vector<Vec3>& x, const std::vector<Vec3>& v,
                   const NewtonSettings& s, bool regularized, double floor_factor);

// ---- e7_ptc.rs:9-189 ------------------------------------------------------------------------
struct PtcRules {
  double h0 = 1.0 / 240.0, h_max = 1.0, h_min_factor = 0.1, growth = 2.0, shrink = 0.5;
  double inner_tolerance = 0.1;
  int inner_max = 10, max_rejections = 12, max_steps = 1000;
  double equilibrium = 1e-8, floor_factor = 2.0, floor_guard = 1e-6;
};
struct PtcStep {
  double h, residual;
  int inner_iterations, rejected_attempts;
  double max_mu;
  bool mu_cap, curvature_dropped;
};
struct PtcRun {
  std::vector<std::vector<Vec3>> states;
  std::vector<PtcStep> steps;
  double final_residual = NAN, final_h = 0;
  std::string reason = "cap";
};
double next_step_size(const PtcRules& r, double h, int inner);
double gravity_load(const Drape& d);
std::vector<double> static_gradient(const Drape& d, const std::vector<Vec3>& x);
double static_energy(const Drape& d, const std::vector<Vec3>& x);
std::vector<Vec3> ptc_prediction(const Drape& d, const std::vector<Vec3>& x, double h);
double static_floor(Drape& d, const std::vector<Vec3>& x);
bool ptc_failed(StepStatus s, int iterations, const std::vector<Vec3>& x);
// `max_steps_override` > 0 stops after that many accepted steps (a bounded sample).
PtcRun ptc_solve(Drape& d, std::vector<Vec3> start, double h_start, const PtcRules& r,
                 int max_steps_override = 0, bool keep_states = true, std::string* log = nullptr);

// ---- cases.rs:64-161 (E1 linear systems) ----------------------------------------------------
enum CaseKind { Tension = 0, Slack = 1, Dynamic = 2 };
struct CaseProblem {
  Mesh mesh;
  Bsr matrix;
  std::vector<double> rhs;
};
CaseProblem build_case(CaseKind kind, int columns);

// ---- e3_generate.rs:84-166 + e5_checks.rs:58-86 (problem arrays) ----------------------------
struct ProblemArrays {
  std::vector<float> rest_positions, elem_geom, hinge_stiffness, row_mass, row_unit_mass,
      colliders, kernel_state, contact_state;
  std::vector<uint32_t> elem_vertices, hinge_vertices, row_vertex, row_ptr, col_idx, diag_idx,
      block_contrib_ptr, block_contrib_idx, force_contrib_ptr, force_contrib_idx, fixed_vertex;
};
ProblemArrays problem_arrays(const Drape& d);
std::vector<Vec3> bent_state(const Drape& d, Lcg& rng);    // e3_checks.rs:164-169

// e3_checks.rs:136-204: FD checks of Phi at the seed-53 bent state (exact tangent, pinned sheet).
struct DerivativeCheck {
  double gradient_error, hessian_error, hinge_energy_change_error, energy_change_error;
};
DerivativeCheck drape_derivative_check(int columns);
std::vector<Vec3> contact_state(const Drape& d);           // e5_checks.rs:60-77

}  // namespace oracle
