// Unit tests of the oracle: the reference's own Rust #[test]s (closed forms, identities and finite
// differences on tiny hand-built inputs), restated in C++ with the same inputs and tolerances.
// `selftest` runs every test and prints "PASS <name>" or "FAIL <name>: <detail>"; exit status 1 on
// any failure. `selftest <name>` runs one.
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>

#include "oracle.hpp"

using namespace oracle;

struct Failure {
  std::string what;
};
#define CHECK(cond, ...)                                          \
  do {                                                            \
    if (!(cond)) {                                                \
      char _b[512];                                               \
      std::snprintf(_b, sizeof _b, __VA_ARGS__);                  \
      throw Failure{std::string(#cond) + " | " + _b};             \
    }                                                             \
  } while (0)

// testing.rs:4-37
static std::vector<Vec3> random_positions(const std::vector<Vec2>& rest, Lcg& rng, double stretch, double noise) {
  std::vector<Vec3> x;
  for (const auto& p : rest) {
    double a = stretch * p[0] + noise * rng.uniform();
    double b = p[1] + noise * rng.uniform();
    double c = noise * rng.uniform();
    x.push_back({a, b, c});
  }
  return x;
}
static std::vector<Vec3> stretched(const std::vector<Vec2>& rest, double wale, double course, Lcg& rng, double noise) {
  std::vector<Vec3> x;
  for (const auto& p : rest) {
    double a = wale * p[0] + noise * rng.uniform();
    double b = course * p[1] + noise * rng.uniform();
    double c = noise * rng.uniform();
    x.push_back({a, b, c});
  }
  return x;
}
static double central(const std::vector<Vec3>& x, int v, int axis, double h, const std::function<double(const std::vector<Vec3>&)>& f) {
  auto plus = x, minus = x;
  plus[v][axis] += h;
  minus[v][axis] -= h;
  return (f(plus) - f(minus)) / (2.0 * h);
}
template <size_t N>
static std::array<Vec3, N> gradient_derivative(const std::vector<Vec3>& x, int v, int axis, double h,
                                               const std::function<std::array<Vec3, N>(const std::vector<Vec3>&)>& g) {
  auto plus = x, minus = x;
  plus[v][axis] += h;
  minus[v][axis] -= h;
  auto gp = g(plus), gm = g(minus);
  std::array<Vec3, N> out;
  for (size_t k = 0; k < N; ++k)
    for (int r = 0; r < 3; ++r) out[k][r] = (gp[k][r] - gm[k][r]) / (2.0 * h);
  return out;
}

// ---- membrane.rs:147-233 --------------------------------------------------------------------
static std::pair<MembraneElement, std::vector<Vec2>> sample_element() {
  std::vector<Vec2> rest = {{0.0, 0.0}, {0.004, 0.0005}, {0.001, 0.0045}};
  return {MembraneElement::make({0, 1, 2}, rest), rest};
}
static void membrane_rest_state_is_stress_free() {
  auto [e, rest] = sample_element();
  std::vector<Vec3> x;
  for (auto& p : rest) x.push_back({p[0], p[1], 0.0});
  auto g = e.gradient(x, Material::knit_baseline());
  for (auto& v : g)
    for (double c : v) CHECK(std::fabs(c) < 1e-14, "%g", c);
}
static void membrane_gradient_matches_energy_fd() {
  auto [e, rest] = sample_element();
  Material m = Material::knit_baseline();
  Lcg rng(7);
  auto x = random_positions(rest, rng, 1.15, 0.0008);
  auto g = e.gradient(x, m);
  for (int v = 0; v < 3; ++v)
    for (int d = 0; d < 3; ++d) {
      double fd = central(x, v, d, 1e-8, [&](const std::vector<Vec3>& y) { return e.energy(y, m); });
      CHECK(std::fabs(fd - g[v][d]) <= 1e-6 * (1.0 + std::fabs(g[v][d])), "v%d d%d %g vs %g", v, d, fd, g[v][d]);
    }
}
static void membrane_hessian_fd(bool exact, double wale, double course, uint64_t seed) {
  auto [e, rest] = sample_element();
  Material m = Material::knit_baseline();
  Lcg rng(seed);
  auto x = stretched(rest, wale, course, rng, 1e-5);
  auto h = exact ? e.tangent(x, m) : e.hessian(x, m);
  for (int l = 0; l < 3; ++l)
    for (int d = 0; d < 3; ++d) {
      auto fd = gradient_derivative<3>(x, l, d, 1e-8, [&](const std::vector<Vec3>& y) { return e.gradient(y, m); });
      for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) {
          double ex = h[k][l][r][d];
          CHECK(std::fabs(fd[k][r] - ex) <= 1e-5 * (1.0 + std::fabs(ex)), "k%d l%d r%d d%d", k, l, r, d);
        }
    }
}
static void membrane_hessian_matches_fd_in_tension() { membrane_hessian_fd(false, 1.04, 1.07, 3); }
static void membrane_exact_tangent_matches_fd_in_compression() { membrane_hessian_fd(true, 0.93, 1.05, 11); }
static void membrane_energy_change_equals_energy_difference() {
  auto [e, rest] = sample_element();
  Material m = Material::knit_baseline();
  Lcg r5(5), r6(6);
  auto x = random_positions(rest, r5, 1.1, 0.0005);
  auto d = random_positions(rest, r6, 0.0, 0.0003);
  std::vector<Vec3> moved;
  for (size_t i = 0; i < x.size(); ++i) moved.push_back(add(x[i], scale(d[i], 0.7)));
  double direct = e.energy(moved, m) - e.energy(x, m);
  double change = e.energy_change(x, d, 0.7, m);
  CHECK(std::fabs(direct - change) <= 1e-12 * std::fmax(std::fabs(direct), e.energy(x, m)), "%g vs %g", direct, change);
}
static void membrane_rigid_rotation_is_stress_free() {
  auto [e, rest] = sample_element();
  double c = 0.6, s = 0.8;
  std::vector<Vec3> x;
  for (auto& p : rest) x.push_back({c * p[0] - s * p[1], s * p[0] + c * p[1], 0.3});
  for (double v : e.strain(x)) CHECK(std::fabs(v) < 1e-13, "%g", v);
}

// ---- hinge.rs:122-206, dual.rs:96-102 -------------------------------------------------------
static std::pair<HingeElement, std::vector<Vec3>> bent_hinge(double height) {
  std::vector<Vec3> x = {{0.5, 0.9, height}, {0.45, -1.1, -0.2 * height}, {0.0, 0.0, 0.0}, {1.0, 0.05, 0.0}};
  return {HingeElement{{0, 1, 2, 3}, 1.3}, x};
}
static void hinge_flat_has_zero_angle() {
  auto [h, x] = bent_hinge(0.0);
  CHECK(std::fabs(h.angle(x)) < 1e-14, "%g", h.angle(x));
}
static void hinge_angle_gradient_matches_fd() {
  auto [h, x] = bent_hinge(0.4);
  auto g = h.angle_gradient(x);
  for (int v = 0; v < 4; ++v)
    for (int d = 0; d < 3; ++d) {
      double fd = central(x, v, d, 1e-7, [&](const std::vector<Vec3>& y) { return h.angle(y); });
      CHECK(std::fabs(fd - g[v][d]) < 1e-6, "v%d d%d %g vs %g", v, d, fd, g[v][d]);
    }
}
static void hinge_generic_gradient_matches_f64() {
  auto [h, x] = bent_hinge(0.4);
  std::array<std::array<Dual, 3>, 4> p;
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) p[k][d] = {x[k][d], 0.0};
  auto generic = angle_gradient_of<Dual>(p);
  auto plain = h.angle_gradient(x);
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) CHECK(generic[k][d].v == plain[k][d], "k%d d%d", k, d);
}
static void hinge_exact_tangent_matches_fd() {
  auto [h, x] = bent_hinge(0.4);
  auto t = h.tangent(x);
  for (int l = 0; l < 4; ++l)
    for (int d = 0; d < 3; ++d) {
      auto fd = gradient_derivative<4>(x, l, d, 1e-7, [&](const std::vector<Vec3>& y) { return h.gradient(y); });
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r)
          CHECK(std::fabs(fd[k][r] - t[k][l][r][d]) < 1e-6 * (1.0 + std::fabs(t[k][l][r][d])), "k%d l%d r%d d%d", k, l, r, d);
    }
}
static void hinge_angle_hessian_is_symmetric() {
  auto [h, x] = bent_hinge(0.3);
  auto a = h.angle_hessian(x);
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) CHECK(std::fabs(a[k][l][r][c] - a[l][k][c][r]) < 1e-10, "k%d l%d", k, l);
}
static void hinge_energy_change_equals_difference() {
  auto [h, x] = bent_hinge(0.4);
  std::vector<Vec3> d = {{0.01, -0.02, 0.03}, {0.0, 0.01, -0.02}, {0.005, 0.0, 0.01}, {-0.01, 0.02, 0.0}};
  std::vector<Vec3> moved;
  for (int i = 0; i < 4; ++i) moved.push_back(add(x[i], scale(d[i], 0.6)));
  double direct = h.energy(moved) - h.energy(x);
  CHECK(std::fabs(h.energy_change(x, d, 0.6) - direct) < 1e-14, "%g", direct);
}
static void hinge_gauss_newton_exact_at_flat() {
  auto [h, x] = bent_hinge(0.0);
  auto gn = h.hessian(x);
  for (int l = 0; l < 4; ++l)
    for (int d = 0; d < 3; ++d) {
      auto fd = gradient_derivative<4>(x, l, d, 1e-7, [&](const std::vector<Vec3>& y) { return h.gradient(y); });
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) CHECK(std::fabs(fd[k][r] - gn[k][l][r][d]) < 1e-5, "k%d l%d", k, l);
    }
}
static void dual_derivatives_match_calculus() {
  Dual x{2.0, 1.0};
  Dual y = (x * x + Dual{3.0, 0.0}) / dsqrt(x);
  double exact = (2.0 * 2.0 * std::sqrt(2.0) - (4.0 + 3.0) / (2.0 * std::sqrt(2.0))) / 2.0;
  CHECK(std::fabs(y.t - exact) < 1e-14, "%g vs %g", y.t, exact);
}

// ---- contact.rs:386-440 ---------------------------------------------------------------------
static Contact test_contact() {
  Contact c;
  c.colliders = {{1, {0.0, 0.0, 1.0}, 0.0}, {2, {0.25, 0.25, 0.125}, 0.125}};
  c.barrier = {5e-4, 2.0};
  return c;
}
static void barrier_derivatives_match_fd() {
  Barrier b{5e-4, 3.0};
  for (double d : {3e-5, 1e-4, 2.5e-4, 4.9e-4}) {
    double h = 1e-9;
    double first = (b.value(d + h) - b.value(d - h)) / (2.0 * h);
    double second = (b.first(d + h) - b.first(d - h)) / (2.0 * h);
    CHECK(std::fabs(first - b.first(d)) <= 1e-6 * std::fabs(b.first(d)), "%g", d);
    CHECK(std::fabs(second - b.second(d)) <= 1e-6 * std::fabs(b.second(d)), "%g", d);
  }
}
static void resting_kappa_balances_gravity() {
  double kappa = Barrier::resting_kappa(5e-4, 2e-5, 9.81);
  CHECK(std::fabs(Barrier{5e-4, kappa}.first(2.5e-4) + 2e-5 * 9.81) < 1e-18, "%g", kappa);
}
// e5_checks.rs:233-241: the contact state puts every free vertex inside the barrier (in [0.05 dhat, dhat)) and
// some on each collider.
static void contact_state_inside_the_barrier_on_both_colliders() {
  Drape d(reference_scene(9, true, 1.0, 1.0, 1.0, false));
  auto x = contact_state(d);
  const double dhat = d.contact.barrier.dhat;
  int sphere = 0, floor = 0;
  for (int i : d.free_) {
    const double m = d.contact.min_distance(x[i]);
    CHECK(m >= 0.05 * dhat * 0.999 && m < dhat, "vertex %d at %g", i, m);
    for (const auto& c : d.contact.colliders)
      if (c.distance(x[i]) < dhat) (c.kind == 2 ? sphere : floor)++;
  }
  CHECK(sphere > 0 && floor > 0, "%d %d", sphere, floor);
}

// e5_checks.rs:109-200 contact_derivative_check, patch directions with the Phi step scaled to the patch's
// smallest distance (patch_scaled_phi_step): Phi gradient and Hessian, barrier gradient and Hessian, barrier
// energy-change identity.
namespace e5fd {
constexpr double FD_STEP = 1e-7, BARRIER_FD_FRACTION = 1e-3;
constexpr int DIRECTIONS = 6, PATCH = 16;
std::vector<Vec3> shifted(const std::vector<Vec3>& x, const std::vector<Vec3>& v, double h) {
  std::vector<Vec3> out(x.size());
  for (size_t i = 0; i < x.size(); ++i)
    for (int k = 0; k < 3; ++k) out[i][k] = x[i][k] + h * v[i][k];
  return out;
}
double local_value(const DrapeStep& s, const std::vector<Vec3>& x, const std::vector<int>& patch) {  // drape.rs:273-283
  const Drape& d = *s.drape;
  std::vector<bool> touched(d.mesh.rest.size(), false);
  for (int i : patch) touched[i] = true;
  double inertia = 0.0;
  for (int i : patch) {
    double o2 = 0.0;
    for (int k = 0; k < 3; ++k) o2 += (x[i][k] - s.x_hat[i][k]) * (x[i][k] - s.x_hat[i][k]);
    inertia += 0.5 * d.mass_over_h2(i) * o2 + d.contact.energy(x[i]);
  }
  double membrane = 0.0, hinge = 0.0;
  for (const auto& e : d.membranes)
    if (touched[e.vertices[0]] || touched[e.vertices[1]] || touched[e.vertices[2]]) membrane += e.energy(x, d.material);
  for (const auto& h : d.hinges)
    if (touched[h.vertices[0]] || touched[h.vertices[1]] || touched[h.vertices[2]] || touched[h.vertices[3]])
      hinge += h.energy(x);
  return inertia + membrane + hinge;
}
std::vector<double> barrier_rows(const Drape& d, const std::vector<Vec3>& x) {
  std::vector<double> g;
  for (int i : d.free_) {
    Vec3 v = d.contact.gradient(x[i]);
    g.insert(g.end(), v.begin(), v.end());
  }
  return g;
}
std::vector<double> barrier_hessian_times(const Drape& d, const std::vector<Vec3>& x, const std::vector<Vec3>& v) {
  std::vector<double> out;
  for (int i : d.free_) {
    Mat3 h = d.contact.hessian_with(x[i], true);
    for (int r = 0; r < 3; ++r) out.push_back(h[r][0] * v[i][0] + h[r][1] * v[i][1] + h[r][2] * v[i][2]);
  }
  return out;
}
double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double directional(const std::function<double(const std::vector<Vec3>&)>& f, const std::vector<double>& g,
                   const std::vector<double>& rows, const std::vector<Vec3>& x, const std::vector<Vec3>& v, double h) {
  const double fd = (f(shifted(x, v, h)) - f(shifted(x, v, -h))) / (2.0 * h);
  const double exact = dot(g, rows);
  return std::fabs(fd - exact) / std::fabs(exact);
}
double hessian(const std::function<std::vector<double>(const std::vector<Vec3>&)>& g, const std::vector<double>& exact,
               const std::vector<Vec3>& x, const std::vector<Vec3>& v, double h) {
  auto p = g(shifted(x, v, h)), m = g(shifted(x, v, -h));
  std::vector<double> fd(p.size());
  for (size_t i = 0; i < p.size(); ++i) fd[i] = (p[i] - m[i]) / (2.0 * h);
  return difference_norm(fd, exact) / vector_norm(exact);
}
}  // namespace e5fd

static void active_barrier_derivatives_match_fd() {
  using namespace e5fd;
  Drape d(reference_scene(9, true, 1.0, 1.0, 1.0, false));
  Lcg rng(59);
  const auto x = contact_state(d);
  std::vector<Vec3> x_hat = x;
  for (auto& p : x_hat)
    for (double& v : p) v += 1e-4 * rng.uniform();
  DrapeStep step{&d, x_hat};
  const int span = (int)d.free_.size() - PATCH;
  double worst[5] = {0, 0, 0, 0, 0}, worst_fixed[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < DIRECTIONS; ++k) {
    const int start = (int)(((rng.uniform() + 1.0) / 2.0) * span);
    std::vector<int> patch(d.free_.begin() + start, d.free_.begin() + start + PATCH);
    std::vector<Vec3> v(d.mesh.rest.size(), Vec3{0, 0, 0});
    for (int i : patch) v[i] = {rng.uniform(), rng.uniform(), rng.uniform()};
    const auto rows = step.gather(v);
    double smallest = INFINITY;
    for (int i : patch) smallest = std::fmin(smallest, d.contact.min_distance(x[i]));
    const double barrier_step = BARRIER_FD_FRACTION * smallest;
    auto local_barrier = [&](const std::vector<Vec3>& z) {
      double e = 0.0;
      for (int i : patch) e += d.contact.energy(z[i]);
      return e;
    };
    const double alpha = 0.5 * std::fmin(step.max_step(x, v), d.contact.barrier.dhat);
    const double direct = local_barrier(shifted(x, v, alpha)) - local_barrier(x);
    const double identity = std::fabs(step.barrier_change(x, v, alpha) - direct) / std::fabs(direct);
    for (int scaled = 0; scaled < 2; ++scaled) {
      const double phi_step = scaled ? barrier_step : FD_STEP;
      const double e[5] = {
          directional([&](const std::vector<Vec3>& z) { return local_value(step, z, patch); }, step.residual(x), rows, x,
                      v, phi_step),
          hessian([&](const std::vector<Vec3>& z) { return step.residual(z); }, step.tangent(x).multiply(rows), x, v,
                  phi_step),
          directional(local_barrier, barrier_rows(d, x), rows, x, v, barrier_step),
          hessian([&](const std::vector<Vec3>& z) { return barrier_rows(d, z); }, barrier_hessian_times(d, x, v), x, v,
                  barrier_step),
          identity};
      double* w = scaled ? worst : worst_fixed;
      for (int j = 0; j < 5; ++j) w[j] = std::fmax(w[j], e[j]);
    }
  }
  CHECK(worst[0] <= 1e-6 && worst[2] <= 1e-6, "%g %g", worst[0], worst[2]);
  CHECK(worst[1] <= 1e-5 && worst[3] <= 1e-5, "%g %g", worst[1], worst[3]);
  CHECK(worst[4] <= 1e-12, "%g", worst[4]);
  CHECK(worst_fixed[2] <= 1e-6 && worst_fixed[3] <= 1e-5, "%g %g", worst_fixed[2], worst_fixed[3]);
}

// newton.rs:302-314 (armijo_backtracks_an_overlong_newton_step), restated on the drape step system (the
// reference's test uses its E2 coupon, which is not restated here): four times the Newton direction of the
// first backward-Euler step is backtracked at least once, and the accepted step satisfies Armijo.
static void armijo_backtracks_an_overlong_newton_step() {
  Drape d(reference_scene(9, false, 1.0, 1.0, 1.0, false));
  std::vector<Vec3> x = d.rest_positions(), v(x.size(), Vec3{0, 0, 0});
  DrapeStep step{&d, d.prediction(x, v)};
  const auto start = step.x_hat;
  const auto g = step.residual(start);
  BandCholesky f;
  CHECK(f.factor(step.regularized_tangent(start, 0.0)), "factor");
  std::vector<double> neg(g.size());
  for (size_t i = 0; i < g.size(); ++i) neg[i] = -4.0 * g[i];
  const auto direction = step.scatter(f.solve(neg));
  LineSearch ls;
  NewtonSettings settings{1e-12, 30, 1e-4, 30};
  CHECK(line_search(step, start, g, direction, settings, &ls), "line search failed");
  CHECK(ls.halvings >= 1, "%d", ls.halvings);
  const double slope = e5fd::dot(g, step.gather(direction));
  CHECK(step.energy_change(start, direction, ls.step_length) <= 1e-4 * ls.step_length * slope, "armijo");
}

static void contact_gradient_and_hessian_match_fd() {
  Contact c = test_contact();
  Vec3 x = {0.27, 0.24, 0.2502};
  double h = 1e-9;
  for (int k = 0; k < 3; ++k) {
    Vec3 plus = x, minus = x;
    plus[k] += h;
    minus[k] -= h;
    double fd = (c.energy(plus) - c.energy(minus)) / (2.0 * h);
    CHECK(std::fabs(fd - c.gradient(x)[k]) <= 1e-6 * std::fmax(std::fabs(c.gradient(x)[k]), 1e-9), "%d", k);
    for (int i = 0; i < 3; ++i) {
      double col = (c.gradient(plus)[i] - c.gradient(minus)[i]) / (2.0 * h);
      CHECK(std::fabs(col - c.hessian_with(x, true)[i][k]) <= 1e-5 * std::fabs(c.hessian_with(x, true)[k][k]), "%d%d", i, k);
    }
  }
}
static void max_step_keeps_a_tenth() {
  Contact c = test_contact();
  Vec3 x = {0.4, 0.4, 2e-4};
  double alpha = c.max_step(x, {0.0, 0.0, -1e-3});
  CHECK(std::fabs(x[2] - alpha * 1e-3 - 2e-5) < 1e-15, "%g", alpha);
  Vec3 top = {0.25, 0.25, 0.25 + 1e-4};
  double beta = c.max_step(top, {0.0, 0.0, -1e-3});
  CHECK(std::fabs(c.colliders[1].distance({0.25, 0.25, top[2] - beta * 1e-3}) - 1e-5) < 1e-14, "%g", beta);
  CHECK(c.max_step(top, {0.0, 0.0, 1e-3}) == 1.0, "upward");
}

// ---- drape.rs:441-488 -----------------------------------------------------------------------
static void residual_is_gradient_of_incremental_potential() {
  Drape d(reference_scene(7, false, 1.0, 1.0, 1.0, false));
  auto x = d.rest_positions();
  std::vector<Vec3> v;
  for (size_t i = 0; i < x.size(); ++i) v.push_back({0.1 * (double)(i % 3), -0.05, 0.2});
  DrapeStep step{&d, d.prediction(x, v)};
  auto y = x;
  for (size_t k = 0; k < d.free_.size(); ++k) y[d.free_[k]][2] += 1e-3 * ((double)(k % 5) - 2.0);
  auto residual = step.residual(y);
  for (size_t row = 0; row < d.free_.size(); row += 4)
    for (int dd = 0; dd < 3; ++dd) {
      double fd = central(y, d.free_[row], dd, 1e-7, [&](const std::vector<Vec3>& z) { return step.value(z); });
      double ex = residual[3 * row + dd];
      CHECK(std::fabs(fd - ex) <= 1e-6 * (1.0 + std::fabs(ex)), "row %zu d%d: %g vs %g", row, dd, fd, ex);
    }
}
static void feasible_start_keeps_a_tenth() {
  Drape d(reference_scene(9, true, 1.0, 1.0, 1.0, false));
  std::vector<Vec3> x;
  for (auto& p : d.rest_positions()) x.push_back({p[0], p[1], 0.2501});
  std::vector<Vec3> v(x.size(), Vec3{0.0, 0.0, -2.0});
  DrapeStep step{&d, d.prediction(x, v)};
  auto start = feasible_start(step, x);
  double kept = INFINITY;
  for (int i : d.free_) kept = std::fmin(kept, d.contact.min_distance(start[i]) / d.contact.min_distance(x[i]));
  CHECK(std::fabs(kept - 0.1) < 1e-9, "%g", kept);
}
static void first_step_falls_h2g_and_energy_decreases() {
  SceneSpec s = reference_scene(9, false, 1.0, 1.0, 1.0, false);
  s.exact_tangent = false;
  Drape d(s);
  NewtonSettings settings{1e-10, 30, 1e-4, 30};
  auto x = d.rest_positions();
  std::vector<Vec3> v(x.size(), Vec3{0, 0, 0});
  std::vector<double> energies = {d.total_energy(x, v)};
  std::vector<std::vector<Vec3>> positions = {x};
  for (int n = 0; n < 3; ++n) {
    TimeStep t = time_step(d, x, v, settings, false, 0.0);
    x = t.solution.positions;
    v = t.velocity;
    positions.push_back(x);
    energies.push_back(d.total_energy(x, v));
  }
  size_t centre = d.mesh.rest.size() / 2;
  double drop = positions[1][centre][2];
  double h = 1.0 / 240.0;
  CHECK(std::fabs(drop + h * h * 9.81) < 1e-6, "%g", drop);
  for (int n = 0; n < 3; ++n) CHECK(energies[n + 1] <= energies[n], "%d: %g > %g", n, energies[n + 1], energies[n]);
}

// ---- newton.rs:310-318, e4_rules.rs:398-430 -------------------------------------------------
static void regularization_schedule() {
  CHECK(raise_mu(0.0) == MU_START, "0");
  CHECK(raise_mu(MU_START) == 10.0 * MU_START, "start");
  CHECK(raise_mu(500.0) == MU_CAP, "cap");
  CHECK(std::fabs(relax_mu(0.1) - 0.01) < 1e-18, "relax");
  CHECK(relax_mu(MU_START) == 0.0, "to zero");
  CHECK(relax_mu(0.0) == 0.0, "zero");
}
static void forcing_follows_squared_ratio() {
  double e0 = 1e-2;
  double e1 = next_forcing(e0, 1e-1 / 1.0);
  double e2 = next_forcing(e1, 1e-3 / 1e-1);
  double e3 = next_forcing(e2, 1e-6 / 1e-3);
  CHECK(std::fabs(e1 - 9e-3) < 1e-15, "%g", e1);
  CHECK(std::fabs(e2 - 9e-5) < 1e-17, "%g", e2);
  CHECK(std::fabs(e3 - 9e-7) < 1e-19, "%g", e3);
}
static void forcing_is_clamped() {
  CHECK(next_forcing(1e-2, 0.9) == 1e-2, "max");
  CHECK(next_forcing(1e-2, 1e-6) == 1e-7, "min");
}
static void update_rule_compares_with_step() {
  CHECK(update_is_small(1e-9, 1e-4, 0.0), "small");
  CHECK(!update_is_small(2e-8, 1e-4, 0.0), "not small");
  CHECK(update_is_small(1e-7, 1e-4, f32_spacing(0.5)), "resolution");
}
static void f32_spacing_reads_the_binade() {
  CHECK(f32_spacing(0.5) == std::ldexp(1.0, -24), "0.5");
  CHECK(f32_spacing(0.49999997) == std::ldexp(1.0, -25), "below 0.5");
  CHECK(f32_spacing(1.0) == std::ldexp(1.0, -23), "1");
}

// ---- pcg.rs:494-534, band.rs:198-211, cases.rs:406-440 --------------------------------------
static void cholesky_solves_tension_case() {
  CaseProblem p = build_case(Tension, 11);
  BandCholesky f;
  CHECK(f.factor(p.matrix), "factor");
  auto x = f.solve(p.rhs);
  double res = difference_norm(p.matrix.multiply(x), p.rhs) / vector_norm(p.rhs);
  CHECK(res < 1e-12, "%g", res);
}
static void f64_pcg_matches_cholesky() {
  CaseProblem p = build_case(Tension, 13);
  BandCholesky f;
  f.factor(p.matrix);
  auto exact = f.solve(p.rhs);
  PcgReport r = pcg_solve<double>(p.matrix, p.rhs, {1e-12, 5000, 10, 0, 0.0});
  CHECK(r.converged, "converged");
  double err = difference_norm(r.solution, exact) / vector_norm(exact);
  CHECK(err < 1e-9, "%g", err);
}
static void residual_replacement_keeps_f32_pcg_on_true_residual() {
  CaseProblem p = build_case(Tension, 33);
  auto true_res = [&](const PcgReport& r) { return difference_norm(p.matrix.multiply(r.solution), p.rhs) / vector_norm(p.rhs); };
  PcgReport plain = pcg_solve<float>(p.matrix, p.rhs, {1e-7, 5000, 10, 0, 1e-2});
  PcgReport replaced = pcg_solve<float>(p.matrix, p.rhs, {1e-7, 5000, 10, 10, 1e-2});
  CHECK(replaced.converged, "converged");
  CHECK(true_res(replaced) < true_res(plain), "%g vs %g", true_res(replaced), true_res(plain));
}
static void inverse_times_matrix_is_identity() {
  double m[9] = {4.0, 1.0, 0.5, 1.0, 3.0, 0.2, 0.5, 0.2, 2.0};
  auto inv = invert3<double>(m);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += m[3 * r + k] * inv[3 * k + c];
      CHECK(std::fabs(v - (r == c ? 1.0 : 0.0)) < 1e-14, "%d%d", r, c);
    }
}
static void assembled_matrices_are_symmetric() {
  for (CaseKind k : {Tension, Slack, Dynamic}) {
    CaseProblem p = build_case(k, 9);
    CHECK(p.matrix.asymmetry() < 1e-12, "kind %d", (int)k);
  }
}
static void uniform_prestretch_has_no_interior_force() {
  Mesh mesh = Mesh::grid(7, 7, 0.5, 0.5);
  Material m = Material::knit_baseline();
  std::vector<Vec3> x;
  for (auto& p : mesh.rest) x.push_back({1.03 * p[0], 1.08 * p[1], 0.0});
  std::vector<Vec3> force(mesh.rest.size(), Vec3{0, 0, 0});
  for (auto& t : mesh.triangles) {
    auto g = MembraneElement::make(t, mesh.rest).gradient(x, m);
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) force[t[k]][d] += g[k][d];
  }
  for (int v = 0; v < (int)mesh.rest.size(); ++v)
    if (!mesh.is_boundary(v))
      for (double f : force[v]) CHECK(std::fabs(f) < 1e-12, "v%d %g", v, f);
}
static void material_is_convex() {
  CHECK(Material::knit_baseline().is_convex(), "knit");
  CHECK(Material::isotropic(1e6, 0.3, 5e-4, 1e-5, 0.2).is_convex(), "isotropic");
}

// ---- e7_ptc.rs:191-242 ----------------------------------------------------------------------
static void ptc_step_size_rule() {
  PtcRules r;
  CHECK(next_step_size(r, 0.01, 3) == 0.02, "double");
  CHECK(next_step_size(r, 0.01, r.inner_max) == 0.005, "halve");
  CHECK(next_step_size(r, 0.9, 1) == r.h_max, "max");
  CHECK(next_step_size(r, r.h0 * r.h_min_factor, r.inner_max) == r.h0 * r.h_min_factor, "min");
}
static void ptc_start_residual_is_static_residual() {
  Drape d(reference_scene(9, true, 1.0, 1.0, 1.0, false));
  std::vector<Vec3> x;
  for (auto& p : d.rest_positions()) x.push_back({p[0], p[1], p[2] + 0.002 * std::sin(p[0] * 31.0)});
  d.time_step = 0.05;
  DrapeStep system{&d, ptc_prediction(d, x, 0.05)};
  auto dynamic = system.residual(x);
  auto statics = static_gradient(d, x);
  for (size_t i = 0; i < dynamic.size(); ++i)
    CHECK(std::fabs(dynamic[i] - statics[i]) <= 1e-12 * (1.0 + std::fabs(statics[i])), "%zu", i);
}
static void ptc_never_raises_static_energy() {
  Drape d(reference_scene(9, true, 1.0, 1.0, 1.0, false));
  PtcRules r;
  PtcRun run = ptc_solve(d, d.rest_positions(), r.h0, r, 6, true, nullptr);
  CHECK(run.steps.size() == 6, "%zu steps", run.steps.size());
  for (size_t k = 0; k + 1 < run.states.size(); ++k) {
    double a = static_energy(d, run.states[k]), b = static_energy(d, run.states[k + 1]);
    CHECK(b <= a + 1e-15, "step %zu: %g > %g", k, b, a);
  }
}
static void ptc_failed_attempt_classification() {
  std::vector<Vec3> finite = {{0.0, 0.0, 0.0}};
  CHECK(ptc_failed(RegularizationCap, 0, finite), "cap before update");
  CHECK(!ptc_failed(RegularizationCap, 2, finite), "cap after updates");
  CHECK(!ptc_failed(Converged, 0, finite), "converged");
  CHECK(ptc_failed(Converged, 3, {{NAN, 0.0, 0.0}}), "nan");
}

// ---- irregular mesh (net-new; pinned by invariants) -----------------------------------------
static void irregular_mesh_is_valid() {
  Mesh m = Mesh::irregular(17, 0.5, 0.2, 0.1, 5);
  Mesh g = Mesh::grid(17, 17, 0.5, 0.5);
  CHECK(m.triangles.size() == g.triangles.size(), "triangle count");
  CHECK(m.hinges.size() == g.hinges.size(), "interior edges of a triangulated grid: %zu vs %zu", m.hinges.size(), g.hinges.size());
  double total = 0.0;
  for (auto& t : m.triangles) {
    double a = rest_area(m.rest, t);
    CHECK(a > 0.0, "degenerate");
    total += a;
  }
  CHECK(std::fabs(total - 0.25) < 1e-12, "area %g", total);
  int flipped = 0;
  for (size_t c = 0; c < g.triangles.size(); c += 2) flipped += m.triangles[c] != g.triangles[c];
  CHECK(flipped > 0 && flipped < (int)g.triangles.size() / 2 / 4, "flips %d", flipped);
}

int main(int argc, char** argv) {
  std::vector<std::pair<const char*, void (*)()>> tests = {
      {"membrane_rest_state_is_stress_free", membrane_rest_state_is_stress_free},
      {"membrane_gradient_matches_energy_fd", membrane_gradient_matches_energy_fd},
      {"membrane_hessian_matches_fd_in_tension", membrane_hessian_matches_fd_in_tension},
      {"membrane_exact_tangent_matches_fd_in_compression", membrane_exact_tangent_matches_fd_in_compression},
      {"membrane_energy_change_equals_energy_difference", membrane_energy_change_equals_energy_difference},
      {"membrane_rigid_rotation_is_stress_free", membrane_rigid_rotation_is_stress_free},
      {"hinge_flat_has_zero_angle", hinge_flat_has_zero_angle},
      {"hinge_angle_gradient_matches_fd", hinge_angle_gradient_matches_fd},
      {"hinge_generic_gradient_matches_f64", hinge_generic_gradient_matches_f64},
      {"hinge_exact_tangent_matches_fd", hinge_exact_tangent_matches_fd},
      {"hinge_angle_hessian_is_symmetric", hinge_angle_hessian_is_symmetric},
      {"hinge_energy_change_equals_difference", hinge_energy_change_equals_difference},
      {"hinge_gauss_newton_exact_at_flat", hinge_gauss_newton_exact_at_flat},
      {"dual_derivatives_match_calculus", dual_derivatives_match_calculus},
      {"barrier_derivatives_match_fd", barrier_derivatives_match_fd},
      {"resting_kappa_balances_gravity", resting_kappa_balances_gravity},
      {"contact_gradient_and_hessian_match_fd", contact_gradient_and_hessian_match_fd},
      {"contact_state_inside_the_barrier_on_both_colliders", contact_state_inside_the_barrier_on_both_colliders},
      {"active_barrier_derivatives_match_fd", active_barrier_derivatives_match_fd},
      {"armijo_backtracks_an_overlong_newton_step", armijo_backtracks_an_overlong_newton_step},
      {"max_step_keeps_a_tenth", max_step_keeps_a_tenth},
      {"residual_is_gradient_of_incremental_potential", residual_is_gradient_of_incremental_potential},
      {"feasible_start_keeps_a_tenth", feasible_start_keeps_a_tenth},
      {"first_step_falls_h2g_and_energy_decreases", first_step_falls_h2g_and_energy_decreases},
      {"regularization_schedule", regularization_schedule},
      {"forcing_follows_squared_ratio", forcing_follows_squared_ratio},
      {"forcing_is_clamped", forcing_is_clamped},
      {"update_rule_compares_with_step", update_rule_compares_with_step},
      {"f32_spacing_reads_the_binade", f32_spacing_reads_the_binade},
      {"cholesky_solves_tension_case", cholesky_solves_tension_case},
      {"f64_pcg_matches_cholesky", f64_pcg_matches_cholesky},
      {"residual_replacement_keeps_f32_pcg_on_true_residual", residual_replacement_keeps_f32_pcg_on_true_residual},
      {"inverse_times_matrix_is_identity", inverse_times_matrix_is_identity},
      {"assembled_matrices_are_symmetric", assembled_matrices_are_symmetric},
      {"uniform_prestretch_has_no_interior_force", uniform_prestretch_has_no_interior_force},
      {"material_is_convex", material_is_convex},
      {"ptc_step_size_rule", ptc_step_size_rule},
      {"ptc_start_residual_is_static_residual", ptc_start_residual_is_static_residual},
      {"ptc_never_raises_static_energy", ptc_never_raises_static_energy},
      {"ptc_failed_attempt_classification", ptc_failed_attempt_classification},
      {"irregular_mesh_is_valid", irregular_mesh_is_valid},
  };
  if (argc > 1 && std::strcmp(argv[1], "--list") == 0) {
    for (auto& t : tests) std::printf("%s\n", t.first);
    return 0;
  }
  int failed = 0;
  for (auto& t : tests) {
    if (argc > 1 && std::strcmp(argv[1], t.first) != 0) continue;
    try {
      t.second();
      std::printf("PASS %s\n", t.first);
    } catch (const Failure& f) {
      std::printf("FAIL %s: %s\n", t.first, f.what.c_str());
      ++failed;
    }
  }
  return failed ? 1 : 0;
}
