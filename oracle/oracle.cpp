// Oracle implementation. Test infrastructure only (see oracle.hpp). Each block cites the Rust
// file:line of engine/oracle/src it restates; operation order is kept term by term.
#include "oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>

namespace oracle {

static const double PI = 3.14159265358979323846;
static const double EPS64 = std::numeric_limits<double>::epsilon();

// ---- linalg.rs:54-60, coupon.rs:248-250 -----------------------------------------------------
double vector_norm(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}
double difference_norm(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
  return std::sqrt(s);
}
double free_dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i + 2 < a.size(); i += 3) s += dot({a[i], a[i + 1], a[i + 2]}, {b[i], b[i + 1], b[i + 2]});
  return s;
}

// ---- mesh.rs:14-107 -------------------------------------------------------------------------
static std::array<std::array<int, 3>, 2> cell_triangles(int columns, int i, int j, bool flip) {
  int v00 = j * columns + i, v10 = v00 + 1, v01 = v00 + columns, v11 = v01 + 1;
  bool even = ((i + j) % 2 == 0) != flip;
  if (even) return {{{v00, v10, v11}, {v00, v11, v01}}};
  return {{{v00, v10, v01}, {v10, v11, v01}}};
}

static double cross2(Vec2 a, Vec2 b) { return a[0] * b[1] - a[1] * b[0]; }
static Vec2 sub2(Vec2 a, Vec2 b) { return {a[0] - b[0], a[1] - b[1]}; }

double rest_area(const std::vector<Vec2>& rest, const std::array<int, 3>& t) {
  Vec2 a = rest[t[0]], b = rest[t[1]], c = rest[t[2]];
  return 0.5 * std::fabs(cross2(sub2(b, a), sub2(c, a)));
}

std::vector<std::array<int, 4>> interior_hinges(const std::vector<Vec2>& rest,
                                                const std::vector<std::array<int, 3>>& tris) {
  std::map<std::pair<int, int>, std::vector<int>> edges;
  for (const auto& t : tris)
    for (int k = 0; k < 3; ++k) {
      int a = t[k], b = t[(k + 1) % 3], opp = t[(k + 2) % 3];
      edges[{std::min(a, b), std::max(a, b)}].push_back(opp);
    }
  std::vector<std::array<int, 4>> hinges;
  for (const auto& [key, opposites] : edges) {
    if (opposites.size() != 2) continue;
    int a = key.first, b = key.second, first = opposites[0], second = opposites[1];
    bool is_left = cross2(sub2(rest[b], rest[a]), sub2(rest[first], rest[a])) > 0.0;
    hinges.push_back(is_left ? std::array<int, 4>{first, second, a, b}
                             : std::array<int, 4>{second, first, a, b});
  }
  std::sort(hinges.begin(), hinges.end());
  return hinges;
}

Mesh Mesh::grid(int columns, int rows, double width, double height) {
  Mesh m;
  m.columns = columns; m.rows = rows; m.width = width; m.height = height;
  double dx = width / (double)(columns - 1), dy = height / (double)(rows - 1);
  for (int j = 0; j < rows; ++j)
    for (int i = 0; i < columns; ++i) m.rest.push_back({(double)i * dx, (double)j * dy});
  for (int j = 0; j + 1 < rows; ++j)
    for (int i = 0; i + 1 < columns; ++i) {
      auto two = cell_triangles(columns, i, j, false);
      m.triangles.push_back(two[0]);
      m.triangles.push_back(two[1]);
    }
  m.hinges = interior_hinges(m.rest, m.triangles);
  return m;
}

Mesh Mesh::irregular(int columns, double size, double jitter, double flip_fraction, uint64_t seed) {
  Mesh m;
  m.columns = columns; m.rows = columns; m.width = size; m.height = size;
  double dx = size / (double)(columns - 1);
  Lcg rng(seed);
  for (int j = 0; j < columns; ++j)
    for (int i = 0; i < columns; ++i) {
      Vec2 p = {(double)i * dx, (double)j * dx};
      bool boundary = i == 0 || j == 0 || i == columns - 1 || j == columns - 1;
      if (!boundary) {
        p[0] += jitter * dx * rng.uniform();
        p[1] += jitter * dx * rng.uniform();
      }
      m.rest.push_back(p);
    }
  for (int j = 0; j + 1 < columns; ++j)
    for (int i = 0; i + 1 < columns; ++i) {
      bool flip = 0.5 * (rng.uniform() + 1.0) < flip_fraction;
      auto two = cell_triangles(columns, i, j, flip);
      m.triangles.push_back(two[0]);
      m.triangles.push_back(two[1]);
    }
  m.hinges = interior_hinges(m.rest, m.triangles);
  return m;
}

bool Mesh::is_boundary(int v) const {
  int i = v % columns, j = v / columns;
  return i == 0 || j == 0 || i == columns - 1 || j == rows - 1;
}

std::vector<double> Mesh::vertex_areas() const {
  std::vector<double> areas(rest.size(), 0.0);
  for (const auto& t : triangles) {
    double area = rest_area(rest, t);
    for (int v : t) areas[v] += area / 3.0;
  }
  return areas;
}

// ---- material: isotropic plane stress mapped onto W of material.rs:39-45 ----------------------
Material Material::isotropic(double E, double nu, double t, double bend, double density) {
  double membrane = E * t / (1.0 - nu * nu);
  return {membrane, membrane, nu * membrane, E * t / (2.0 * (1.0 + nu)), bend, bend, density};
}

// ---- membrane.rs:15-140 ---------------------------------------------------------------------
MembraneElement MembraneElement::make(std::array<int, 3> v, const std::vector<Vec2>& rest) {
  Vec2 p0 = rest[v[0]], p1 = rest[v[1]], p2 = rest[v[2]];
  double dm00 = p1[0] - p0[0], dm01 = p2[0] - p0[0], dm10 = p1[1] - p0[1], dm11 = p2[1] - p0[1];
  double det = dm00 * dm11 - dm01 * dm10;
  Vec2 g1 = {dm11 / det, -dm01 / det};
  Vec2 g2 = {-dm10 / det, dm00 / det};
  Vec2 g0 = {-g1[0] - g2[0], -g1[1] - g2[1]};
  return {v, {g0, g1, g2}, 0.5 * std::fabs(det)};
}

void MembraneElement::columns(const std::vector<Vec3>& x, Vec3& wale, Vec3& course) const {
  Vec3 e1 = sub(x[vertices[1]], x[vertices[0]]), e2 = sub(x[vertices[2]], x[vertices[0]]);
  Vec2 g1 = grad_n[1], g2 = grad_n[2];
  wale = add(scale(e1, g1[0]), scale(e2, g2[0]));
  course = add(scale(e1, g1[1]), scale(e2, g2[1]));
}

std::array<double, 3> MembraneElement::strain(const std::vector<Vec3>& x) const {
  Vec3 fw, fc;
  columns(x, fw, fc);
  return {0.5 * (dot(fw, fw) - 1.0), 0.5 * (dot(fc, fc) - 1.0), 0.5 * dot(fw, fc)};
}

double MembraneElement::energy(const std::vector<Vec3>& x, const Material& m) const {
  return area * m.energy_density(strain(x));
}

std::array<Vec3, 3> MembraneElement::gradient(const std::vector<Vec3>& x, const Material& m) const {
  Vec3 fw, fc;
  columns(x, fw, fc);
  auto s = m.stress(strain(x));
  Vec3 pw = add(scale(fw, s[0]), scale(fc, s[2]));
  Vec3 pc = add(scale(fw, s[2]), scale(fc, s[1]));
  std::array<Vec3, 3> g;
  for (int k = 0; k < 3; ++k) g[k] = scale(add(scale(pw, grad_n[k][0]), scale(pc, grad_n[k][1])), area);
  return g;
}

static std::array<Vec3, 3> strain_vectors(Vec2 g, Vec3 fw, Vec3 fc) {
  Vec3 shear = scale(add(scale(fw, g[1]), scale(fc, g[0])), 0.5);
  return {scale(fw, g[0]), scale(fc, g[1]), shear};
}

static double geometric_factor(Vec2 gk, Vec2 gl, std::array<double, 3> s) {
  return gk[0] * (s[0] * gl[0] + s[2] * gl[1]) + gk[1] * (s[2] * gl[0] + s[1] * gl[1]);
}

ElementHessian MembraneElement::hessian_with_stress(const std::vector<Vec3>& x, const Material& m,
                                                    std::array<double, 3> s) const {
  Vec3 fw, fc;
  columns(x, fw, fc);
  std::array<std::array<Vec3, 3>, 3> rows;
  for (int k = 0; k < 3; ++k) rows[k] = strain_vectors(grad_n[k], fw, fc);
  ElementHessian h;
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      const auto &rk = rows[k], &rl = rows[l];
      Mat3 b = identity_scaled(geometric_factor(grad_n[k], grad_n[l], s));
      mat_add_scaled(b, outer(rk[0], rl[0]), m.wale);
      mat_add_scaled(b, outer(rk[1], rl[1]), m.course);
      mat_add_scaled(b, outer(rk[0], rl[1]), m.coupling);
      mat_add_scaled(b, outer(rk[1], rl[0]), m.coupling);
      mat_add_scaled(b, outer(rk[2], rl[2]), 4.0 * m.shear);
      for (auto& row : b)
        for (auto& v : row) v = v * area;
      h[k][l] = b;
    }
  return h;
}

ElementHessian MembraneElement::hessian(const std::vector<Vec3>& x, const Material& m) const {
  return hessian_with_stress(x, m, clamp_negative_stress(m.stress(strain(x))));
}
ElementHessian MembraneElement::tangent(const std::vector<Vec3>& x, const Material& m) const {
  return hessian_with_stress(x, m, m.stress(strain(x)));
}

double MembraneElement::energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d,
                                      double alpha, const Material& m) const {
  Vec3 fw, fc, dw, dc;
  columns(x, fw, fc);
  columns(d, dw, dc);
  dw = scale(dw, alpha);
  dc = scale(dc, alpha);
  std::array<double, 3> change = {dot(fw, dw) + 0.5 * dot(dw, dw), dot(fc, dc) + 0.5 * dot(dc, dc),
                                  0.5 * (dot(fw, dc) + dot(dw, fc) + dot(dw, dc))};
  auto s = m.stress(strain(x));
  double first = s[0] * change[0] + s[1] * change[1] + 2.0 * s[2] * change[2];
  return area * (first + m.energy_density(change));
}

static Vec2 principal_direction(double a, double b, double c, double e) {
  Vec2 first = {b, e - a}, second = {e - c, b};
  Vec2 pick = std::hypot(first[0], first[1]) >= std::hypot(second[0], second[1]) ? first : second;
  double length = std::hypot(pick[0], pick[1]);
  if (length == 0.0) return a >= c ? Vec2{1.0, 0.0} : Vec2{0.0, 1.0};
  return {pick[0] / length, pick[1] / length};
}

std::array<double, 3> clamp_negative_stress(std::array<double, 3> s) {
  double a = s[0], c = s[1], b = s[2];
  double mean = 0.5 * (a + c);
  double radius = std::sqrt(0.25 * (a - c) * (a - c) + b * b);
  double high = mean + radius, low = mean - radius;
  if (low >= 0.0) return s;
  if (high <= 0.0) return {0.0, 0.0, 0.0};
  Vec2 v = principal_direction(a, b, c, high);
  return {high * v[0] * v[0], high * v[1] * v[1], high * v[0] * v[1]};
}

// ---- hinge.rs:13-120 ------------------------------------------------------------------------
double hinge_angle_of(const std::array<Vec3, 4>& p) {
  Vec3 x1 = p[0], x2 = p[1], x3 = p[2], x4 = p[3];
  Vec3 edge = sub(x4, x3);
  Vec3 n1 = cross(sub(x1, x3), sub(x1, x4));
  Vec3 n2 = cross(sub(x2, x4), sub(x2, x3));
  Vec3 u1 = scale(n1, 1.0 / norm(n1)), u2 = scale(n2, 1.0 / norm(n2));
  double sine = dot(cross(u1, u2), scale(edge, 1.0 / norm(edge)));
  return std::atan2(sine, dot(u1, u2));
}

double HingeElement::angle(const std::vector<Vec3>& x) const {
  return hinge_angle_of({x[vertices[0]], x[vertices[1]], x[vertices[2]], x[vertices[3]]});
}

template <class T> using V3 = std::array<T, 3>;
template <class T> static V3<T> sub3(V3<T> a, V3<T> b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
template <class T> static V3<T> add3(V3<T> a, V3<T> b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
template <class T> static V3<T> scale3(V3<T> a, T s) { return {a[0] * s, a[1] * s, a[2] * s}; }
template <class T> static T dot3(V3<T> a, V3<T> b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class T> static V3<T> cross3(V3<T> a, V3<T> b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// hinge.rs:94-106: written once for f64 and Dual with the same operation order.
template <class T>
std::array<std::array<T, 3>, 4> angle_gradient_of(const std::array<std::array<T, 3>, 4>& p) {
  V3<T> x1 = p[0], x2 = p[1], x3 = p[2], x4 = p[3];
  V3<T> edge = sub3(x4, x3);
  T length = dsqrt(dot3(edge, edge));
  V3<T> n1 = cross3(sub3(x1, x3), sub3(x1, x4));
  V3<T> n2 = cross3(sub3(x2, x4), sub3(x2, x3));
  V3<T> a1 = scale3(n1, constant_of<T>(1.0) / dot3(n1, n1));
  V3<T> a2 = scale3(n2, constant_of<T>(1.0) / dot3(n2, n2));
  V3<T> u1 = scale3(a1, length), u2 = scale3(a2, length);
  V3<T> u3 = add3(scale3(a1, dot3(sub3(x1, x4), edge) / length), scale3(a2, dot3(sub3(x2, x4), edge) / length));
  V3<T> u4 = add3(scale3(a1, -dot3(sub3(x1, x3), edge) / length), scale3(a2, -dot3(sub3(x2, x3), edge) / length));
  T m1 = constant_of<T>(-1.0);
  return {scale3(u1, m1), scale3(u2, m1), scale3(u3, m1), scale3(u4, m1)};
}
template std::array<std::array<double, 3>, 4> angle_gradient_of<double>(const std::array<std::array<double, 3>, 4>&);
template std::array<std::array<Dual, 3>, 4> angle_gradient_of<Dual>(const std::array<std::array<Dual, 3>, 4>&);

std::array<Vec3, 4> HingeElement::angle_gradient(const std::vector<Vec3>& x) const {
  return angle_gradient_of<double>({x[vertices[0]], x[vertices[1]], x[vertices[2]], x[vertices[3]]});
}

double HingeElement::energy(const std::vector<Vec3>& x) const {
  double theta = angle(x);
  return stiffness * theta * theta;
}

std::array<Vec3, 4> HingeElement::gradient(const std::vector<Vec3>& x) const {
  double factor = 2.0 * stiffness * angle(x);
  auto g = angle_gradient(x);
  for (auto& v : g) v = scale(v, factor);
  return g;
}

HingeBlocks HingeElement::hessian(const std::vector<Vec3>& x) const {
  auto g = angle_gradient(x);
  double factor = 2.0 * stiffness;
  HingeBlocks b;
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l) {
      Mat3 o = outer(g[k], g[l]);
      for (auto& row : o)
        for (auto& v : row) v = v * factor;
      b[k][l] = o;
    }
  return b;
}

HingeBlocks HingeElement::angle_hessian(const std::vector<Vec3>& x) const {
  HingeBlocks b{};
  for (int l = 0; l < 4; ++l)
    for (int c = 0; c < 3; ++c) {
      std::array<std::array<Dual, 3>, 4> seeded;
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) seeded[k][d] = {x[vertices[k]][d], (k == l && d == c) ? 1.0 : 0.0};
      auto column = angle_gradient_of<Dual>(seeded);
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) b[k][l][r][c] = column[k][r].t;
    }
  return b;
}

HingeBlocks HingeElement::tangent(const std::vector<Vec3>& x) const {
  double theta = angle(x);
  auto g = angle_gradient(x);
  auto h = angle_hessian(x);
  double factor = 2.0 * stiffness;
  HingeBlocks b;
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) b[k][l][r][c] = factor * (g[k][r] * g[l][c] + theta * h[k][l][r][c]);
  return b;
}

double HingeElement::energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const {
  std::array<Vec3, 4> moved;
  for (int k = 0; k < 4; ++k) moved[k] = add(x[vertices[k]], scale(d[vertices[k]], alpha));
  double before = angle(x), after = hinge_angle_of(moved);
  return stiffness * (after - before) * (after + before);
}

double hinge_stiffness(const std::vector<Vec2>& rest, std::array<int, 4> v, double bending) {
  Vec2 p1 = rest[v[0]], p2 = rest[v[1]], p3 = rest[v[2]], p4 = rest[v[3]];
  Vec2 edge = {p4[0] - p3[0], p4[1] - p3[1]};
  double length2 = edge[0] * edge[0] + edge[1] * edge[1];
  auto area = [&](Vec2 p) {
    return 0.5 * std::fabs((p3[0] - p[0]) * (p4[1] - p[1]) - (p3[1] - p[1]) * (p4[0] - p[0]));
  };
  return bending * 3.0 * length2 / (area(p1) + area(p2));
}

// cases.rs:94-105
static std::vector<HingeElement> hinge_elements(const Mesh& mesh, const Material& material) {
  std::vector<HingeElement> out;
  for (const auto& v : mesh.hinges) {
    Vec2 a = mesh.rest[v[2]], b = mesh.rest[v[3]];
    double length = std::hypot(b[0] - a[0], b[1] - a[1]);
    Vec2 direction = {(b[0] - a[0]) / length, (b[1] - a[1]) / length};
    out.push_back({v, hinge_stiffness(mesh.rest, v, material.hinge_bending(direction))});
  }
  return out;
}

// ---- contact.rs:14-150 ----------------------------------------------------------------------
double Collider::distance(Vec3 x) const {
  if (kind == 1) return dot(a, x) - w;
  return norm(sub(x, a)) - w;
}
Vec3 Collider::normal(Vec3 x) const {
  if (kind == 1) return a;
  Vec3 q = sub(x, a);
  return {q[0] / norm(q), q[1] / norm(q), q[2] / norm(q)};
}
double Collider::curvature(Vec3 x) const { return kind == 1 ? 0.0 : 1.0 / norm(sub(x, a)); }

static double plane_step(double d0, double rate) {
  double s = rate >= 0.0 ? 1.0 : (1.0 - KEEP_FRACTION) * d0 / -rate;
  return std::min(s, 1.0);
}
static double sphere_step(double d0, Vec3 q, Vec3 p, double radius) {
  double a = dot(p, p), b = 2.0 * dot(q, p);
  double target = radius + KEEP_FRACTION * d0;
  double c = (1.0 - KEEP_FRACTION) * d0 * (norm(q) + target);
  double disc = b * b - 4.0 * a * c;
  if (b >= 0.0 || disc < 0.0) return 1.0;
  return std::min(2.0 * c / (-b + std::sqrt(disc)), 1.0);
}
double Collider::max_step(Vec3 x, Vec3 p) const {
  double d0 = distance(x);
  if (kind == 1) return plane_step(d0, dot(a, p));
  return sphere_step(d0, sub(x, a), p, w);
}

double Barrier::value(double d) const {
  if (d >= dhat) return 0.0;
  if (d <= 0.0) return INFINITY;
  return -kappa * ((d - dhat) * (d - dhat)) * std::log(d / dhat);
}
double Barrier::first(double d) const {
  if (d >= dhat) return 0.0;
  double t = d - dhat;
  return -kappa * (2.0 * t * std::log(d / dhat) + t * t / d);
}
double Barrier::second(double d) const {
  if (d >= dhat) return 0.0;
  double t = d - dhat;
  return -kappa * (2.0 * std::log(d / dhat) + 4.0 * t / d - t * t / (d * d));
}
double Barrier::resting_kappa(double dhat, double mass, double gravity) {
  Barrier unit{dhat, 1.0};
  return mass * gravity / std::fabs(unit.first(0.5 * dhat));
}

double Contact::energy(Vec3 x) const {
  double s = 0.0;
  for (const auto& c : colliders) s += barrier.value(c.distance(x));
  return s;
}
Vec3 Contact::gradient(Vec3 x) const {
  Vec3 sum = {0, 0, 0};
  for (const auto& c : colliders) {
    double slope = barrier.first(c.distance(x));
    Vec3 n = c.normal(x);
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + slope * n[k];
  }
  return sum;
}
Mat3 Contact::hessian_with(Vec3 x, bool curvature_term) const {
  Mat3 sum{};
  for (const auto& c : colliders) {
    double d = c.distance(x);
    double first = barrier.first(d), second = barrier.second(d);
    double curvature = curvature_term ? c.curvature(x) : 0.0;
    Mat3 nn = outer(c.normal(x), c.normal(x));
    Mat3 next;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double identity = i == j ? 1.0 : 0.0;
        next[i][j] = sum[i][j] + second * nn[i][j] + first * curvature * (identity - nn[i][j]);
      }
    sum = next;
  }
  return sum;
}
double Contact::max_step(Vec3 x, Vec3 p) const {
  double a = 1.0;
  for (const auto& c : colliders) a = std::fmin(a, c.max_step(x, p));
  return a;
}
double Contact::max_step_limited(Vec3 x, Vec3 p, int* at) const {
  double alpha = 1.0;
  *at = -1;
  for (size_t k = 0; k < colliders.size(); ++k) {
    double s = colliders[k].max_step(x, p);
    if (s < alpha) { alpha = s; *at = (int)k; }
  }
  return alpha;
}
double Contact::min_distance(Vec3 x) const {
  double m = INFINITY;
  for (const auto& c : colliders) m = std::fmin(m, c.distance(x));
  return m;
}

// ---- bsr.rs:12-93 ---------------------------------------------------------------------------
Bsr Bsr::with_pattern(int rows, const std::vector<std::vector<int>>& groups) {
  std::vector<std::vector<int>> columns(rows);
  for (int r = 0; r < rows; ++r) columns[r].push_back(r);
  for (const auto& g : groups)
    for (int r : g) columns[r].insert(columns[r].end(), g.begin(), g.end());
  Bsr m;
  m.row_ptr.push_back(0);
  for (auto& row : columns) {
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    m.col_idx.insert(m.col_idx.end(), row.begin(), row.end());
    m.row_ptr.push_back((int)m.col_idx.size());
  }
  m.vals.assign(9 * m.col_idx.size(), 0.0);
  return m;
}
int Bsr::position(int row, int column) const {
  auto b = col_idx.begin() + row_ptr[row], e = col_idx.begin() + row_ptr[row + 1];
  auto it = std::lower_bound(b, e, column);
  if (it == e || *it != column) { std::fprintf(stderr, "block outside sparsity pattern\n"); std::abort(); }
  return (int)(it - col_idx.begin());
}
void Bsr::add(int row, int column, const Mat3& b) {
  size_t base = 9 * (size_t)position(row, column);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) vals[base + 3 * r + c] += b[r][c];
}
std::vector<double> Bsr::multiply(const std::vector<double>& x) const {
  std::vector<double> y(x.size(), 0.0);
  for (int row = 0; row < rows(); ++row)
    for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k) {
      int column = col_idx[k];
      for (int r = 0; r < 3; ++r) {
        const double* v = &vals[9 * (size_t)k + 3 * r];
        y[3 * row + r] += v[0] * x[3 * column] + v[1] * x[3 * column + 1] + v[2] * x[3 * column + 2];
      }
    }
  return y;
}
std::vector<int> Bsr::diagonal_positions() const {
  std::vector<int> d(rows());
  for (int r = 0; r < rows(); ++r) d[r] = position(r, r);
  return d;
}
double Bsr::asymmetry() const {
  double largest = 0.0;
  for (double v : vals) largest = std::fmax(largest, std::fabs(v));
  double worst = 0.0;
  for (int row = 0; row < rows(); ++row)
    for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k) {
      int mirror = position(col_idx[k], row);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          worst = std::fmax(worst, std::fabs(vals[9 * (size_t)k + 3 * r + c] - vals[9 * (size_t)mirror + 3 * c + r]));
    }
  return worst / largest;
}
Bsr Bsr::rounded_to_f32() const {
  Bsr m = *this;
  for (auto& v : m.vals) v = (double)(float)v;
  return m;
}

// ---- band.rs:11-100 -------------------------------------------------------------------------
static int band_of(const Bsr& m) {
  int widest = 0;
  for (int row = 0; row < m.rows(); ++row)
    for (int k = m.row_ptr[row]; k < m.row_ptr[row + 1]; ++k) {
      int column = m.col_idx[k];
      if (column <= row) widest = std::max(widest, 3 * (row - column) + 2);
    }
  return widest;
}

bool BandCholesky::factor(const Bsr& m, std::string* error) {
  size = 3 * m.rows();
  bandwidth = band_of(m);
  lower.assign((size_t)size * (bandwidth + 1), 0.0);
  for (int row = 0; row < m.rows(); ++row)
    for (int k = m.row_ptr[row]; k < m.row_ptr[row + 1]; ++k) {
      int column = m.col_idx[k];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          int i = 3 * row + r, j = 3 * column + c;
          if (j <= i) lower[index(i, j)] = m.vals[9 * (size_t)k + 3 * r + c];
        }
    }
  for (int i = 0; i < size; ++i) {
    for (int j = first(i); j <= i; ++j) {
      int start = std::max(first(i), first(j));
      size_t ri = index(i, start), rj = index(j, start);
      int overlap = j - start;
      double product = 0.0;
      for (int k = 0; k < overlap; ++k) product += lower[ri + k] * lower[rj + k];
      double value = lower[index(i, j)] - product;
      if (i == j) {
        if (!(value > 0.0)) {
          if (error) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "matrix not positive definite at row %d (pivot %e)", i, value);
            *error = buf;
          }
          return false;
        }
        lower[index(i, j)] = std::sqrt(value);
      } else {
        lower[index(i, j)] = value / lower[index(j, j)];
      }
    }
  }
  return true;
}

std::vector<double> BandCholesky::solve(const std::vector<double>& rhs) const {
  std::vector<double> y = rhs;
  for (int i = 0; i < size; ++i) {
    double s = 0.0;
    for (int k = first(i); k < i; ++k) s += lower[index(i, k)] * y[k];
    y[i] = (y[i] - s) / lower[index(i, i)];
  }
  for (int i = size - 1; i >= 0; --i) {
    int end = std::min(i + bandwidth + 1, size);
    double s = 0.0;
    for (int k = i + 1; k < end; ++k) s += lower[index(k, i)] * y[k];
    y[i] = (y[i] - s) / lower[index(i, i)];
  }
  return y;
}

// ---- pcg.rs:60-281 --------------------------------------------------------------------------
template <class T> static T tree(std::vector<T>& lane) {
  for (int stride = WORKGROUP / 2; stride > 0; stride /= 2)
    for (int t = 0; t < stride; ++t) lane[t] = lane[t] + lane[t + stride];
  return lane[0];
}

template <class T> T reduce_rows_of(int rows, const std::vector<T>& values) {
  int groups = (rows + WORKGROUP - 1) / WORKGROUP;
  std::vector<T> partials(groups);
  std::vector<T> lane(WORKGROUP);
  for (int g = 0; g < groups; ++g) {
    for (int t = 0; t < WORKGROUP; ++t) {
      int row = g * WORKGROUP + t;
      lane[t] = row < rows ? values[row] : T(0);
    }
    partials[g] = tree(lane);
  }
  for (int t = 0; t < WORKGROUP; ++t) {
    T acc = T(0);
    for (int j = t; j < groups; j += WORKGROUP) acc = acc + partials[j];
    lane[t] = acc;
  }
  return tree(lane);
}
template float reduce_rows_of<float>(int, const std::vector<float>&);
template double reduce_rows_of<double>(int, const std::vector<double>&);

template <class T> std::array<T, 9> invert3(const T* m) {
  T a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  T c00 = e * i - f * h, c01 = c * h - b * i, c02 = b * f - c * e;
  T c10 = f * g - d * i, c11 = a * i - c * g, c12 = c * d - a * f;
  T c20 = d * h - e * g, c21 = b * g - a * h, c22 = a * e - b * d;
  T det = a * c00 + b * c10 + c * c20;
  return {c00 / det, c01 / det, c02 / det, c10 / det, c11 / det, c12 / det, c20 / det, c21 / det, c22 / det};
}
template std::array<float, 9> invert3<float>(const float*);
template std::array<double, 9> invert3<double>(const double*);

namespace {
template <class T> struct PcgSystem {
  const Bsr* m;
  std::vector<T> vals, dinv, rhs;
  int rows;
  PcgSystem(const Bsr& matrix, const std::vector<double>& b) : m(&matrix), rows(matrix.rows()) {
    vals.resize(matrix.vals.size());
    for (size_t i = 0; i < vals.size(); ++i) vals[i] = (T)matrix.vals[i];
    auto diag = matrix.diagonal_positions();
    dinv.resize(9 * (size_t)rows);
    for (int r = 0; r < rows; ++r) {
      auto inv = invert3<T>(&vals[9 * (size_t)diag[r]]);
      for (int k = 0; k < 9; ++k) dinv[9 * (size_t)r + k] = inv[k];
    }
    rhs.resize(b.size());
    for (size_t i = 0; i < b.size(); ++i) rhs[i] = (T)b[i];
  }
  void multiply_row(int row, const std::vector<T>& p, T acc[3]) const {
    acc[0] = acc[1] = acc[2] = T(0);
    for (int k = m->row_ptr[row]; k < m->row_ptr[row + 1]; ++k) {
      int c = 3 * m->col_idx[k];
      const T* v = &vals[9 * (size_t)k];
      for (int r = 0; r < 3; ++r) acc[r] = acc[r] + (v[3 * r] * p[c] + v[3 * r + 1] * p[c + 1] + v[3 * r + 2] * p[c + 2]);
    }
  }
  void precondition_row(int row, const std::vector<T>& r, T* out) const {
    const T* d = &dinv[9 * (size_t)row];
    T a = r[3 * row], b = r[3 * row + 1], c = r[3 * row + 2];
    out[0] = d[0] * a + d[1] * b + d[2] * c;
    out[1] = d[3] * a + d[4] * b + d[5] * c;
    out[2] = d[6] * a + d[7] * b + d[8] * c;
  }
};
template <class T> static T row_dot(const std::vector<T>& a, const std::vector<T>& b, int row) {
  int i = 3 * row;
  return a[i] * b[i] + a[i + 1] * b[i + 1] + a[i + 2] * b[i + 2];
}
template <class T> static T reduce_dot(int rows, const std::vector<T>& a, const std::vector<T>& b) {
  std::vector<T> v(rows);
  for (int r = 0; r < rows; ++r) v[r] = row_dot(a, b, r);
  return reduce_rows_of<T>(rows, v);
}
}  // namespace

template <class T>
PcgReport pcg_solve(const Bsr& matrix, const std::vector<double>& rhs, const PcgSettings& s, bool breakdown_check) {
  auto started = std::chrono::steady_clock::now();
  PcgSystem<T> sys(matrix, rhs);
  int n = (int)rhs.size(), rows = sys.rows;
  std::vector<T> x(n, T(0)), r = sys.rhs, z(n), p, q(n, T(0));
  for (int row = 0; row < rows; ++row) sys.precondition_row(row, r, &z[3 * row]);
  T rz = reduce_dot(rows, r, z), rr = reduce_dot(rows, r, r), rr0 = rr;
  p = z;
  PcgReport rep;
  T tol2 = (T)(s.tolerance * s.tolerance);
  T above2 = (T)(s.replace_above * s.replace_above);
  auto relative = [&]() { return std::sqrt((double)rr / (double)rr0); };
  bool converged = rr <= T(0);
  rep.status = converged ? 1 : 0;
  while (!converged && rep.iterations < s.max_iterations) {
    for (int row = 0; row < rows; ++row) sys.multiply_row(row, p, &q[3 * row]);
    T pq = reduce_dot(rows, p, q);
    T alpha = rz / pq;
    if (breakdown_check && !(pq > T(0) && std::isfinite((double)alpha))) { rep.status = 3; break; }
    for (int i = 0; i < n; ++i) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * q[i];
    }
    for (int row = 0; row < rows; ++row) sys.precondition_row(row, r, &z[3 * row]);
    T rz_new = reduce_dot(rows, r, z);
    rr = reduce_dot(rows, r, r);
    T beta = rz_new / rz;
    rz = rz_new;
    rep.iterations += 1;
    rep.alphas.push_back((double)alpha);
    rep.betas.push_back((double)beta);
    if (s.history_stride > 0 && rep.iterations % s.history_stride == 0) rep.history.push_back({rep.iterations, relative()});
    converged = rr <= tol2 * rr0;
    if (!converged)
      for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    bool due = s.replace_every > 0 && rep.iterations % s.replace_every == 0;
    if (!converged && due && rr > above2 * rr0) {
      for (int row = 0; row < rows; ++row) {
        T ax[3];
        sys.multiply_row(row, x, ax);
        for (int d = 0; d < 3; ++d) r[3 * row + d] = sys.rhs[3 * row + d] - ax[d];
      }
      for (int row = 0; row < rows; ++row) sys.precondition_row(row, r, &z[3 * row]);
      rz = reduce_dot(rows, r, z);
      rr = reduce_dot(rows, r, r);
    }
  }
  if (rep.status != 3) rep.status = converged ? 1 : 2;
  rep.converged = converged;
  rep.recursive_relative_residual = relative();
  // residual.wgsl (REPLACE = false): true residual s = b - Ax in the solver's precision.
  std::vector<T> ss(rows);
  for (int row = 0; row < rows; ++row) {
    T ax[3];
    sys.multiply_row(row, x, ax);
    T s0 = sys.rhs[3 * row] - ax[0], s1 = sys.rhs[3 * row + 1] - ax[1], s2 = sys.rhs[3 * row + 2] - ax[2];
    ss[row] = s0 * s0 + s1 * s1 + s2 * s2;
  }
  T true_rr = reduce_rows_of<T>(rows, ss);
  rep.true_relative_residual_f = std::sqrt((double)true_rr / (double)rr0);
  rep.solution.resize(n);
  for (int i = 0; i < n; ++i) rep.solution[i] = (double)x[i];
  rep.history.push_back({rep.iterations, relative()});
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
  return rep;
}
template PcgReport pcg_solve<float>(const Bsr&, const std::vector<double>&, const PcgSettings&, bool);
template PcgReport pcg_solve<double>(const Bsr&, const std::vector<double>&, const PcgSettings&, bool);

// ---- newton.rs:112-125, e4_rules.rs:15-45 ---------------------------------------------------
const char* step_status_name(StepStatus s) {
  switch (s) {
    case Converged: return "converged";
    case IterationLimit: return "iteration_limit";
    case LineSearchFailed: return "line_search_failed";
    case RegularizationCap: return "regularization_cap";
    case ConvergedAtFloor: return "converged_at_floor";
  }
  return "?";
}
double raise_mu(double mu) { return std::fmin(std::fmax(10.0 * mu, MU_START), MU_CAP); }
double relax_mu(double mu) {
  double next = mu / 10.0;
  return next < MU_START ? 0.0 : next;
}
double next_forcing(double previous, double ratio) {
  double eta = 0.9 * ratio * ratio;
  double safeguard = 0.9 * previous * previous;
  double guarded = safeguard > 0.1 ? std::fmax(eta, safeguard) : eta;
  return std::clamp(guarded, 1e-7, 1e-2);
}
double f32_spacing(double max_abs) {
  float f = (float)max_abs;
  uint32_t bits;
  std::memcpy(&bits, &f, 4);
  int exponent = (int)((bits >> 23) & 0xff);
  return std::ldexp(1.0, exponent - 127 - 23);
}
bool update_is_small(double update_inf, double step_inf, double resolution) {
  return update_inf <= std::fmax(1e-4 * step_inf, 2.0 * resolution);
}

// ---- drape.rs:53-386 ------------------------------------------------------------------------
SceneSpec reference_scene(int columns, bool sphere_on_floor, double material_factor,
                          double radius_factor, double rest_factor, bool placed) {
  const double SHEET = 0.5;
  SceneSpec s;
  s.columns = columns;
  s.sheet_size = SHEET;
  s.rest_factor = rest_factor;
  Material base = Material::knit_baseline();
  s.material = {material_factor * base.wale, material_factor * base.course, material_factor * base.coupling,
                material_factor * base.shear, material_factor * base.bend_wale,
                material_factor * base.bend_course, base.density};
  s.gravity = 9.81;
  s.time_step = 1.0 / 240.0;
  s.pinned = !sphere_on_floor;
  if (sphere_on_floor) {
    const double activation = 1e-3 * SHEET;
    double radius = 0.25 * SHEET * radius_factor;
    s.colliders.push_back({1, {0.0, 0.0, 1.0}, 0.0});
    s.colliders.push_back({2, {0.5 * SHEET, 0.5 * SHEET, radius}, radius});
    s.dhat = activation;
    s.kappa = 0.0;  // resting kappa of the heaviest vertex
    double clearance = placed ? 2.0 * activation : 0.02 * SHEET;
    s.lift = 2.0 * radius + clearance;
  }
  return s;
}

Drape::Drape(const SceneSpec& sp) : spec(sp) {
  if (sp.jitter > 0.0 || sp.flip_fraction > 0.0)
    mesh = Mesh::irregular(sp.columns, sp.sheet_size, sp.jitter, sp.flip_fraction, sp.mesh_seed);
  else
    mesh = Mesh::grid(sp.columns, sp.columns, sp.sheet_size, sp.sheet_size);
  if (sp.rest_factor != 1.0) {
    double centre = 0.5 * sp.sheet_size;
    for (auto& p : mesh.rest) p = {centre + sp.rest_factor * (p[0] - centre), centre + sp.rest_factor * (p[1] - centre)};
  }
  material = sp.material;
  if (sp.pinned) fixed = {(mesh.rows - 1) * mesh.columns, mesh.rows * mesh.columns - 1};
  int nv = (int)mesh.rest.size();
  row_of.assign(nv, -1);
  for (int v = 0; v < nv; ++v)
    if (std::find(fixed.begin(), fixed.end(), v) == fixed.end()) {
      row_of[v] = (int)free_.size();
      free_.push_back(v);
    }
  std::vector<std::vector<int>> groups;
  for (const auto& t : mesh.triangles) {
    std::vector<int> g;
    for (int v : t) if (row_of[v] >= 0) g.push_back(row_of[v]);
    groups.push_back(g);
  }
  for (const auto& h : mesh.hinges) {
    std::vector<int> g;
    for (int v : h) if (row_of[v] >= 0) g.push_back(row_of[v]);
    groups.push_back(g);
  }
  pattern = Bsr::with_pattern((int)free_.size(), groups);
  for (const auto& t : mesh.triangles) membranes.push_back(MembraneElement::make(t, mesh.rest));
  hinges = hinge_elements(mesh, material);
  auto areas = mesh.vertex_areas();
  for (double a : areas) mass.push_back(material.density * a);
  exact_tangent = sp.exact_tangent;
  time_step = sp.time_step;
  lift = sp.lift;
  if (!sp.colliders.empty()) {
    contact.colliders = sp.colliders;
    double heaviest = 0.0;
    for (double m : mass) heaviest = std::fmax(heaviest, m);
    double kappa = sp.kappa > 0.0 ? sp.kappa : Barrier::resting_kappa(sp.dhat, heaviest, sp.gravity);
    contact.barrier = {sp.dhat, kappa};
  }
}

std::vector<Vec3> Drape::rest_positions() const {
  std::vector<Vec3> x;
  for (const auto& p : mesh.rest) x.push_back({p[0], p[1], lift});
  return x;
}

std::vector<Vec3> Drape::prediction(const std::vector<Vec3>& x, const std::vector<Vec3>& v) const {
  std::vector<Vec3> out = x;
  const double h = spec.time_step;
  Vec3 gravity = {0.0, 0.0, -spec.gravity};
  for (int i : free_)
    for (int d = 0; d < 3; ++d) out[i][d] = x[i][d] + h * v[i][d] + h * h * gravity[d];
  return out;
}

double Drape::barrier_energy(const std::vector<Vec3>& x) const {
  double s = 0.0;
  for (int i : free_) s += contact.energy(x[i]);
  return s;
}

double Drape::potential(const std::vector<Vec3>& x) const {
  double membrane = 0.0;
  for (const auto& e : membranes) membrane += e.energy(x, material);
  double hinge = 0.0;
  for (const auto& h : hinges) hinge += h.energy(x);
  double elastic = membrane + hinge;
  return has_contact() ? elastic + barrier_energy(x) : elastic;
}

std::vector<Vec3> Drape::potential_gradient(const std::vector<Vec3>& x) const {
  std::vector<Vec3> g(x.size(), Vec3{0, 0, 0});
  for (const auto& e : membranes) {
    auto ge = e.gradient(x, material);
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) g[e.vertices[k]][d] += ge[k][d];
  }
  if (has_contact())
    for (int i : free_) {
      Vec3 c = contact.gradient(x[i]);
      for (int d = 0; d < 3; ++d) g[i][d] += c[d];
    }
  for (const auto& h : hinges) {
    auto gh = h.gradient(x);
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) g[h.vertices[k]][d] += gh[k][d];
  }
  return g;
}

double Drape::total_energy(const std::vector<Vec3>& x, const std::vector<Vec3>& v) const {
  double kinetic = 0.0, gravitational = 0.0;
  for (size_t i = 0; i < x.size(); ++i) kinetic += 0.5 * mass[i] * dot(v[i], v[i]);
  for (size_t i = 0; i < x.size(); ++i) gravitational += mass[i] * spec.gravity * x[i][2];
  return kinetic + gravitational + potential(x);
}

double DrapeStep::value(const std::vector<Vec3>& x) const {
  double inertia = 0.0;
  for (int i : drape->free_) {
    Vec3 off = sub(x[i], x_hat[i]);
    inertia += 0.5 * drape->mass_over_h2(i) * dot(off, off);
  }
  return inertia + drape->potential(x);
}

std::vector<double> DrapeStep::residual(const std::vector<Vec3>& x) const {
  auto g = drape->potential_gradient(x);
  std::vector<double> out;
  out.reserve(3 * drape->free_.size());
  for (int i : drape->free_)
    for (int d = 0; d < 3; ++d) out.push_back(drape->mass_over_h2(i) * (x[i][d] - x_hat[i][d]) + g[i][d]);
  return out;
}

template <size_t N, class B>
static void add_blocks(const Drape& d, Bsr& m, const std::array<int, N>& vertices, const B& blocks) {
  for (size_t k = 0; k < N; ++k)
    for (size_t l = 0; l < N; ++l) {
      int r = d.row_of[vertices[k]], c = d.row_of[vertices[l]];
      if (r >= 0 && c >= 0) m.add(r, c, blocks[k][l]);
    }
}

Bsr DrapeStep::tangent(const std::vector<Vec3>& x) const {
  const Drape& d = *drape;
  Bsr m = d.pattern;
  for (size_t row = 0; row < d.free_.size(); ++row) m.add((int)row, (int)row, identity_scaled(d.mass_over_h2(d.free_[row])));
  for (const auto& e : d.membranes)
    add_blocks<3>(d, m, e.vertices, d.exact_tangent ? e.tangent(x, d.material) : e.hessian(x, d.material));
  for (const auto& h : d.hinges) add_blocks<4>(d, m, h.vertices, d.exact_tangent ? h.tangent(x) : h.hessian(x));
  if (d.has_contact())
    for (size_t row = 0; row < d.free_.size(); ++row)
      m.add((int)row, (int)row, d.contact.hessian_with(x[d.free_[row]], d.barrier_curvature));
  return m;
}

Bsr DrapeStep::regularized_tangent(const std::vector<Vec3>& x, double mu) const {
  Bsr m = tangent(x);
  for (size_t row = 0; row < drape->free_.size(); ++row)
    m.add((int)row, (int)row, identity_scaled(mu * drape->mass_over_h2(drape->free_[row])));
  return m;
}

double DrapeStep::barrier_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const {
  double s = 0.0;
  for (int i : drape->free_) {
    Vec3 moved = {x[i][0] + alpha * d[i][0], x[i][1] + alpha * d[i][1], x[i][2] + alpha * d[i][2]};
    s += drape->contact.energy(moved) - drape->contact.energy(x[i]);
  }
  return s;
}

double DrapeStep::energy_change(const std::vector<Vec3>& x, const std::vector<Vec3>& d, double alpha) const {
  const Drape& dr = *drape;
  double inertia = 0.0;
  for (int i : dr.free_) {
    Vec3 step = {alpha * d[i][0], alpha * d[i][1], alpha * d[i][2]};
    Vec3 offset = {x[i][0] - x_hat[i][0], x[i][1] - x_hat[i][1], x[i][2] - x_hat[i][2]};
    inertia += dr.mass_over_h2(i) * (dot(offset, step) + 0.5 * dot(step, step));
  }
  double membrane = 0.0;
  for (const auto& e : dr.membranes) membrane += e.energy_change(x, d, alpha, dr.material);
  double hinge = 0.0;
  for (const auto& h : dr.hinges) hinge += h.energy_change(x, d, alpha);
  double elastic = inertia + membrane + hinge;
  return dr.has_contact() ? elastic + barrier_change(x, d, alpha) : elastic;
}

std::vector<Vec3> DrapeStep::scatter(const std::vector<double>& rows) const {
  std::vector<Vec3> v(drape->mesh.rest.size(), Vec3{0, 0, 0});
  for (size_t row = 0; row < drape->free_.size(); ++row)
    v[drape->free_[row]] = {rows[3 * row], rows[3 * row + 1], rows[3 * row + 2]};
  return v;
}

std::vector<double> DrapeStep::gather(const std::vector<Vec3>& v) const {
  std::vector<double> out;
  for (int i : drape->free_) out.insert(out.end(), v[i].begin(), v[i].end());
  return out;
}

double DrapeStep::max_step(const std::vector<Vec3>& x, const std::vector<Vec3>& d) const {
  double a = 1.0;
  for (int i : drape->free_) a = std::fmin(a, drape->contact.max_step(x[i], d[i]));
  return a;
}

std::vector<double> DrapeStep::stiffness_times_positions(const std::vector<Vec3>& x) const {
  Bsr t = tangent(x);
  for (auto& v : t.vals) v = std::fabs(v);
  auto g = gather(x);
  for (auto& v : g) v = std::fabs(v);
  return t.multiply(g);
}

std::vector<double> DrapeStep::absolute_contributions(const std::vector<Vec3>& x) const {
  const Drape& d = *drape;
  std::vector<Vec3> sums(x.size(), Vec3{0, 0, 0});
  for (const auto& e : d.membranes) {
    auto g = e.gradient(x, d.material);
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) sums[e.vertices[k]][c] += std::fabs(g[k][c]);
  }
  for (const auto& h : d.hinges) {
    auto g = h.gradient(x);
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) sums[h.vertices[k]][c] += std::fabs(g[k][c]);
  }
  for (int i : d.free_) {
    Vec3 g = d.contact.gradient(x[i]);
    for (int c = 0; c < 3; ++c) sums[i][c] += std::fabs(g[c]);
  }
  std::vector<double> out;
  for (int i : d.free_)
    for (int c = 0; c < 3; ++c) out.push_back(std::fabs(d.mass_over_h2(i) * (x[i][c] - x_hat[i][c])) + sums[i][c]);
  return out;
}

double DrapeStep::residual_floor(const std::vector<Vec3>& x) const {
  return EPS64 * (vector_norm(absolute_contributions(x)) + vector_norm(stiffness_times_positions(x)));
}

std::vector<Vec3> feasible_start(const DrapeStep& step, const std::vector<Vec3>& x) {
  std::vector<Vec3> towards(x.size());
  for (size_t i = 0; i < x.size(); ++i) towards[i] = sub(step.x_hat[i], x[i]);
  double alpha = step.max_step(x, towards);
  std::vector<Vec3> out(x.size());
  for (size_t i = 0; i < x.size(); ++i)
    for (int k = 0; k < 3; ++k) out[i][k] = x[i][k] + alpha * towards[i][k];
  return out;
}

// ---- newton.rs:128-264 ----------------------------------------------------------------------
static void advance(std::vector<Vec3>& x, const std::vector<Vec3>& d, double step_length) {
  for (size_t i = 0; i < x.size(); ++i)
    for (int k = 0; k < 3; ++k) x[i][k] += step_length * d[i][k];
}

bool line_search(const DrapeStep& s, const std::vector<Vec3>& x, const std::vector<double>& g,
                 const std::vector<Vec3>& d, const NewtonSettings& settings, LineSearch* out) {
  double slope = free_dot(g, s.gather(d));
  if (slope >= 0.0) return false;
  double start = s.max_step(x, d);
  double factor = 1.0;
  for (int halvings = 0; halvings <= settings.max_halvings; ++halvings) {
    double step_length = start * factor;
    if (s.energy_change(x, d, step_length) <= settings.armijo * step_length * slope) {
      *out = {step_length, halvings};
      return true;
    }
    factor *= 0.5;
  }
  return false;
}

static std::vector<double> negated(const std::vector<double>& g) {
  std::vector<double> n(g.size());
  for (size_t i = 0; i < g.size(); ++i) n[i] = -g[i];
  return n;
}

StepSolution solve_step(const DrapeStep& s, std::vector<Vec3> start, double scale, const NewtonSettings& settings) {
  StepSolution step;
  step.positions = std::move(start);
  for (;;) {
    auto g = s.residual(step.positions);
    double relative = vector_norm(g) / scale;
    step.residuals.push_back(relative);
    if (relative <= settings.tolerance) { step.status = Converged; return step; }
    if ((int)step.searches.size() == settings.max_iterations) return step;
    BandCholesky f;
    std::string err;
    if (!f.factor(s.tangent(step.positions), &err)) { step.status = RegularizationCap; return step; }
    auto direction = s.scatter(f.solve(negated(g)));
    LineSearch ls;
    if (!line_search(s, step.positions, g, direction, settings, &ls)) { step.status = LineSearchFailed; return step; }
    advance(step.positions, direction, ls.step_length);
    step.searches.push_back(ls);
  }
}

static bool regularized_direction(const DrapeStep& s, const std::vector<Vec3>& x, const std::vector<double>& g,
                                  double start, Regularization& record, std::vector<Vec3>& direction, double& used) {
  auto neg = negated(g);
  double mu = start;
  for (;;) {
    record.max_mu = std::fmax(record.max_mu, mu);
    BandCholesky f;
    if (f.factor(s.regularized_tangent(x, mu))) {
      direction = s.scatter(f.solve(neg));
      if (free_dot(g, s.gather(direction)) < 0.0) { used = mu; return true; }
    }
    if (mu >= MU_CAP) { record.cap_hit = true; return false; }
    mu = raise_mu(mu);
    record.retries += 1;
  }
}

StepSolution solve_step_regularized(const DrapeStep& s, std::vector<Vec3> start, double scale,
                                    const NewtonSettings& settings, double floor_factor) {
  StepSolution step;
  step.positions = std::move(start);
  double mu = 0.0;
  for (;;) {
    auto g = s.residual(step.positions);
    double relative = vector_norm(g) / scale;
    step.residuals.push_back(relative);
    if (relative <= settings.tolerance) { step.status = Converged; return step; }
    if (floor_factor > 0.0) {
      double floor = s.residual_floor(step.positions) / scale;
      if (relative <= floor_factor * floor) { step.status = ConvergedAtFloor; return step; }
    }
    if ((int)step.searches.size() == settings.max_iterations) return step;
    std::vector<Vec3> direction;
    double used = 0.0;
    if (!regularized_direction(s, step.positions, g, mu, step.regularization, direction, used)) {
      step.status = RegularizationCap;
      return step;
    }
    LineSearch ls;
    if (!line_search(s, step.positions, g, direction, settings, &ls)) { step.status = LineSearchFailed; return step; }
    advance(step.positions, direction, ls.step_length);
    mu = ls.halvings == 0 ? relax_mu(used) : used;
    step.searches.push_back(ls);
  }
}

TimeStep time_step(const Drape& d, const std::vector<Vec3>& x, const std::vector<Vec3>& v, const NewtonSettings& s,
                   bool regularized, double floor_factor) {
  DrapeStep step{&d, d.prediction(x, v)};
  double scale = vector_norm(step.residual(x));
  auto start = d.has_contact() ? feasible_start(step, x) : step.x_hat;
  TimeStep out;
  out.solution = regularized ? solve_step_regularized(step, start, scale, s, floor_factor) : solve_step(step, start, scale, s);
  out.scale = scale;
  const double h = d.spec.time_step;
  for (size_t i = 0; i < x.size(); ++i)
    out.velocity.push_back({(out.solution.positions[i][0] - x[i][0]) / h, (out.solution.positions[i][1] - x[i][1]) / h,
                            (out.solution.positions[i][2] - x[i][2]) / h});
  return out;
}

// ---- e7_ptc.rs:32-189 -----------------------------------------------------------------------
double next_step_size(const PtcRules& r, double h, int inner) {
  double factor = inner < r.inner_max ? r.growth : r.shrink;
  return std::clamp(h * factor, r.h0 * r.h_min_factor, r.h_max);
}
double gravity_load(const Drape& d) {
  std::vector<double> loads;
  for (int i : d.free_) loads.push_back(d.mass[i] * d.spec.gravity);
  return vector_norm(loads);
}
std::vector<double> static_gradient(const Drape& d, const std::vector<Vec3>& x) {
  auto g = d.potential_gradient(x);
  std::vector<double> out;
  for (int i : d.free_) {
    out.push_back(g[i][0]);
    out.push_back(g[i][1]);
    out.push_back(g[i][2] + d.mass[i] * d.spec.gravity);
  }
  return out;
}
double static_energy(const Drape& d, const std::vector<Vec3>& x) {
  double s = 0.0;
  for (size_t i = 0; i < x.size(); ++i) s += d.mass[i] * d.spec.gravity * x[i][2];
  return d.potential(x) + s;
}
std::vector<Vec3> ptc_prediction(const Drape& d, const std::vector<Vec3>& x, double h) {
  auto p = x;
  for (int i : d.free_) p[i][2] = x[i][2] - h * h * d.spec.gravity;
  return p;
}
double static_floor(Drape& d, const std::vector<Vec3>& x) {
  DrapeStep step{&d, x};
  auto contributions = step.absolute_contributions(x);
  for (size_t row = 0; row < d.free_.size(); ++row) {
    int i = d.free_[row];
    contributions[3 * row + 2] += d.mass[i] * d.spec.gravity - std::fabs(d.mass_over_h2(i) * (x[i][2] - step.x_hat[i][2]));
  }
  Bsr t = step.tangent(x);
  for (size_t row = 0; row < d.free_.size(); ++row) t.add((int)row, (int)row, identity_scaled(-d.mass_over_h2(d.free_[row])));
  for (auto& v : t.vals) v = std::fabs(v);
  auto g = step.gather(x);
  for (auto& v : g) v = std::fabs(v);
  auto stiffness = t.multiply(g);
  return EPS64 * (vector_norm(contributions) + vector_norm(stiffness));
}
bool ptc_failed(StepStatus status, int iterations, const std::vector<Vec3>& x) {
  bool stuck = iterations == 0 && (status == RegularizationCap || status == LineSearchFailed);
  if (stuck) return true;
  for (const auto& p : x)
    for (double v : p)
      if (!std::isfinite(v)) return true;
  return false;
}

static StepSolution inner_solve(Drape& d, const std::vector<Vec3>& x, double h, double residual, const PtcRules& r) {
  d.time_step = h;
  DrapeStep system{&d, ptc_prediction(d, x, h)};
  NewtonSettings settings{r.inner_tolerance, r.inner_max, 1e-4, 30};
  return solve_step_regularized(system, feasible_start(system, x), residual, settings, 0.0);
}

static bool attempt(Drape& d, const std::vector<Vec3>& x, double h, double residual, const PtcRules& r,
                    std::vector<Vec3>& next, PtcStep& step) {
  StepSolution solution = inner_solve(d, x, h, residual, r);
  bool capped = solution.regularization.cap_hit && !solution.searches.empty();
  if (capped) {
    d.barrier_curvature = false;
    StepSolution retry = inner_solve(d, x, h, residual, r);
    d.barrier_curvature = true;
    solution = std::move(retry);
  }
  int iterations = (int)solution.searches.size();
  if (ptc_failed(solution.status, iterations, solution.positions)) return false;
  step = {h, residual, iterations, 0, solution.regularization.max_mu, solution.regularization.cap_hit, capped};
  next = std::move(solution.positions);
  return true;
}

static bool take_step(Drape& d, const std::vector<Vec3>& x, double h, double residual, const PtcRules& r,
                      std::vector<Vec3>& next, PtcStep& step) {
  double size = h;
  for (int rejected = 0; rejected <= r.max_rejections; ++rejected) {
    if (attempt(d, x, size, residual, r, next, step)) {
      step.rejected_attempts = rejected;
      return true;
    }
    size = std::fmax(size * r.shrink, r.h0 * r.h_min_factor);
  }
  return false;
}

// Rust `{:.3e}`: three decimals, exponent without '+' or leading zeros.
static std::string rust_exp3(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.3e", v);
  std::string s = buf;
  size_t e = s.find('e');
  if (e == std::string::npos) return s;
  std::string mant = s.substr(0, e), ex = s.substr(e + 1);
  bool neg = !ex.empty() && ex[0] == '-';
  size_t i = (ex[0] == '-' || ex[0] == '+') ? 1 : 0;
  while (i + 1 < ex.size() && ex[i] == '0') ++i;
  return mant + "e" + (neg ? "-" : "") + ex.substr(i);
}

PtcRun ptc_solve(Drape& d, std::vector<Vec3> start, double h_start, const PtcRules& r, int max_steps_override,
                 bool keep_states, std::string* log) {
  double load = gravity_load(d);
  PtcRun run;
  run.states.push_back(std::move(start));
  run.final_h = h_start;
  int max_steps = max_steps_override > 0 ? max_steps_override : r.max_steps;
  for (;;) {
    std::vector<Vec3> x = run.states.back();
    run.final_residual = vector_norm(static_gradient(d, x));
    double rel = run.final_residual / load;
    if (rel <= r.equilibrium) { run.reason = "tolerance"; break; }
    if (rel <= r.floor_guard && run.final_residual <= r.floor_factor * static_floor(d, x)) { run.reason = "floor"; break; }
    if ((int)run.steps.size() == max_steps) break;
    double h = run.steps.empty() ? h_start : next_step_size(r, run.steps.back().h, run.steps.back().inner_iterations);
    std::vector<Vec3> next;
    PtcStep step;
    if (!take_step(d, x, h, run.final_residual, r, next, step)) { run.reason = "stalled"; break; }
    if (log) {
      char mu[64];
      std::snprintf(mu, sizeof mu, "%g", step.max_mu);
      *log += "ptc step " + std::to_string(run.steps.size() + 1) + " h=" + rust_exp3(step.h) + " r/L=" +
              rust_exp3(run.final_residual / load) + " inner=" + std::to_string(step.inner_iterations) +
              " rejected=" + std::to_string(step.rejected_attempts) + " mu=" + mu + "\n";
    }
    run.steps.push_back(step);
    if (keep_states) run.states.push_back(std::move(next));
    else run.states.back() = std::move(next);
  }
  run.final_h = run.steps.empty() ? h_start : next_step_size(r, run.steps.back().h, run.steps.back().inner_iterations);
  return run;
}

// ---- cases.rs:64-161 ------------------------------------------------------------------------
CaseProblem build_case(CaseKind kind, int columns) {
  const double SHEET = 0.5, TIME_STEP = 1.0 / 240.0, GRAVITY = 9.81, IN_PLANE_LOAD = 2.0;
  CaseProblem p;
  p.mesh = Mesh::grid(columns, columns, SHEET, SHEET);
  Material material = Material::knit_baseline();
  std::vector<int> row_of(p.mesh.rest.size(), -1), free_v;
  for (int v = 0; v < (int)p.mesh.rest.size(); ++v)
    if (!p.mesh.is_boundary(v)) { row_of[v] = (int)free_v.size(); free_v.push_back(v); }
  double sw = kind == Tension ? 1.03 : 1.0, sc = kind == Tension ? 1.08 : 1.0;
  std::vector<Vec3> x;
  for (const auto& q : p.mesh.rest) x.push_back({sw * q[0], sc * q[1], 0.0});
  std::vector<std::vector<int>> groups;
  for (const auto& t : p.mesh.triangles) {
    std::vector<int> g;
    for (int v : t) if (row_of[v] >= 0) g.push_back(row_of[v]);
    groups.push_back(g);
  }
  for (const auto& h : p.mesh.hinges) {
    std::vector<int> g;
    for (int v : h) if (row_of[v] >= 0) g.push_back(row_of[v]);
    groups.push_back(g);
  }
  p.matrix = Bsr::with_pattern((int)free_v.size(), groups);
  for (const auto& t : p.mesh.triangles) {
    auto e = MembraneElement::make(t, p.mesh.rest);
    auto h = e.hessian(x, material);
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l)
        if (row_of[t[k]] >= 0 && row_of[t[l]] >= 0) p.matrix.add(row_of[t[k]], row_of[t[l]], h[k][l]);
  }
  for (const auto& he : hinge_elements(p.mesh, material)) {
    auto h = he.hessian(x);
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l)
        if (row_of[he.vertices[k]] >= 0 && row_of[he.vertices[l]] >= 0)
          p.matrix.add(row_of[he.vertices[k]], row_of[he.vertices[l]], h[k][l]);
  }
  auto areas = p.mesh.vertex_areas();
  if (kind == Dynamic)
    for (size_t row = 0; row < free_v.size(); ++row)
      p.matrix.add((int)row, (int)row, identity_scaled(material.density * areas[free_v[row]] / (TIME_STEP * TIME_STEP)));
  for (int v : free_v) {
    double px = p.mesh.rest[v][0], py = p.mesh.rest[v][1];
    double shape = std::sin(PI * px / SHEET) * std::sin(PI * py / SHEET);
    double skew = std::sin(2.0 * PI * px / SHEET) * std::sin(PI * py / SHEET);
    double area = areas[v];
    p.rhs.push_back(area * IN_PLANE_LOAD * shape);
    p.rhs.push_back(0.5 * area * IN_PLANE_LOAD * skew);
    p.rhs.push_back(-material.density * area * GRAVITY);
  }
  return p;
}

// ---- e3_generate.rs:84-166, e5_checks.rs:58-86 ----------------------------------------------
std::vector<Vec3> bent_state(const Drape& d, Lcg& rng) {
  double w = d.mesh.width, h = d.mesh.height, s = d.mesh.width / (double)(d.mesh.columns - 1);
  std::vector<Vec3> x;
  for (const auto& p : d.mesh.rest) {
    double a = p[0] + 0.1 * s * rng.uniform();
    double b = p[1] + 0.1 * s * rng.uniform();
    double fold = 0.05 * std::sin(PI * p[0] / w) * std::cos(2.0 * PI * p[1] / h);
    double c = fold + 0.1 * s * rng.uniform();
    x.push_back({a, b, c});
  }
  return x;
}

static std::vector<Vec3> shifted(const std::vector<Vec3>& x, const std::vector<Vec3>& v, double h) {
  std::vector<Vec3> out(x.size());
  for (size_t i = 0; i < x.size(); ++i)
    for (int k = 0; k < 3; ++k) out[i][k] = x[i][k] + h * v[i][k];
  return out;
}

DerivativeCheck drape_derivative_check(int columns) {
  const double FD_STEP = 1e-7, LINE_STEP = 1e-3;
  Drape drape(reference_scene(columns, false, 1.0, 1.0, 1.0, false));
  double spacing = drape.mesh.width / (double)(drape.mesh.columns - 1);
  Lcg rng(53);
  auto x = bent_state(drape, rng);
  std::vector<Vec3> x_hat = x;
  for (auto& p : x_hat)
    for (auto& v : p) v = v + 0.02 * spacing * rng.uniform();
  DrapeStep step{&drape, x_hat};
  std::vector<std::vector<Vec3>> directions;
  for (int k = 0; k < 6; ++k) {
    std::vector<Vec3> v(drape.mesh.rest.size(), Vec3{0, 0, 0});
    for (int i : drape.free_) {
      double a = rng.uniform(), b = rng.uniform(), c = rng.uniform();
      v[i] = {a, b, c};
    }
    directions.push_back(v);
  }
  DerivativeCheck out{0, 0, 0, 0};
  auto residual = step.residual(x);
  Bsr tangent = step.tangent(x);
  for (const auto& v : directions) {
    double fd = (step.value(shifted(x, v, FD_STEP)) - step.value(shifted(x, v, -FD_STEP))) / (2.0 * FD_STEP);
    auto gv = step.gather(v);
    double exact = 0.0;
    for (size_t j = 0; j < residual.size(); ++j) exact += residual[j] * gv[j];
    out.gradient_error = std::fmax(out.gradient_error, std::fabs(fd - exact) / std::fabs(exact));
    auto plus = step.residual(shifted(x, v, FD_STEP)), minus = step.residual(shifted(x, v, -FD_STEP));
    std::vector<double> fdv(plus.size());
    for (size_t j = 0; j < plus.size(); ++j) fdv[j] = (plus[j] - minus[j]) / (2.0 * FD_STEP);
    auto tv = tangent.multiply(gv);
    out.hessian_error = std::fmax(out.hessian_error, difference_norm(fdv, tv) / vector_norm(tv));
    auto moved = shifted(x, v, LINE_STEP);
    double direct = 0.0, change = 0.0;
    for (const auto& h : drape.hinges) direct += h.energy(moved) - h.energy(x);
    for (const auto& h : drape.hinges) change += h.energy_change(x, v, LINE_STEP);
    out.hinge_energy_change_error = std::fmax(out.hinge_energy_change_error, std::fabs(change - direct) / std::fabs(direct));
    double total = step.value(moved) - step.value(x);
    out.energy_change_error =
        std::fmax(out.energy_change_error, std::fabs(step.energy_change(x, v, LINE_STEP) - total) / std::fabs(total));
  }
  return out;
}

std::vector<Vec3> contact_state(const Drape& d) {
  Lcg rng(131);
  double dhat = d.contact.barrier.dhat;
  std::vector<Vec3> x;
  for (const auto& p : d.mesh.rest) {
    double dist = dhat * std::pow(0.05, 0.5 * (1.0 + rng.uniform()));
    bool placed = false;
    for (const auto& c : d.contact.colliders) {
      if (c.kind != 2) continue;
      double dx = p[0] - c.a[0], dy = p[1] - c.a[1], radius = c.w;
      if (std::sqrt(dx * dx + dy * dy) >= 0.6 * radius) continue;
      Vec3 n = {dx / radius, dy / radius, std::sqrt(1.0 - (dx * dx + dy * dy) / (radius * radius))};
      x.push_back({c.a[0] + (radius + dist) * n[0], c.a[1] + (radius + dist) * n[1], c.a[2] + (radius + dist) * n[2]});
      placed = true;
      break;
    }
    if (!placed) x.push_back({p[0], p[1], dist});
  }
  return x;
}

ProblemArrays problem_arrays(const Drape& d) {
  ProblemArrays a;
  int T = (int)d.membranes.size(), H = (int)d.hinges.size(), R = d.rows();
  for (const auto& p : d.rest_positions())
    for (double v : p) a.rest_positions.push_back((float)v);
  for (const auto& e : d.membranes) {
    for (int v : e.vertices) a.elem_vertices.push_back((uint32_t)v);
    double geom[7] = {e.grad_n[0][0], e.grad_n[0][1], e.grad_n[1][0], e.grad_n[1][1], e.grad_n[2][0], e.grad_n[2][1], e.area};
    for (double g : geom) a.elem_geom.push_back((float)g);
  }
  for (const auto& h : d.hinges) {
    for (int v : h.vertices) a.hinge_vertices.push_back((uint32_t)v);
    a.hinge_stiffness.push_back((float)h.stiffness);
  }
  for (int v : d.free_) {
    a.row_vertex.push_back((uint32_t)v);
    a.row_mass.push_back((float)d.mass_over_h2(v));
    a.row_unit_mass.push_back((float)d.mass[v]);
  }
  for (int v : d.pattern.row_ptr) a.row_ptr.push_back((uint32_t)v);
  for (int v : d.pattern.col_idx) a.col_idx.push_back((uint32_t)v);
  for (int v : d.pattern.diagonal_positions()) a.diag_idx.push_back((uint32_t)v);
  std::vector<std::vector<uint32_t>> blocks(d.pattern.blocks());
  for (int row = 0; row < R; ++row) blocks[d.pattern.position(row, row)].push_back((uint32_t)(9 * T + 16 * H + row));
  for (int e = 0; e < T; ++e) {
    const auto& v = d.membranes[e].vertices;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) {
        int r = d.row_of[v[k]], c = d.row_of[v[l]];
        if (r >= 0 && c >= 0) blocks[d.pattern.position(r, c)].push_back((uint32_t)(9 * e + 3 * k + l));
      }
  }
  for (int i = 0; i < H; ++i) {
    const auto& v = d.hinges[i].vertices;
    for (int k = 0; k < 4; ++k)
      for (int l = 0; l < 4; ++l) {
        int r = d.row_of[v[k]], c = d.row_of[v[l]];
        if (r >= 0 && c >= 0) blocks[d.pattern.position(r, c)].push_back((uint32_t)(9 * T + 16 * i + 4 * k + l));
      }
  }
  std::vector<std::vector<uint32_t>> forces(R);
  for (int row = 0; row < R; ++row) forces[row].push_back((uint32_t)(3 * T + 4 * H + row));
  for (int e = 0; e < T; ++e)
    for (int k = 0; k < 3; ++k)
      if (d.row_of[d.membranes[e].vertices[k]] >= 0) forces[d.row_of[d.membranes[e].vertices[k]]].push_back((uint32_t)(3 * e + k));
  for (int i = 0; i < H; ++i)
    for (int k = 0; k < 4; ++k)
      if (d.row_of[d.hinges[i].vertices[k]] >= 0) forces[d.row_of[d.hinges[i].vertices[k]]].push_back((uint32_t)(3 * T + 4 * i + k));
  auto flatten = [](const std::vector<std::vector<uint32_t>>& lists, std::vector<uint32_t>& ptr, std::vector<uint32_t>& idx) {
    ptr.push_back(0);
    for (const auto& l : lists) {
      idx.insert(idx.end(), l.begin(), l.end());
      ptr.push_back((uint32_t)idx.size());
    }
  };
  flatten(blocks, a.block_contrib_ptr, a.block_contrib_idx);
  flatten(forces, a.force_contrib_ptr, a.force_contrib_idx);
  for (int v : d.fixed) a.fixed_vertex.push_back((uint32_t)v);
  Lcg rng(97);
  for (const auto& p : bent_state(d, rng))
    for (double v : p) a.kernel_state.push_back((float)v);
  if (d.has_contact()) {
    double header[4] = {(double)d.contact.colliders.size(), d.contact.barrier.dhat, d.contact.barrier.kappa, 0.0};
    for (double h : header) a.colliders.push_back((float)h);
    for (const auto& c : d.contact.colliders)
      for (double v : c.packed()) a.colliders.push_back((float)v);
    for (const auto& p : contact_state(d))
      for (double v : p) a.contact_state.push_back((float)v);
  }
  return a;
}

}  // namespace oracle
