// Problem builder: a scene (sheet grid, material, colliders) to the problem arrays the step consumes.
//
// The reference builds these arrays in its Rust generator (mesh.rs:14-107, drape.rs:73-155,
// cases.rs:94-105, e3_generate.rs:84-116, e7_run.rs:46, e7_ptc.rs:38-40) and ships them in problem.bin;
// this is the library's own builder so that callers can go from a scene to a solve without the
// generator. Geometry is formed in double precision and rounded once to f32, so the arrays agree with
// problem.bin entry for entry. Irregular meshes (jittered interior vertices, flipped cell diagonals) are
// an extension with no reference counterpart.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../internal.h"

namespace drape {

struct SceneArrays {
  std::vector<float> rest_positions, elem_geom, hinge_stiffness, row_unit_mass, colliders;
  std::vector<uint32_t> elem_vertices, hinge_vertices, row_vertex, row_ptr, col_idx;
  // device builds only: diagonal positions and the contribution lists (e3_generate.rs:111-161)
  bool device_built = false;
  std::vector<uint32_t> diag_idx, block_contrib_ptr, block_contrib_idx, force_contrib_ptr, force_contrib_idx;
  double gravity_load = 0.0;
  drape_problem problem{};
};

namespace {

struct Lcg {
  uint64_t s;
  explicit Lcg(uint64_t seed) : s(seed * 6364136223846793005ULL + 1442695040888963407ULL) {}
  double uniform() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) / (double)(1ULL << 53) * 2.0 - 1.0;
  }
};

using P2 = std::array<double, 2>;

double cross2(P2 a, P2 b) { return a[0] * b[1] - a[1] * b[0]; }
P2 sub2(P2 a, P2 b) { return {a[0] - b[0], a[1] - b[1]}; }

// The sheet: rest positions (f64), triangles and pinned vertices.
struct SceneMesh {
  int n = 0;
  std::vector<P2> rest;
  std::vector<std::array<int, 3>> tris;
  std::vector<int> fixed;
};

int make_mesh(const drape_scene* sc, SceneMesh& m) {
  const int n = sc->columns;
  if (n < 3 || n > 8193 || !(sc->sheet_size > 0)) {
    set_error("scene: columns must be in [3, 8193] and the sheet size positive");
    return DRAPE_E_INVALID;
  }
  m.n = n;
  const double size = sc->sheet_size;
  const double dx = size / (double)(n - 1);
  const bool irregular = sc->jitter > 0.0 || sc->flip_fraction > 0.0;
  Lcg rng(sc->mesh_seed);
  m.rest.reserve((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      P2 p = {(double)i * dx, (double)j * dx};
      if (irregular && !(i == 0 || j == 0 || i == n - 1 || j == n - 1)) {
        p[0] += sc->jitter * dx * rng.uniform();
        p[1] += sc->jitter * dx * rng.uniform();
      }
      m.rest.push_back(p);
    }
  if (sc->rest_factor != 1.0) {
    const double centre = 0.5 * size;
    for (auto& p : m.rest) p = {centre + sc->rest_factor * (p[0] - centre), centre + sc->rest_factor * (p[1] - centre)};
  }
  // triangles: alternating diagonals (mesh.rs:57-67), optionally flipped per cell
  m.tris.reserve(2 * (size_t)(n - 1) * (n - 1));
  for (int j = 0; j + 1 < n; ++j)
    for (int i = 0; i + 1 < n; ++i) {
      const bool flip = irregular && 0.5 * (rng.uniform() + 1.0) < sc->flip_fraction;
      const int v00 = j * n + i, v10 = v00 + 1, v01 = v00 + n, v11 = v01 + 1;
      if ((((i + j) % 2) == 0) != flip) {
        m.tris.push_back({v00, v10, v11});
        m.tris.push_back({v00, v11, v01});
      } else {
        m.tris.push_back({v00, v10, v01});
        m.tris.push_back({v10, v11, v01});
      }
    }
  // pinned corners (drape.rs:148)
  if (sc->pinned) m.fixed = {(n - 1) * n, n * n - 1};
  return DRAPE_OK;
}

// Hinges, pattern, geometry and rows on the host; `mass` gets density times the lumped area of every vertex.
void host_topology(const drape_scene* sc, const SceneMesh& m, SceneArrays& out, std::vector<double>& mass) {
  const auto& rest = m.rest;
  const auto& tris = m.tris;
  // interior hinges [left opposite, right opposite, edge start, edge end], sorted (mesh.rs:83-107)
  std::map<std::pair<int, int>, std::vector<int>> edges;
  for (const auto& t : tris)
    for (int k = 0; k < 3; ++k) {
      const int a = t[k], b = t[(k + 1) % 3];
      edges[{std::min(a, b), std::max(a, b)}].push_back(t[(k + 2) % 3]);
    }
  std::vector<std::array<int, 4>> hinges;
  for (const auto& [key, opp] : edges) {
    if (opp.size() != 2) continue;
    const int a = key.first, b = key.second;
    const bool left = cross2(sub2(rest[b], rest[a]), sub2(rest[opp[0]], rest[a])) > 0.0;
    hinges.push_back(left ? std::array<int, 4>{opp[0], opp[1], a, b} : std::array<int, 4>{opp[1], opp[0], a, b});
  }
  std::sort(hinges.begin(), hinges.end());
  const int V = (int)rest.size(), T = (int)tris.size(), H = (int)hinges.size();
  // lumped masses (mesh.rs:27-36)
  std::vector<double> area(V, 0.0);
  for (const auto& t : tris) {
    const double a = 0.5 * std::fabs(cross2(sub2(rest[t[1]], rest[t[0]]), sub2(rest[t[2]], rest[t[0]])));
    for (int v : t) area[v] += a / 3.0;
  }
  mass.resize(V);
  for (int v = 0; v < V; ++v) mass[v] = sc->density * area[v];
  std::vector<int> row_of(V, -1);
  std::vector<int> free_v;
  for (int v = 0; v < V; ++v)
    if (std::find(m.fixed.begin(), m.fixed.end(), v) == m.fixed.end()) {
      row_of[v] = (int)free_v.size();
      free_v.push_back(v);
    }
  const int R = (int)free_v.size();
  // BSR pattern from coupling groups (bsr.rs:12-29)
  std::vector<std::vector<int>> cols(R);
  for (int r = 0; r < R; ++r) cols[r].push_back(r);
  auto couple = [&](const int* vs, int k) {
    int rows[4], q = 0;
    for (int i = 0; i < k; ++i)
      if (row_of[vs[i]] >= 0) rows[q++] = row_of[vs[i]];
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) cols[rows[i]].push_back(rows[j]);
  };
  for (const auto& t : tris) couple(t.data(), 3);
  for (const auto& h : hinges) couple(h.data(), 4);
  out.row_ptr.assign(1, 0);
  for (auto& c : cols) {
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    out.col_idx.insert(out.col_idx.end(), c.begin(), c.end());
    out.row_ptr.push_back((uint32_t)out.col_idx.size());
  }
  // element geometry (e3_generate.rs:108-111)
  out.elem_geom.resize(7 * (size_t)T);
  for (int e = 0; e < T; ++e) {
    const auto& t = tris[e];
    const P2 p0 = rest[t[0]], p1 = rest[t[1]], p2 = rest[t[2]];
    const double m00 = p1[0] - p0[0], m01 = p2[0] - p0[0], m10 = p1[1] - p0[1], m11 = p2[1] - p0[1];
    const double det = m00 * m11 - m01 * m10;
    const double g1[2] = {m11 / det, -m01 / det}, g2[2] = {-m10 / det, m00 / det};
    const double g0[2] = {-g1[0] - g2[0], -g1[1] - g2[1]};
    const double geom[7] = {g0[0], g0[1], g1[0], g1[1], g2[0], g2[1], 0.5 * std::fabs(det)};
    for (int k = 0; k < 7; ++k) out.elem_geom[7 * e + k] = (float)geom[k];
  }
  // hinge stiffness k_b(e) 3 |e|^2 / (A1 + A2) (cases.rs:94-105)
  out.hinge_vertices.resize(4 * (size_t)H);
  out.hinge_stiffness.resize(H);
  for (int h = 0; h < H; ++h) {
    const auto& hv = hinges[h];
    const P2 p1 = rest[hv[0]], p2 = rest[hv[1]], p3 = rest[hv[2]], p4 = rest[hv[3]];
    const double length = std::hypot(p4[0] - p3[0], p4[1] - p3[1]);
    const double d[2] = {(p4[0] - p3[0]) / length, (p4[1] - p3[1]) / length};
    const double bending = sc->bend_wale * d[1] * d[1] + sc->bend_course * d[0] * d[0];
    const double e0 = p4[0] - p3[0], e1 = p4[1] - p3[1];
    const double length2 = e0 * e0 + e1 * e1;
    auto tri_area = [&](P2 p) {
      return 0.5 * std::fabs((p3[0] - p[0]) * (p4[1] - p[1]) - (p3[1] - p[1]) * (p4[0] - p[0]));
    };
    const double stiffness = bending * 3.0 * length2 / (tri_area(p1) + tri_area(p2));
    for (int k = 0; k < 4; ++k) out.hinge_vertices[4 * h + k] = (uint32_t)hv[k];
    out.hinge_stiffness[h] = (float)stiffness;
  }
  out.row_vertex.assign(free_v.begin(), free_v.end());
}

// The same on the device (builder.hip).
int device_topology(const drape_scene* sc, const SceneMesh& m, SceneArrays& out, std::vector<double>& mass) {
  const int V = (int)m.rest.size(), T = (int)m.tris.size();
  std::vector<double> rest_xy(2 * (size_t)V);
  for (int v = 0; v < V; ++v) {
    rest_xy[2 * v] = m.rest[v][0];
    rest_xy[2 * v + 1] = m.rest[v][1];
  }
  std::vector<uint32_t> tris(3 * (size_t)T);
  for (int e = 0; e < T; ++e)
    for (int k = 0; k < 3; ++k) tris[3 * e + k] = (uint32_t)m.tris[e][k];
  std::vector<uint8_t> free_mask(V, 1);
  for (int v : m.fixed) free_mask[v] = 0;
  MeshTopology topo;
  TRY_RC(build_topology_device(V, rest_xy.data(), T, tris.data(), free_mask.data(), sc->density, sc->bend_wale,
                               sc->bend_course, topo));
  out.hinge_vertices = std::move(topo.hinge_vertices);
  out.hinge_stiffness = std::move(topo.hinge_stiffness);
  out.elem_geom = std::move(topo.elem_geom);
  out.row_vertex = std::move(topo.row_vertex);
  out.row_ptr = std::move(topo.row_ptr);
  out.col_idx = std::move(topo.col_idx);
  out.diag_idx = std::move(topo.diag_idx);
  out.block_contrib_ptr = std::move(topo.block_contrib_ptr);
  out.block_contrib_idx = std::move(topo.block_contrib_idx);
  out.force_contrib_ptr = std::move(topo.force_contrib_ptr);
  out.force_contrib_idx = std::move(topo.force_contrib_idx);
  out.device_built = true;
  mass = std::move(topo.mass);
  return DRAPE_OK;
}

// Rest positions, triangles, row masses, gravity load, colliders and the problem view (e3_generate.rs:84-104).
void finish_scene(const drape_scene* sc, const SceneMesh& m, const std::vector<double>& mass, SceneArrays& out) {
  const int V = (int)m.rest.size(), T = (int)m.tris.size(), H = (int)out.hinge_stiffness.size();
  const int R = (int)out.row_vertex.size();
  out.rest_positions.resize(3 * (size_t)V);
  for (int v = 0; v < V; ++v) {
    out.rest_positions[3 * v] = (float)m.rest[v][0];
    out.rest_positions[3 * v + 1] = (float)m.rest[v][1];
    out.rest_positions[3 * v + 2] = (float)sc->lift;
  }
  out.elem_vertices.resize(3 * (size_t)T);
  for (int e = 0; e < T; ++e)
    for (int k = 0; k < 3; ++k) out.elem_vertices[3 * e + k] = (uint32_t)m.tris[e][k];
  out.row_unit_mass.resize(R);
  double load2 = 0.0, heaviest = 0.0;
  for (int v = 0; v < V; ++v) heaviest = std::fmax(heaviest, mass[v]);
  for (int r = 0; r < R; ++r) {
    const double mr = mass[out.row_vertex[r]];
    out.row_unit_mass[r] = (float)mr;
    load2 += (mr * sc->gravity) * (mr * sc->gravity);
  }
  out.gravity_load = std::sqrt(load2);
  // colliders: [count, dhat, kappa, 0, (kind a b c w 0 0 0) x K]; kappa <= 0 selects the resting
  // stiffness of the heaviest vertex at dhat/2 (contact.rs:104-107).
  const int K = std::max(0, std::min(sc->n_colliders, 4));
  if (K > 0) {
    double kappa = sc->kappa;
    if (!(kappa > 0.0)) {
      const double dh = sc->dhat, d = 0.5 * dh, t = d - dh;
      const double unit_first = -1.0 * (2.0 * t * std::log(d / dh) + t * t / d);
      kappa = heaviest * sc->gravity / std::fabs(unit_first);
    }
    out.colliders = {(float)K, (float)sc->dhat, (float)kappa, 0.0f};
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < 8; ++j) out.colliders.push_back(j < 5 ? (float)sc->colliders[8 * k + j] : 0.0f);
  }
  drape_problem& p = out.problem;
  p.vertices = V; p.triangles = T; p.hinges = H; p.rows = R; p.blocks = (int32_t)out.col_idx.size();
  p.rest_positions = out.rest_positions.data();
  p.elem_vertices = out.elem_vertices.data();
  p.elem_geom = out.elem_geom.data();
  p.hinge_vertices = out.hinge_vertices.data();
  p.hinge_stiffness = out.hinge_stiffness.data();
  p.row_vertex = out.row_vertex.data();
  p.row_unit_mass = out.row_unit_mass.data();
  p.row_ptr = out.row_ptr.data();
  p.col_idx = out.col_idx.data();
  p.colliders = out.colliders.empty() ? nullptr : out.colliders.data();
  p.collider_floats = (int32_t)out.colliders.size();
  p.wale = sc->wale; p.course = sc->course; p.coupling = sc->coupling; p.shear = sc->shear;
  p.gravity = sc->gravity;
  p.time_step = sc->time_step;
  p.grid_columns = sc->columns;  // irregular sheets keep the grid's vertex numbering
}

int build_scene(const drape_scene* sc, SceneArrays& out, bool on_device) {
  SceneMesh m;
  TRY_RC(make_mesh(sc, m));
  std::vector<double> mass;
  if (on_device) TRY_RC(device_topology(sc, m, out, mass));
  else host_topology(sc, m, out, mass);
  finish_scene(sc, m, mass, out);
  return DRAPE_OK;
}

}  // namespace

}  // namespace drape

using namespace drape;

extern "C" {

static int scene_build(const drape_scene* scene, drape_scene_arrays** out, bool on_device) {
  if (!scene || !out) {
    set_error("null argument");
    return DRAPE_E_INVALID;
  }
  auto* a = new SceneArrays();
  int rc = build_scene(scene, *a, on_device);
  if (rc != DRAPE_OK) {
    delete a;
    *out = nullptr;
    return rc;
  }
  *out = reinterpret_cast<drape_scene_arrays*>(a);
  return DRAPE_OK;
}

int drape_scene_build(const drape_scene* scene, drape_scene_arrays** out) { return scene_build(scene, out, false); }

int drape_scene_build_device(const drape_scene* scene, drape_scene_arrays** out) {
  return scene_build(scene, out, true);
}

const drape_problem* drape_scene_problem(const drape_scene_arrays* a, double* gravity_load) {
  const auto* s = reinterpret_cast<const SceneArrays*>(a);
  if (gravity_load) *gravity_load = s->gravity_load;
  return &s->problem;
}

int drape_scene_contributions(const drape_scene_arrays* a, int64_t* sizes, uint32_t* diag_idx,
                              uint32_t* block_contrib_ptr, uint32_t* block_contrib_idx, uint32_t* force_contrib_ptr,
                              uint32_t* force_contrib_idx) {
  const auto* s = reinterpret_cast<const SceneArrays*>(a);
  if (!s || !sizes) {
    set_error("null argument");
    return DRAPE_E_INVALID;
  }
  if (!s->device_built) {
    set_error("the contribution lists come with drape_scene_build_device only");
    return DRAPE_E_INVALID;
  }
  sizes[0] = (int64_t)s->block_contrib_idx.size();
  sizes[1] = (int64_t)s->force_contrib_idx.size();
  auto put = [](const std::vector<uint32_t>& v, uint32_t* dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  put(s->diag_idx, diag_idx);
  put(s->block_contrib_ptr, block_contrib_ptr);
  put(s->block_contrib_idx, block_contrib_idx);
  put(s->force_contrib_ptr, force_contrib_ptr);
  put(s->force_contrib_idx, force_contrib_idx);
  return DRAPE_OK;
}

void drape_scene_free(drape_scene_arrays* a) { delete reinterpret_cast<SceneArrays*>(a); }

}  // extern "C"
