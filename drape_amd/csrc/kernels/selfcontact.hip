// Self-contact of the sheet (no reference counterpart: the reference's contact is vertex-collider only): the
// IPC log barrier b(d) (the colliders' dhat and kappa) over every point-triangle and edge-edge pair of the
// mesh closer than dhat, in FP64.
//
// Broad phase: a uniform grid of cell size h = (longest edge) + (inflation), hashed (Teschner et al. 2003:
// 73856093 x ^ 19349663 y ^ 83492791 z) and sorted by key (hipCUB radix sort, stable). A vertex within the
// inflation distance of a triangle lies within h of its centroid, and two edges within it have midpoints
// within h, so each triangle queries the vertices, and each edge the edges, of the 27 cells around its
// centroid (midpoint) cell, one thread per (primitive, cell); the true cell of every visited primitive is
// checked, so hash collisions neither add nor duplicate pairs. Counting pass, scan, then the emitting pass in
// the same order: the candidate list is deterministic (point-triangle pairs by triangle and cell, then
// edge-edge pairs by first edge and cell).
//
// Narrow phase: closest points by Voronoi-region classification (point-triangle) and clamped parameters
// (segment-segment), following Ericson, Real-Time Collision Detection, 5.1.5 and 5.1.9. The distance
// gradient comes from the envelope theorem: with closest points x1 = sum a_k p_k on one primitive and
// x2 on the other and n = (x1 - x2)/d, grad_k d = c_k n with c_k the barycentric weight of vertex k (negated
// on the second primitive). A pair contributes b(d) to the energy, b'(d) c_k n to the gradient of its vertex
// k and the Gauss-Newton block b''(d) c_k c_l n n^T to the Hessian (positive semidefinite).
//
// Step bound: conservative advancement per pair (Mirtich 1996; as in additive CCD, Li et al. 2021): the
// distance of a pair shrinks at most lp t along t -> x + t d, lp = the largest displacement of a vertex of one
// primitive relative to a vertex of the other, so steps of 0.9 (d - g)/lp keep it above g = 0.1 d(x), the
// reference's collider rule. The line-search set holds the pairs closer than dhat + lp; the grid covers them
// with an inflation of dhat + 2 max |displacement - mean displacement|.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../internal.h"
#include "barrier.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace drape {

struct SelfContact {
  int V = 0, T = 0, E = 0;
  int2* d_edges = nullptr;       // [E] unique mesh edges (a < b)
  int32_t* d_row_of = nullptr;   // [V] row of each vertex, -1 when pinned
  // broad phase
  uint32_t *d_vkey = nullptr, *d_vid = nullptr, *d_vkey_s = nullptr, *d_vid_s = nullptr;  // [V]
  uint32_t *d_ekey = nullptr, *d_eid = nullptr, *d_ekey_s = nullptr, *d_eid_s = nullptr;  // [E]
  uint32_t *d_tcount = nullptr, *d_toff = nullptr;  // [27 T + 1] per (triangle, neighbour cell)
  uint32_t *d_ecount = nullptr, *d_eoff = nullptr;  // [27 E + 1] per (edge, neighbour cell)
  // the step's swept boxes: (cell key, id) of every cell a vertex's / an edge's box over the step overlaps
  int vcap = 0, ecap = 0;        // entries allocated
  uint32_t *d_bvkey = nullptr, *d_bvid = nullptr, *d_bvkey_s = nullptr, *d_bvid_s = nullptr;
  uint32_t *d_bekey = nullptr, *d_beid = nullptr, *d_bekey_s = nullptr, *d_beid_s = nullptr;
  uint32_t* d_over = nullptr;    // [1] set when a box spans more cells than its limit
  // per query slot (triangle or edge, queried cell) of the step: candidate counts and offsets
  int ptcap = 0, eecap = 0;
  uint32_t *d_ptcount = nullptr, *d_ptoff = nullptr, *d_eecount = nullptr, *d_eeoff = nullptr;
  uint32_t *d_tq = nullptr, *d_tqoff = nullptr;  // [T + 1] queried cells per triangle, their offsets
  float* d_part = nullptr;       // [5 * groups] block edge maxima and displacement sums, then deviations
  float* d_grid = nullptr;       // [5] cell size, inflation, mean displacement
  void* d_tmp = nullptr;         // hipCUB scratch
  size_t tmp_bytes = 0;
  uint32_t* h_totals = nullptr;  // pinned [8]
  // candidates of the evaluated state (their pair terms below) and of the current step (the line-search set)
  struct Set {
    int cap = 0, npt = 0, nee = 0;
    int4* d_cand = nullptr;      // [cap] PT: (p, a, b, c); EE: (p1, q1, p2, q2)
  } eval, step;
  bool step_ready = false;       // `step` belongs to the current positions and direction
  int tcap = 0;                  // pairs the term arrays hold
  // pair terms of the last evaluation (active pairs only contribute)
  int4* d_crow = nullptr;        // [cap] rows of the pair's vertices (-1: pinned)
  float4* d_cdat = nullptr;      // [cap] n, b''
  float4* d_ccoef = nullptr;     // [cap] c_0..c_3
  float* d_cfirst = nullptr;     // [cap] b'
  double* d_dist = nullptr;      // [cap] distance
  uint32_t *d_ikey = nullptr, *d_ival = nullptr, *d_ikey_s = nullptr, *d_ival_s = nullptr;  // [4 cap]
  uint32_t* d_iptr = nullptr;    // [R + 1]
  bool terms = false;            // pair terms and incidence lists belong to the current assembly
  double dhat = 0.0, kappa = 0.0;
  // line-search sums
  double* d_epart = nullptr;     // [2 * groups]
  double* d_eout = nullptr;      // [2]
  float* d_apart = nullptr;      // [groups] step bound per group
  int32_t* d_aidx = nullptr;     // [groups] pair of the group's bound
  int parts = 0;                 // groups the partial buffers hold
};

void self_contact_destroy(SelfContact* sc) {
  for (void* p : {(void*)sc->d_edges, (void*)sc->d_row_of, (void*)sc->d_vkey, (void*)sc->d_vid, (void*)sc->d_vkey_s,
                  (void*)sc->d_vid_s, (void*)sc->d_ekey, (void*)sc->d_eid, (void*)sc->d_ekey_s, (void*)sc->d_eid_s,
                  (void*)sc->d_tcount, (void*)sc->d_toff, (void*)sc->d_ecount, (void*)sc->d_eoff, (void*)sc->d_part,
                  (void*)sc->d_grid, sc->d_tmp, (void*)sc->eval.d_cand, (void*)sc->step.d_cand, (void*)sc->d_crow, (void*)sc->d_cdat,
                  (void*)sc->d_ccoef, (void*)sc->d_cfirst, (void*)sc->d_dist, (void*)sc->d_ikey, (void*)sc->d_ival,
                  (void*)sc->d_ikey_s, (void*)sc->d_ival_s, (void*)sc->d_iptr, (void*)sc->d_epart,
                  (void*)sc->d_eout, (void*)sc->d_apart, (void*)sc->d_aidx, (void*)sc->d_bvkey, (void*)sc->d_bvid,
                  (void*)sc->d_bvkey_s, (void*)sc->d_bvid_s, (void*)sc->d_bekey, (void*)sc->d_beid,
                  (void*)sc->d_bekey_s, (void*)sc->d_beid_s, (void*)sc->d_over, (void*)sc->d_ptcount,
                  (void*)sc->d_ptoff, (void*)sc->d_eecount, (void*)sc->d_eeoff, (void*)sc->d_tq, (void*)sc->d_tqoff})
    if (p) (void)hipFree(p);
  if (sc->h_totals) (void)hipHostFree(sc->h_totals);
  delete sc;
}

namespace {

struct P3 {
  double x, y, z;
};
__device__ inline P3 p3(const float4& v) { return {v.x, v.y, v.z}; }
__device__ inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline P3 axpy(P3 a, double s, P3 b) { return {a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }
__device__ inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Closest point of triangle (a, b, c) to p as barycentric weights w (Voronoi regions: vertices, edges,
// interior).
__device__ inline void closest_on_triangle(P3 p, P3 a, P3 b, P3 c, double w[3]) {
  const P3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) { w[0] = 1.0; w[1] = 0.0; w[2] = 0.0; return; }
  const P3 bp = sub(p, b);
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) { w[0] = 0.0; w[1] = 1.0; w[2] = 0.0; return; }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    w[0] = 1.0 - v; w[1] = v; w[2] = 0.0;
    return;
  }
  const P3 cp = sub(p, c);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) { w[0] = 0.0; w[1] = 0.0; w[2] = 1.0; return; }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double t = d2 / (d2 - d6);
    w[0] = 1.0 - t; w[1] = 0.0; w[2] = t;
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[0] = 0.0; w[1] = 1.0 - t; w[2] = t;
    return;
  }
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, t = vc * den;
  w[0] = 1.0 - v - t; w[1] = v; w[2] = t;
}

// Closest points of segments p1 + s (q1 - p1) and p2 + t (q2 - p2), s, t in [0, 1].
__device__ inline void closest_on_segments(P3 p1, P3 q1, P3 p2, P3 q2, double* s_out, double* t_out) {
  const P3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const double tiny = 1e-30;
  double s = 0.0, t = 0.0;
  if (a <= tiny && e <= tiny) {
    s = t = 0.0;
  } else if (a <= tiny) {
    s = 0.0;
    t = fmin(fmax(f / e, 0.0), 1.0);
  } else {
    const double c = dot(d1, r);
    if (e <= tiny) {
      t = 0.0;
      s = fmin(fmax(-c / a, 0.0), 1.0);
    } else {
      const double b = dot(d1, d2), den = a * e - b * b;
      s = den > 0.0 ? fmin(fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
      t = (b * s + f) / e;
      if (t < 0.0) {
        t = 0.0;
        s = fmin(fmax(-c / a, 0.0), 1.0);
      } else if (t > 1.0) {
        t = 1.0;
        s = fmin(fmax((b - c) / a, 0.0), 1.0);
      }
    }
  }
  *s_out = s;
  *t_out = t;
}

// Distance of a pair, its unit direction n and vertex coefficients c (grad_k d = c_k n).
struct PairGeom {
  double d;
  P3 n;
  double c[4];
};
__device__ inline PairGeom pair_geometry(bool pt, P3 x0, P3 x1, P3 x2, P3 x3) {
  PairGeom g;
  P3 diff;
  if (pt) {
    double w[3];
    closest_on_triangle(x0, x1, x2, x3, w);
    const P3 q = {w[0] * x1.x + w[1] * x2.x + w[2] * x3.x, w[0] * x1.y + w[1] * x2.y + w[2] * x3.y,
                  w[0] * x1.z + w[1] * x2.z + w[2] * x3.z};
    diff = sub(x0, q);
    g.c[0] = 1.0; g.c[1] = -w[0]; g.c[2] = -w[1]; g.c[3] = -w[2];
  } else {
    double s, t;
    closest_on_segments(x0, x1, x2, x3, &s, &t);
    const P3 a = axpy(x0, s, sub(x1, x0)), b = axpy(x2, t, sub(x3, x2));
    diff = sub(a, b);
    g.c[0] = 1.0 - s; g.c[1] = s; g.c[2] = -(1.0 - t); g.c[3] = -t;
  }
  g.d = sqrt(dot(diff, diff));
  const double inv = g.d > 0.0 ? 1.0 / g.d : 0.0;
  g.n = {diff.x * inv, diff.y * inv, diff.z * inv};
  return g;
}

__device__ inline int3 cell_of(P3 x, double h) {
  return make_int3((int)floor(x.x / h), (int)floor(x.y / h), (int)floor(x.z / h));
}
__device__ inline uint32_t cell_key(int3 c) {
  return ((uint32_t)c.x * 73856093u) ^ ((uint32_t)c.y * 19349663u) ^ ((uint32_t)c.z * 83492791u);
}
__device__ inline bool same_cell(int3 a, int3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ---- broad phase -------------------------------------------------------------------------------

// Displacement of vertex v along the step (free vertices only; pinned ones stay).
__device__ inline P3 disp(const float4* __restrict__ dir, const int32_t* __restrict__ row_of, int v) {
  return row_of[v] >= 0 ? p3(dir[v]) : P3{0.0, 0.0, 0.0};
}
__device__ inline double dist3(P3 a, P3 b) { return sqrt(dot(sub(a, b), sub(a, b))); }

// The most a pair's distance can shrink along x + t dir, t in [0, 1]: the largest relative displacement of a
// vertex of one primitive against one of the other (the closest points are convex combinations).
__device__ inline double approach(const float4* dir, const int32_t* row_of, bool pt, int4 v) {
  const P3 a0 = disp(dir, row_of, v.x), a1 = disp(dir, row_of, v.y), b0 = disp(dir, row_of, v.z),
           b1 = disp(dir, row_of, v.w);
  if (pt) return fmax(dist3(a0, a1), fmax(dist3(a0, b0), dist3(a0, b1)));
  return fmax(fmax(dist3(a0, b0), dist3(a0, b1)), fmax(dist3(a1, b0), dist3(a1, b1)));
}

// block maxima of the edge lengths and block sums of the free vertices' displacements (dir may be null)
__global__ void __launch_bounds__(256) k_extent(const int2* __restrict__ edges, int E, const float4* __restrict__ pos,
                                                const float4* __restrict__ dir, const int32_t* __restrict__ row_of,
                                                int V, float* __restrict__ part) {
  __shared__ float la[256];
  __shared__ float4 ls[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float len = 0.0f;
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < E) {
    const int2 e = edges[i];
    const float4 a = pos[e.x], b = pos[e.y];
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    len = sqrtf(dx * dx + dy * dy + dz * dz);
  }
  if (dir && i < V && row_of[i] >= 0) {
    const float4 d = dir[i];
    sum = make_float4(d.x, d.y, d.z, 1.0f);
  }
  la[threadIdx.x] = len;
  ls[threadIdx.x] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      la[threadIdx.x] = fmaxf(la[threadIdx.x], la[threadIdx.x + s]);
      const float4 u = ls[threadIdx.x], w = ls[threadIdx.x + s];
      ls[threadIdx.x] = make_float4(u.x + w.x, u.y + w.y, u.z + w.z, u.w + w.w);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[5 * blockIdx.x] = la[0];
    part[5 * blockIdx.x + 1] = ls[0].x;
    part[5 * blockIdx.x + 2] = ls[0].y;
    part[5 * blockIdx.x + 3] = ls[0].z;
    part[5 * blockIdx.x + 4] = ls[0].w;
  }
}

// grid[0] = longest edge, grid[2..4] = mean displacement of the free vertices (fixed-order sums)
__global__ void k_mean(const float* __restrict__ part, int groups, float* __restrict__ grid) {
  __shared__ float la[256];
  __shared__ float4 ls[256];
  float a = 0.0f;
  float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int j = threadIdx.x; j < groups; j += 256) {
    a = fmaxf(a, part[5 * j]);
    u = make_float4(u.x + part[5 * j + 1], u.y + part[5 * j + 2], u.z + part[5 * j + 3], u.w + part[5 * j + 4]);
  }
  la[threadIdx.x] = a;
  ls[threadIdx.x] = u;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      la[threadIdx.x] = fmaxf(la[threadIdx.x], la[threadIdx.x + s]);
      const float4 p = ls[threadIdx.x], q = ls[threadIdx.x + s];
      ls[threadIdx.x] = make_float4(p.x + q.x, p.y + q.y, p.z + q.z, p.w + q.w);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float n = fmaxf(ls[0].w, 1.0f);
    grid[0] = la[0];
    grid[2] = ls[0].x / n;
    grid[3] = ls[0].y / n;
    grid[4] = ls[0].z / n;
  }
}

// block maxima of |displacement - mean| over the free vertices
__global__ void __launch_bounds__(256) k_deviation(const float4* __restrict__ dir, const int32_t* __restrict__ row_of,
                                                   int V, const float* __restrict__ grid, float* __restrict__ part) {
  __shared__ float lb[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float m = 0.0f;
  if (i < V && row_of[i] >= 0) {
    const float4 d = dir[i];
    const float dx = d.x - grid[2], dy = d.y - grid[3], dz = d.z - grid[4];
    m = sqrtf(dx * dx + dy * dy + dz * dz);
  }
  lb[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lb[threadIdx.x] = fmaxf(lb[threadIdx.x], lb[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[5 * blockIdx.x + 1] = lb[0];
}

// grid[1] = inflation = dhat + 2 max |displacement - mean| (any two vertices' relative displacement is at
// most that; 0 without a step), grid[0] = cell size = longest edge + inflation; both rounded up slightly
__global__ void k_grid(const float* __restrict__ part, int groups, int step, float dhat, float* __restrict__ grid) {
  __shared__ float lb[256];
  float b = 0.0f;
  if (step)
    for (int j = threadIdx.x; j < groups; j += 256) b = fmaxf(b, part[5 * j + 1]);
  lb[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lb[threadIdx.x] = fmaxf(lb[threadIdx.x], lb[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float infl = (dhat + 2.0f * lb[0]) * 1.0001f;
    grid[1] = infl;
    grid[0] = (grid[0] + infl) * 1.0001f;
  }
}

__global__ void __launch_bounds__(256) k_vertex_keys(const float4* __restrict__ pos, int V,
                                                     const float* __restrict__ grid, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ id) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  key[v] = cell_key(cell_of(p3(pos[v]), grid[0]));
  id[v] = (uint32_t)v;
}

__device__ inline P3 midpoint(const float4* pos, int2 e) {
  const P3 a = p3(pos[e.x]), b = p3(pos[e.y]);
  return {0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
}

__global__ void __launch_bounds__(256) k_edge_keys(const int2* __restrict__ edges, int E,
                                                   const float4* __restrict__ pos, const float* __restrict__ grid,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ id) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= E) return;
  key[e] = cell_key(cell_of(midpoint(pos, edges[e]), grid[0]));
  id[e] = (uint32_t)e;
}

// point-triangle candidates of triangle t: EMIT 0 counts, 1 writes at out + offset[t]
template <int EMIT>
__global__ void __launch_bounds__(256) k_pt(const int4* __restrict__ tri, int T, const float4* __restrict__ pos,
                                            int V, const float* __restrict__ grid, double limit, double dhat,
                                            const float4* __restrict__ dir, const int32_t* __restrict__ row_of,
                                            const uint32_t* __restrict__ vkey, const uint32_t* __restrict__ vid,
                                            uint32_t* __restrict__ count, const uint32_t* __restrict__ offset,
                                            int4* __restrict__ out) {
  const int slot = blockIdx.x * 256 + threadIdx.x;  // triangle t, neighbour cell k
  if (slot >= 27 * T) return;
  const int t = slot / 27, k = slot % 27;
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  const double h = grid[0], infl = limit > 0.0 ? limit : grid[1];
  const int4 tv = tri[t];
  const P3 a = p3(pos[tv.x]), b = p3(pos[tv.y]), c = p3(pos[tv.z]);
  const P3 cen = {(a.x + b.x + c.x) / 3.0, (a.y + b.y + c.y) / 3.0, (a.z + b.z + c.z) / 3.0};
  const double reach = fmax(dist3(a, cen), fmax(dist3(b, cen), dist3(c, cen)));
  const int3 cc = cell_of(cen, h);
  uint32_t n = 0, at = EMIT ? offset[slot] : 0;
  {
        const int3 q = make_int3(cc.x + dx, cc.y + dy, cc.z + dz);
        const uint32_t key = cell_key(q);
        for (uint32_t i = lower_bound_u32(vkey, V, key); i < (uint32_t)V && vkey[i] == key; ++i) {
          const int v = (int)vid[i];
          if (v == tv.x || v == tv.y || v == tv.z) continue;
          const P3 x = p3(pos[v]);
          if (!same_cell(cell_of(x, h), q)) continue;
          if (dist3(x, cen) - reach >= infl) continue;  // no closer than the bound on any pair
          const int4 quad = make_int4(v, tv.x, tv.y, tv.z);
          // along a step: the pairs that may come within dhat
          const double lim = dir ? (dhat + approach(dir, row_of, true, quad)) * (1.0 + 1e-9) : infl;
          if (dist3(x, cen) - reach >= lim) continue;
          if (pair_geometry(true, x, a, b, c).d >= lim) continue;
          if (EMIT) out[at++] = quad;
          ++n;
        }
      }
  if (!EMIT) count[slot] = n;
}

// edge-edge candidates (e, f) with f > e and no shared vertex
template <int EMIT>
__global__ void __launch_bounds__(256) k_ee(const int2* __restrict__ edges, int E, const float4* __restrict__ pos,
                                            const float* __restrict__ grid, double limit, double dhat,
                                            const float4* __restrict__ dir, const int32_t* __restrict__ row_of,
                                            const uint32_t* __restrict__ ekey,
                                            const uint32_t* __restrict__ eid, uint32_t* __restrict__ count,
                                            const uint32_t* __restrict__ offset, int4* __restrict__ out) {
  const int slot = blockIdx.x * 256 + threadIdx.x;  // edge e, neighbour cell k
  if (slot >= 27 * E) return;
  const int e = slot / 27, k = slot % 27;
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  const double h = grid[0], infl = limit > 0.0 ? limit : grid[1];
  const int2 ev = edges[e];
  const P3 p1 = p3(pos[ev.x]), q1 = p3(pos[ev.y]);
  const P3 m1 = midpoint(pos, ev);
  const double half1 = 0.5 * dist3(p1, q1);
  const int3 cc = cell_of(m1, h);
  uint32_t n = 0, at = EMIT ? offset[slot] : 0;
  {
        const int3 q = make_int3(cc.x + dx, cc.y + dy, cc.z + dz);
        const uint32_t key = cell_key(q);
        for (uint32_t i = lower_bound_u32(ekey, E, key); i < (uint32_t)E && ekey[i] == key; ++i) {
          const int f = (int)eid[i];
          if (f <= e) continue;
          const int2 fv = edges[f];
          if (fv.x == ev.x || fv.x == ev.y || fv.y == ev.x || fv.y == ev.y) continue;
          const P3 m2 = midpoint(pos, fv);
          if (!same_cell(cell_of(m2, h), q)) continue;
          const P3 p2 = p3(pos[fv.x]), q2 = p3(pos[fv.y]);
          if (dist3(m1, m2) - half1 - 0.5 * dist3(p2, q2) >= infl) continue;
          const int4 quad = make_int4(ev.x, ev.y, fv.x, fv.y);
          const double lim = dir ? (dhat + approach(dir, row_of, false, quad)) * (1.0 + 1e-9) : infl;
          if (dist3(m1, m2) - half1 - 0.5 * dist3(p2, q2) >= lim) continue;
          if (pair_geometry(false, p1, q1, p2, q2).d >= lim) continue;
          if (EMIT) out[at++] = quad;
          ++n;
        }
      }
  if (!EMIT) count[slot] = n;
}

// ---- swept boxes of the step ------------------------------------------------------------------------
// A pair that comes within dhat somewhere along x + t dir, t in [0, 1], has its point (edge) inside its own
// box over the step and its triangle (other edge) inside the other's, within dhat: the boxes of the
// vertices and edges are hashed into every cell they overlap, and each triangle (edge) queries the cells of
// its own box grown by dhat, one thread per (primitive, queried cell). A pair is reported from the lowest
// cell common to both boxes' cell ranges only, and only when the queried cell lies in the stored box's range
// (a hash collision otherwise).

// Most cells a stored box may span, and a queried one; past them (a step far longer than the edges) the
// step's set is found on the inflated grid instead.
constexpr int BOX_CELLS = 512;
constexpr int QUERY_CELLS = 4096;

__device__ inline void grow_box(P3 x, P3& lo, P3& hi) {
  lo = {fmin(lo.x, x.x), fmin(lo.y, x.y), fmin(lo.z, x.z)};
  hi = {fmax(hi.x, x.x), fmax(hi.y, x.y), fmax(hi.z, x.z)};
}
// the vertices of primitive i: KIND 0 vertex, 1 edge, 2 triangle
template <int KIND>
__device__ inline int prim_ids(int i, const int2* edges, const int4* tri, int ids[3]) {
  if (KIND == 0) {
    ids[0] = i;
    return 1;
  }
  if (KIND == 1) {
    const int2 e = edges[i];
    ids[0] = e.x;
    ids[1] = e.y;
    return 2;
  }
  const int4 t = tri[i];
  ids[0] = t.x;
  ids[1] = t.y;
  ids[2] = t.z;
  return 3;
}
// Cell range c0..c1 of the box over the step of vertices ids[0..n), grown by pad; false when it is not
// finite or lies past 1e9 cells from the origin.
__device__ inline bool step_range(const float4* pos, const float4* dir, const int32_t* row_of, const int* ids, int n,
                                  double h, double pad, int3& c0, int3& c1) {
  P3 lo = {INFINITY, INFINITY, INFINITY}, hi = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < n; ++k) {
    const P3 x = p3(pos[ids[k]]);
    grow_box(x, lo, hi);
    grow_box(axpy(x, 1.0, disp(dir, row_of, ids[k])), lo, hi);
  }
  const double f0x = floor((lo.x - pad) / h), f0y = floor((lo.y - pad) / h), f0z = floor((lo.z - pad) / h);
  const double f1x = floor((hi.x + pad) / h), f1y = floor((hi.y + pad) / h), f1z = floor((hi.z + pad) / h);
  const double m = fmax(fmax(fmax(fabs(f0x), fabs(f0y)), fmax(fabs(f0z), fabs(f1x))), fmax(fabs(f1y), fabs(f1z)));
  if (!(m < 1e9)) return false;
  c0 = make_int3((int)f0x, (int)f0y, (int)f0z);
  c1 = make_int3((int)f1x, (int)f1y, (int)f1z);
  return true;
}
__device__ inline long long range_cells(int3 c0, int3 c1) {
  return (long long)(c1.x - c0.x + 1) * (c1.y - c0.y + 1) * (c1.z - c0.z + 1);
}
__device__ inline bool in_range(int3 c, int3 a0, int3 a1) {
  return c.x >= a0.x && c.x <= a1.x && c.y >= a0.y && c.y <= a1.y && c.z >= a0.z && c.z <= a1.z;
}

// count[i] = cells of primitive i's box over the step grown by pad; a box past `most` cells, or not finite,
// sets *over
template <int KIND>
__global__ void __launch_bounds__(256) k_span(const int2* __restrict__ edges, const int4* __restrict__ tri, int n,
                                              const float4* __restrict__ pos, const float4* __restrict__ dir,
                                              const int32_t* __restrict__ row_of, const float* __restrict__ grid,
                                              double pad, int most, uint32_t* __restrict__ count,
                                              uint32_t* __restrict__ over) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int ids[3];
  const int k = prim_ids<KIND>(i, edges, tri, ids);
  int3 c0, c1;
  const bool ok = step_range(pos, dir, row_of, ids, k, grid[0], pad, c0, c1);
  const long long cells = ok ? range_cells(c0, c1) : 0;
  const bool fits = ok && cells <= most;
  if (!fits) *over = 1u;
  count[i] = fits ? (uint32_t)cells : 0u;
}

// (cell key, i) for every cell of primitive i's box over the step, at key + offset[i] (KIND 0 vertex, 1 edge)
template <int KIND>
__global__ void __launch_bounds__(256) k_box_cells(const int2* __restrict__ edges, int n,
                                                   const float4* __restrict__ pos, const float4* __restrict__ dir,
                                                   const int32_t* __restrict__ row_of, const float* __restrict__ grid,
                                                   const uint32_t* __restrict__ offset, uint32_t* __restrict__ key,
                                                   uint32_t* __restrict__ id) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int ids[3];
  const int k = prim_ids<KIND>(i, edges, nullptr, ids);
  int3 c0, c1;
  step_range(pos, dir, row_of, ids, k, grid[0], 0.0, c0, c1);  // in range: its span was counted
  uint32_t at = offset[i];
  for (int z = c0.z; z <= c1.z; ++z)
    for (int y = c0.y; y <= c1.y; ++y)
      for (int x = c0.x; x <= c1.x; ++x) {
        key[at] = cell_key(make_int3(x, y, z));
        id[at] = (uint32_t)i;
        ++at;
      }
}

// The primitive of query slot `slot` (the last p with qoff[p] <= slot) and its queried cell.
__device__ inline int slot_owner(const uint32_t* qoff, int n, uint32_t slot) {
  uint32_t lo = 0, hi = (uint32_t)n + 1;  // first index with qoff > slot
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (qoff[mid] <= slot) lo = mid + 1; else hi = mid;
  }
  return (int)lo - 1;
}
__device__ inline int3 slot_cell(int3 q0, int3 q1, uint32_t local) {
  const uint32_t nx = (uint32_t)(q1.x - q0.x + 1), ny = (uint32_t)(q1.y - q0.y + 1);
  return make_int3(q0.x + (int)(local % nx), q0.y + (int)((local / nx) % ny), q0.z + (int)(local / (nx * ny)));
}

// point-triangle candidates of query slot (triangle t, queried cell q) along the step: EMIT 0 counts,
// 1 writes at out + offset[slot]
template <int EMIT>
__global__ void __launch_bounds__(256) k_swept_pt(const int4* __restrict__ tri, int T, const uint32_t* __restrict__ qoff,
                                                  uint32_t slots, const float4* __restrict__ pos,
                                                  const float4* __restrict__ dir, const int32_t* __restrict__ row_of,
                                                  const float* __restrict__ grid, double dhat,
                                                  const uint32_t* __restrict__ key, const uint32_t* __restrict__ vid,
                                                  uint32_t nkeys, uint32_t* __restrict__ count,
                                                  const uint32_t* __restrict__ offset, int4* __restrict__ out) {
  const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
  if (slot >= slots) return;
  const int t = slot_owner(qoff, T, slot);
  const double h = grid[0];
  const int4 tv = tri[t];
  const int tids[3] = {tv.x, tv.y, tv.z};
  int3 q0, q1;
  step_range(pos, dir, row_of, tids, 3, h, dhat, q0, q1);
  const int3 q = slot_cell(q0, q1, slot - qoff[t]);
  const P3 a = p3(pos[tv.x]), b = p3(pos[tv.y]), c = p3(pos[tv.z]);
  const P3 cen = {(a.x + b.x + c.x) / 3.0, (a.y + b.y + c.y) / 3.0, (a.z + b.z + c.z) / 3.0};
  const double reach = fmax(dist3(a, cen), fmax(dist3(b, cen), dist3(c, cen)));
  const uint32_t k = cell_key(q);
  uint32_t n = 0, at = EMIT ? offset[slot] : 0;
  for (uint32_t i = lower_bound_u32(key, nkeys, k); i < nkeys && key[i] == k; ++i) {
    const int v = (int)vid[i];
    if (v == tv.x || v == tv.y || v == tv.z) continue;
    if (i > 0 && key[i - 1] == k && (int)vid[i - 1] == v) continue;  // two of its cells share the key
    int3 v0, v1;
    step_range(pos, dir, row_of, &v, 1, h, 0.0, v0, v1);
    if (!in_range(q, v0, v1)) continue;
    if (q.x != max(v0.x, q0.x) || q.y != max(v0.y, q0.y) || q.z != max(v0.z, q0.z)) continue;
    const int4 quad = make_int4(v, tv.x, tv.y, tv.z);
    const double lim = (dhat + approach(dir, row_of, true, quad)) * (1.0 + 1e-9);
    const P3 xv = p3(pos[v]);
    if (dist3(xv, cen) - reach >= lim) continue;
    if (pair_geometry(true, xv, a, b, c).d >= lim) continue;
    if (EMIT) out[at++] = quad;
    ++n;
  }
  if (!EMIT) count[slot] = n;
}

// edge-edge candidates (e, f), f > e, without a shared vertex, of query slot (edge e, queried cell q)
template <int EMIT>
__global__ void __launch_bounds__(256) k_swept_ee(const int2* __restrict__ edges, int E,
                                                  const uint32_t* __restrict__ qoff, uint32_t slots,
                                                  const float4* __restrict__ pos, const float4* __restrict__ dir,
                                                  const int32_t* __restrict__ row_of, const float* __restrict__ grid,
                                                  double dhat, const uint32_t* __restrict__ key,
                                                  const uint32_t* __restrict__ eid, uint32_t nkeys,
                                                  uint32_t* __restrict__ count, const uint32_t* __restrict__ offset,
                                                  int4* __restrict__ out) {
  const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
  if (slot >= slots) return;
  const int e = slot_owner(qoff, E, slot);
  const double h = grid[0];
  const int2 ev = edges[e];
  const int eids[2] = {ev.x, ev.y};
  int3 q0, q1;
  step_range(pos, dir, row_of, eids, 2, h, dhat, q0, q1);
  const int3 q = slot_cell(q0, q1, slot - qoff[e]);
  const P3 p1 = p3(pos[ev.x]), q1p = p3(pos[ev.y]);
  const P3 m1 = midpoint(pos, ev);
  const double half1 = 0.5 * dist3(p1, q1p);
  const uint32_t k = cell_key(q);
  uint32_t n = 0, at = EMIT ? offset[slot] : 0;
  for (uint32_t i = lower_bound_u32(key, nkeys, k); i < nkeys && key[i] == k; ++i) {
    const int f = (int)eid[i];
    if (f <= e) continue;
    if (i > 0 && key[i - 1] == k && (int)eid[i - 1] == f) continue;  // two of its cells share the key
    const int2 fv = edges[f];
    if (fv.x == ev.x || fv.x == ev.y || fv.y == ev.x || fv.y == ev.y) continue;
    const int fids[2] = {fv.x, fv.y};
    int3 f0, f1;
    step_range(pos, dir, row_of, fids, 2, h, 0.0, f0, f1);
    if (!in_range(q, f0, f1)) continue;
    if (q.x != max(f0.x, q0.x) || q.y != max(f0.y, q0.y) || q.z != max(f0.z, q0.z)) continue;
    const int4 quad = make_int4(ev.x, ev.y, fv.x, fv.y);
    const double lim = (dhat + approach(dir, row_of, false, quad)) * (1.0 + 1e-9);
    const P3 p2 = p3(pos[fv.x]), q2 = p3(pos[fv.y]);
    const P3 m2 = midpoint(pos, fv);
    if (dist3(m1, m2) - half1 - 0.5 * dist3(p2, q2) >= lim) continue;
    if (pair_geometry(false, p1, q1p, p2, q2).d >= lim) continue;
    if (EMIT) out[at++] = quad;
    ++n;
  }
  if (!EMIT) count[slot] = n;
}

// ---- pair terms ----------------------------------------------------------------------------------

// distance, direction, coefficients and barrier derivatives of each candidate; the incidence entries
// (row, 4 pair + k) of the active pairs' free vertices, others keyed past every row
__global__ void __launch_bounds__(256) k_pair_terms(const int4* __restrict__ cand, int n, int npt,
                                                    const float4* __restrict__ pos, const int32_t* __restrict__ row_of,
                                                    double dhat, double kappa, int4* __restrict__ crow,
                                                    float4* __restrict__ cdat, float4* __restrict__ ccoef,
                                                    float* __restrict__ cfirst, double* __restrict__ dist,
                                                    uint32_t* __restrict__ ikey, uint32_t* __restrict__ ival) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int4 v = cand[c];
  const PairGeom g = pair_geometry(c < npt, p3(pos[v.x]), p3(pos[v.y]), p3(pos[v.z]), p3(pos[v.w]));
  const bool active = g.d < dhat && g.d > 0.0;
  const BarrierTerms b = active ? barrier_terms(g.d, dhat, kappa) : BarrierTerms{0.0, 0.0, 0.0};
  const int4 rows = make_int4(row_of[v.x], row_of[v.y], row_of[v.z], row_of[v.w]);
  crow[c] = rows;
  cdat[c] = make_float4((float)g.n.x, (float)g.n.y, (float)g.n.z, (float)b.second);
  ccoef[c] = make_float4((float)g.c[0], (float)g.c[1], (float)g.c[2], (float)g.c[3]);
  cfirst[c] = (float)b.first;
  dist[c] = g.d;
  const int r[4] = {rows.x, rows.y, rows.z, rows.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ikey[4 * c + k] = (active && r[k] >= 0) ? (uint32_t)r[k] : 0xffffffffu;
    ival[4 * c + k] = 4u * (uint32_t)c + (uint32_t)k;
  }
}

__global__ void __launch_bounds__(256) k_row_ptr(const uint32_t* __restrict__ keys, uint32_t n, int R,
                                                 uint32_t* __restrict__ ptr) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r > R) return;
  ptr[r] = lower_bound_u32(keys, n, (uint32_t)r);
}

// per row, in its list order: b -= b' c_k n and the diagonal block += b'' c_k^2 n n^T
__global__ void __launch_bounds__(256) k_contact_assemble(const uint32_t* __restrict__ iptr,
                                                          const uint32_t* __restrict__ ival,
                                                          const float4* __restrict__ cdat,
                                                          const float4* __restrict__ ccoef,
                                                          const float* __restrict__ cfirst,
                                                          const uint32_t* __restrict__ diag_slot, int R,
                                                          float* __restrict__ sell_val, float4* __restrict__ b) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= R) return;
  const uint32_t t0 = iptr[row], t1 = iptr[row + 1];
  if (t0 == t1) return;
  float4 rhs = b[row];
  float* d = sell_val + diag_slot[row];
  float blk[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) blk[e] = d[64 * e];
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t e = ival[t];
    const uint32_t pair = e >> 2;
    const int k = (int)(e & 3u);
    const float4 dn = cdat[pair];
    const float ck = coef_of(ccoef[pair], k);
    const float g = cfirst[pair] * ck;
    rhs.x = rhs.x - g * dn.x;
    rhs.y = rhs.y - g * dn.y;
    rhs.z = rhs.z - g * dn.z;
    const float s = dn.w * ck * ck;
    const float n3[3] = {dn.x, dn.y, dn.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) blk[3 * i + j] = blk[3 * i + j] + s * n3[i] * n3[j];
  }
  b[row] = rhs;
#pragma unroll
  for (int e = 0; e < 9; ++e) d[64 * e] = blk[e];
}

// ---- line search ---------------------------------------------------------------------------------

// FP64 partial sums of (b(d(x + alpha dir)) - b(d(x)), |same|) over the candidates
__global__ void __launch_bounds__(256) k_contact_energy(const int4* __restrict__ cand, int n, int npt,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ dir,
                                                        const int32_t* __restrict__ row_of, float alpha, double dhat,
                                                        double kappa, double* __restrict__ part) {
  __shared__ double2 lds[256];
  const int c = blockIdx.x * 256 + threadIdx.x;
  double delta = 0.0;
  if (c < n) {
    const int4 v = cand[c];
    const bool pt = c < npt;
    const int ids[4] = {v.x, v.y, v.z, v.w};
    P3 x[4], y[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float4 xp = pos[ids[k]];
      x[k] = p3(xp);
      // the FP32 step the advance applies to free vertices: pos + alpha dir
      const float4 d = row_of[ids[k]] >= 0 ? dir[ids[k]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      y[k] = {(double)(xp.x + alpha * d.x), (double)(xp.y + alpha * d.y), (double)(xp.z + alpha * d.z)};
    }
    const double before = pair_geometry(pt, x[0], x[1], x[2], x[3]).d;
    const double after = pair_geometry(pt, y[0], y[1], y[2], y[3]).d;
    delta = barrier_terms(after, dhat, kappa).value - barrier_terms(before, dhat, kappa).value;
  }
  const double2 t = tree256d2(delta, fabs(delta), lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t.x;
    part[2 * blockIdx.x + 1] = t.y;
  }
}

__global__ void k_sum_pairs64(const double* __restrict__ part, int count, double* __restrict__ out) {
  __shared__ double2 lds[256];
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < count; j += 256) {
    a += part[2 * j];
    b += part[2 * j + 1];
  }
  const double2 t = tree256d2(a, b, lds);
  if (threadIdx.x == 0) {
    out[0] = t.x;
    out[1] = t.y;
  }
}

// conservative advancement of each candidate along x + t dir, t in [0, 1]; group minima (lowest pair on
// ties)
__global__ void __launch_bounds__(256) k_contact_ccd(const int4* __restrict__ cand, int n, int npt,
                                                     const float4* __restrict__ pos, const float4* __restrict__ dir,
                                                     const int32_t* __restrict__ row_of, float* __restrict__ apart,
                                                     int32_t* __restrict__ aidx) {
  __shared__ float sa[256];
  __shared__ int ia[256];
  const int c = blockIdx.x * 256 + threadIdx.x;
  float alpha = 1.0f;
  if (c < n) {
    const int4 v = cand[c];
    const bool pt = c < npt;
    const int ids[4] = {v.x, v.y, v.z, v.w};
    P3 x[4], d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      x[k] = p3(pos[ids[k]]);
      d[k] = disp(dir, row_of, ids[k]);
    }
    const double lp = approach(dir, row_of, pt, v);
    const double d0 = pair_geometry(pt, x[0], x[1], x[2], x[3]).d;
    if (lp > 0.0 && d0 > 0.0) {
      const double target = 0.1 * d0;
      double t = 0.0, dist = d0;
      bool done = false;
      for (int it = 0; it < 64 && !done; ++it) {
        const double dt = 0.9 * (dist - target) / lp;
        if (t + dt >= 1.0) {
          t = 1.0;
          done = true;
        } else if (dt <= 1e-9 * (t + dt)) {
          done = true;
        } else {
          t += dt;
          P3 y[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) y[k] = axpy(x[k], t, d[k]);
          dist = pair_geometry(pt, y[0], y[1], y[2], y[3]).d;
        }
      }
      alpha = (float)t;
      if ((double)alpha > t) alpha = nextafterf(alpha, 0.0f);  // never round the bound up
    } else if (d0 <= 0.0) {
      alpha = 0.0f;
    }
  }
  sa[threadIdx.x] = alpha;
  ia[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const int t = threadIdx.x, u = t + s;
      if (sa[u] < sa[t] || (sa[u] == sa[t] && ia[u] < ia[t])) {
        sa[t] = sa[u];
        ia[t] = ia[u];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    apart[blockIdx.x] = sa[0];
    aidx[blockIdx.x] = ia[0];
  }
}

__global__ void k_min_final(const float* __restrict__ apart, const int32_t* __restrict__ aidx, int count,
                            const int4* __restrict__ cand, const int32_t* __restrict__ row_of,
                            double* __restrict__ out) {
  __shared__ float sa[256];
  __shared__ int ia[256];
  float a = 1.0f;
  int ai = 0x7fffffff;
  for (int j = threadIdx.x; j < count; j += 256)
    if (apart[j] < a || (apart[j] == a && aidx[j] < ai)) {
      a = apart[j];
      ai = aidx[j];
    }
  sa[threadIdx.x] = a;
  ia[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const int t = threadIdx.x, u = t + s;
      if (sa[u] < sa[t] || (sa[u] == sa[t] && ia[u] < ia[t])) {
        sa[t] = sa[u];
        ia[t] = ia[u];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sa[0];
    int row = -1;
    if (ia[0] != 0x7fffffff) {
      const int4 v = cand[ia[0]];
      const int ids[4] = {v.x, v.y, v.z, v.w};
      for (int k = 0; k < 4 && row < 0; ++k) row = row_of[ids[k]];
    }
    out[1] = row;
  }
}

template <typename T>
int grow(T** p, size_t n) {
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  HIP_TRY(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
  return DRAPE_OK;
}

int setup(Context& c) {
  if (c.self_contact) return DRAPE_OK;
  if (c.K <= 0) {
    set_error("self-contact takes its barrier (dhat, kappa) from the colliders; the scene has none");
    return DRAPE_E_INVALID;
  }
  SelfContact* sc = new SelfContact();
  c.self_contact = sc;
  sc->V = c.V;
  sc->T = c.T;
  hipStream_t s = c.stream;
  // unique edges of the triangles, and the row of each vertex
  std::vector<int4> tri(c.T);
  HIP_TRY(hipMemcpyAsync(tri.data(), c.tri, (size_t)c.T * sizeof(int4), hipMemcpyDeviceToHost, s));
  float hdr[4];
  HIP_TRY(hipMemcpyAsync(hdr, c.colliders, sizeof(hdr), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  sc->dhat = hdr[1];
  sc->kappa = hdr[2];
  std::vector<uint64_t> keys;
  keys.reserve(3 * (size_t)c.T);
  for (const int4& t : tri) {
    const int v[3] = {t.x, t.y, t.z};
    for (int k = 0; k < 3; ++k) {
      const uint32_t a = (uint32_t)std::min(v[k], v[(k + 1) % 3]), b = (uint32_t)std::max(v[k], v[(k + 1) % 3]);
      keys.push_back(((uint64_t)a << 32) | b);
    }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  std::vector<int2> edges(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) edges[i] = make_int2((int)(keys[i] >> 32), (int)(keys[i] & 0xffffffffu));
  sc->E = (int)edges.size();
  std::vector<int32_t> row_of(c.V, -1);
  for (int r = 0; r < c.R; ++r) row_of[c.row_vertex[r]] = r;
  const int V = c.V, T = c.T, E = sc->E;
  TRY_RC(grow(&sc->d_edges, E));
  TRY_RC(grow(&sc->d_row_of, V));
  for (uint32_t** p : {&sc->d_vkey, &sc->d_vid, &sc->d_vkey_s, &sc->d_vid_s}) TRY_RC(grow(p, V));
  for (uint32_t** p : {&sc->d_ekey, &sc->d_eid, &sc->d_ekey_s, &sc->d_eid_s}) TRY_RC(grow(p, E));
  // per (triangle, neighbour cell), and per vertex for the step's boxes
  TRY_RC(grow(&sc->d_tcount, std::max(27 * (size_t)T, (size_t)V) + 1));
  TRY_RC(grow(&sc->d_toff, std::max(27 * (size_t)T, (size_t)V) + 1));
  TRY_RC(grow(&sc->d_ecount, 27 * (size_t)E + 1));
  TRY_RC(grow(&sc->d_eoff, 27 * (size_t)E + 1));
  TRY_RC(grow(&sc->d_part, 5 * (size_t)blocks_for(std::max(V, E))));
  TRY_RC(grow(&sc->d_grid, 5));
  TRY_RC(grow(&sc->d_iptr, (size_t)c.R + 1));
  TRY_RC(grow(&sc->d_eout, 2));
  TRY_RC(grow(&sc->d_over, 1));
  TRY_RC(grow(&sc->d_tq, (size_t)T + 1));
  TRY_RC(grow(&sc->d_tqoff, (size_t)T + 1));
  HIP_TRY(hipHostMalloc((void**)&sc->h_totals, 8 * sizeof(uint32_t)));
  // scratch for the largest sort and scans
  size_t need = 0, b = 0;
  const int most = std::max(std::max(V + 1, 27 * E + 1), 27 * T + 1);
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (uint32_t*)nullptr, most));
  need = std::max(need, b);
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr, most));
  need = std::max(need, b);
  sc->tmp_bytes = need;
  HIP_TRY(hipMalloc(&sc->d_tmp, std::max<size_t>(need, 1)));
  HIP_TRY(hipMemcpyAsync(sc->d_edges, edges.data(), (size_t)E * sizeof(int2), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(sc->d_row_of, row_of.data(), (size_t)V * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

int ensure_tmp(SelfContact* sc, size_t bytes) {
  if (bytes <= sc->tmp_bytes) return DRAPE_OK;
  HIP_TRY(hipFree(sc->d_tmp));
  sc->d_tmp = nullptr;
  HIP_TRY(hipMalloc(&sc->d_tmp, bytes));
  sc->tmp_bytes = bytes;
  return DRAPE_OK;
}

int sort_pairs(SelfContact* sc, const uint32_t* k_in, uint32_t* k_out, const uint32_t* v_in, uint32_t* v_out, int n,
               hipStream_t s) {
  if (n <= 0) return DRAPE_OK;
  size_t b = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, k_in, k_out, v_in, v_out, n, 0, 32, s));
  TRY_RC(ensure_tmp(sc, b));
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(sc->d_tmp, b, k_in, k_out, v_in, v_out, n, 0, 32, s));
  return DRAPE_OK;
}

int scan(SelfContact* sc, const uint32_t* in, uint32_t* out, int n, hipStream_t s) {
  size_t b = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, in, out, n, s));
  TRY_RC(ensure_tmp(sc, b));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(sc->d_tmp, b, in, out, n, s));
  return DRAPE_OK;
}

// Candidates at the current positions closer than dhat, or (dir not null) than dhat + 2 |dir|max (rounded
// up), into `set`.
int broad_phase(Context& c, const float4* dir, SelfContact::Set& set) {
  SelfContact* sc = c.self_contact;
  const double limit = dir ? 0.0 : sc->dhat;
  hipStream_t s = c.stream;
  const int V = sc->V, T = sc->T, E = sc->E;
  const int groups = blocks_for(std::max(V, E));
  k_extent<<<groups, 256, 0, s>>>(sc->d_edges, E, c.pos, dir, sc->d_row_of, V, sc->d_part);
  k_mean<<<1, 256, 0, s>>>(sc->d_part, groups, sc->d_grid);
  if (dir) k_deviation<<<blocks_for(V), 256, 0, s>>>(dir, sc->d_row_of, V, sc->d_grid, sc->d_part);
  k_grid<<<1, 256, 0, s>>>(sc->d_part, dir ? blocks_for(V) : 0, dir ? 1 : 0, (float)sc->dhat, sc->d_grid);
  k_vertex_keys<<<blocks_for(V), 256, 0, s>>>(c.pos, V, sc->d_grid, sc->d_vkey, sc->d_vid);
  k_edge_keys<<<blocks_for(E), 256, 0, s>>>(sc->d_edges, E, c.pos, sc->d_grid, sc->d_ekey, sc->d_eid);
  HIP_TRY(hipGetLastError());
  TRY_RC(sort_pairs(sc, sc->d_vkey, sc->d_vkey_s, sc->d_vid, sc->d_vid_s, V, s));
  TRY_RC(sort_pairs(sc, sc->d_ekey, sc->d_ekey_s, sc->d_eid, sc->d_eid_s, E, s));
  const int ST = 27 * T, SE = 27 * E;  // (primitive, neighbour cell) slots
  k_pt<0><<<blocks_for(ST), 256, 0, s>>>(c.tri, T, c.pos, V, sc->d_grid, limit, sc->dhat, dir, sc->d_row_of, sc->d_vkey_s, sc->d_vid_s, sc->d_tcount,
                                        nullptr, nullptr);
  k_ee<0><<<blocks_for(SE), 256, 0, s>>>(sc->d_edges, E, c.pos, sc->d_grid, limit, sc->dhat, dir, sc->d_row_of, sc->d_ekey_s, sc->d_eid_s, sc->d_ecount,
                                        nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  // totals: the scans' last entries (the counts' last slots are zero)
  HIP_TRY(hipMemsetAsync(sc->d_tcount + ST, 0, sizeof(uint32_t), s));
  HIP_TRY(hipMemsetAsync(sc->d_ecount + SE, 0, sizeof(uint32_t), s));
  TRY_RC(scan(sc, sc->d_tcount, sc->d_toff, ST + 1, s));
  TRY_RC(scan(sc, sc->d_ecount, sc->d_eoff, SE + 1, s));
  HIP_TRY(hipMemcpyAsync(sc->h_totals, sc->d_toff + ST, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(sc->h_totals + 1, sc->d_eoff + SE, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c.read_bytes += 2 * sizeof(uint32_t);
  const int npt = (int)sc->h_totals[0], nee = (int)sc->h_totals[1], n = npt + nee;
  if (n > set.cap) {
    const int cap = std::max(n, 2 * set.cap);
    TRY_RC(grow(&set.d_cand, cap));
    set.cap = cap;
  }
  set.npt = npt;
  set.nee = nee;
  if (npt > 0)
    k_pt<1><<<blocks_for(ST), 256, 0, s>>>(c.tri, T, c.pos, V, sc->d_grid, limit, sc->dhat, dir, sc->d_row_of, sc->d_vkey_s, sc->d_vid_s, nullptr,
                                          sc->d_toff, set.d_cand);
  if (nee > 0)
    k_ee<1><<<blocks_for(SE), 256, 0, s>>>(sc->d_edges, E, c.pos, sc->d_grid, limit, sc->dhat, dir, sc->d_row_of, sc->d_ekey_s, sc->d_eid_s, nullptr,
                                          sc->d_eoff, set.d_cand + npt);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

// The step's line-search set from the swept boxes (cell size: longest edge + dhat). Returns with *done false,
// and nothing written to `set`, when a box spans more cells than its limit, or the boxes more than 32 (stored)
// or 64 (queried) cells per primitive on average.
int swept_phase(Context& c, SelfContact::Set& set, bool* done) {
  SelfContact* sc = c.self_contact;
  hipStream_t s = c.stream;
  const int V = sc->V, T = sc->T, E = sc->E;
  const float4* dir = c.dir;
  const double dhat = sc->dhat;
  *done = false;
  const int groups = blocks_for(std::max(V, E));
  k_extent<<<groups, 256, 0, s>>>(sc->d_edges, E, c.pos, nullptr, sc->d_row_of, V, sc->d_part);
  k_mean<<<1, 256, 0, s>>>(sc->d_part, groups, sc->d_grid);
  k_grid<<<1, 256, 0, s>>>(sc->d_part, 0, 0, (float)dhat, sc->d_grid);
  HIP_TRY(hipMemsetAsync(sc->d_over, 0, sizeof(uint32_t), s));
  // stored boxes: vertices (d_tcount, [V + 1]) and edges (d_ecount); queried: triangles (d_tq) and edges
  // (d_ecount past the edges' stored counts)
  uint32_t* eq = sc->d_ecount + (E + 1);
  uint32_t* eqoff = sc->d_eoff + (E + 1);
  k_span<0><<<blocks_for(V), 256, 0, s>>>(nullptr, nullptr, V, c.pos, dir, sc->d_row_of, sc->d_grid, 0.0, BOX_CELLS,
                                          sc->d_tcount, sc->d_over);
  k_span<1><<<blocks_for(E), 256, 0, s>>>(sc->d_edges, nullptr, E, c.pos, dir, sc->d_row_of, sc->d_grid, 0.0,
                                          BOX_CELLS, sc->d_ecount, sc->d_over);
  k_span<2><<<blocks_for(T), 256, 0, s>>>(nullptr, c.tri, T, c.pos, dir, sc->d_row_of, sc->d_grid, dhat, QUERY_CELLS,
                                          sc->d_tq, sc->d_over);
  k_span<1><<<blocks_for(E), 256, 0, s>>>(sc->d_edges, nullptr, E, c.pos, dir, sc->d_row_of, sc->d_grid, dhat,
                                          QUERY_CELLS, eq, sc->d_over);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(sc->d_tcount + V, 0, sizeof(uint32_t), s));
  HIP_TRY(hipMemsetAsync(sc->d_ecount + E, 0, sizeof(uint32_t), s));
  HIP_TRY(hipMemsetAsync(sc->d_tq + T, 0, sizeof(uint32_t), s));
  HIP_TRY(hipMemsetAsync(eq + E, 0, sizeof(uint32_t), s));
  TRY_RC(scan(sc, sc->d_tcount, sc->d_toff, V + 1, s));
  TRY_RC(scan(sc, sc->d_ecount, sc->d_eoff, E + 1, s));
  TRY_RC(scan(sc, sc->d_tq, sc->d_tqoff, T + 1, s));
  TRY_RC(scan(sc, eq, eqoff, E + 1, s));
  uint32_t* h = sc->h_totals;
  HIP_TRY(hipMemcpyAsync(h, sc->d_toff + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h + 1, sc->d_eoff + E, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h + 2, sc->d_tqoff + T, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h + 3, eqoff + E, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h + 4, sc->d_over, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c.read_bytes += 5 * sizeof(uint32_t);
  // the per-primitive limits bound every total below 2^32 for up to 2^22 primitives of a kind
  if (h[4] || h[0] > 32u * (uint32_t)V + 1024u || h[1] > 32u * (uint32_t)E + 1024u ||
      h[2] > 64u * (uint32_t)T + 1024u || h[3] > 64u * (uint32_t)E + 1024u)
    return DRAPE_OK;
  const int nbv = (int)h[0], nbe = (int)h[1];
  const uint32_t spt = h[2], see = h[3];
  if (nbv > sc->vcap) {
    const int cap = std::max(nbv, 2 * sc->vcap);
    for (uint32_t** p : {&sc->d_bvkey, &sc->d_bvid, &sc->d_bvkey_s, &sc->d_bvid_s}) TRY_RC(grow(p, cap));
    sc->vcap = cap;
  }
  if (nbe > sc->ecap) {
    const int cap = std::max(nbe, 2 * sc->ecap);
    for (uint32_t** p : {&sc->d_bekey, &sc->d_beid, &sc->d_bekey_s, &sc->d_beid_s}) TRY_RC(grow(p, cap));
    sc->ecap = cap;
  }
  if ((int)spt + 1 > sc->ptcap) {
    const int cap = std::max((int)spt + 1, 2 * sc->ptcap);
    TRY_RC(grow(&sc->d_ptcount, cap));
    TRY_RC(grow(&sc->d_ptoff, cap));
    sc->ptcap = cap;
  }
  if ((int)see + 1 > sc->eecap) {
    const int cap = std::max((int)see + 1, 2 * sc->eecap);
    TRY_RC(grow(&sc->d_eecount, cap));
    TRY_RC(grow(&sc->d_eeoff, cap));
    sc->eecap = cap;
  }
  k_box_cells<0><<<blocks_for(V), 256, 0, s>>>(nullptr, V, c.pos, dir, sc->d_row_of, sc->d_grid, sc->d_toff,
                                               sc->d_bvkey, sc->d_bvid);
  k_box_cells<1><<<blocks_for(E), 256, 0, s>>>(sc->d_edges, E, c.pos, dir, sc->d_row_of, sc->d_grid, sc->d_eoff,
                                               sc->d_bekey, sc->d_beid);
  HIP_TRY(hipGetLastError());
  TRY_RC(sort_pairs(sc, sc->d_bvkey, sc->d_bvkey_s, sc->d_bvid, sc->d_bvid_s, nbv, s));
  TRY_RC(sort_pairs(sc, sc->d_bekey, sc->d_bekey_s, sc->d_beid, sc->d_beid_s, nbe, s));
  const int bpt = blocks_for((int)spt), bee = blocks_for((int)see);
  if (spt > 0)
    k_swept_pt<0><<<bpt, 256, 0, s>>>(c.tri, T, sc->d_tqoff, spt, c.pos, dir, sc->d_row_of, sc->d_grid, dhat,
                                      sc->d_bvkey_s, sc->d_bvid_s, (uint32_t)nbv, sc->d_ptcount, nullptr, nullptr);
  if (see > 0)
    k_swept_ee<0><<<bee, 256, 0, s>>>(sc->d_edges, E, eqoff, see, c.pos, dir, sc->d_row_of, sc->d_grid, dhat,
                                      sc->d_bekey_s, sc->d_beid_s, (uint32_t)nbe, sc->d_eecount, nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(sc->d_ptcount + spt, 0, sizeof(uint32_t), s));
  HIP_TRY(hipMemsetAsync(sc->d_eecount + see, 0, sizeof(uint32_t), s));
  TRY_RC(scan(sc, sc->d_ptcount, sc->d_ptoff, (int)spt + 1, s));
  TRY_RC(scan(sc, sc->d_eecount, sc->d_eeoff, (int)see + 1, s));
  HIP_TRY(hipMemcpyAsync(h, sc->d_ptoff + spt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(h + 1, sc->d_eeoff + see, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c.read_bytes += 2 * sizeof(uint32_t);
  const int npt = (int)h[0], nee = (int)h[1], n = npt + nee;
  if (n > set.cap) {
    const int cap = std::max(n, 2 * set.cap);
    TRY_RC(grow(&set.d_cand, cap));
    set.cap = cap;
  }
  set.npt = npt;
  set.nee = nee;
  if (npt > 0)
    k_swept_pt<1><<<bpt, 256, 0, s>>>(c.tri, T, sc->d_tqoff, spt, c.pos, dir, sc->d_row_of, sc->d_grid, dhat,
                                      sc->d_bvkey_s, sc->d_bvid_s, (uint32_t)nbv, nullptr, sc->d_ptoff, set.d_cand);
  if (nee > 0)
    k_swept_ee<1><<<bee, 256, 0, s>>>(sc->d_edges, E, eqoff, see, c.pos, dir, sc->d_row_of, sc->d_grid, dhat,
                                      sc->d_bekey_s, sc->d_beid_s, (uint32_t)nbe, nullptr, sc->d_eeoff,
                                      set.d_cand + npt);
  HIP_TRY(hipGetLastError());
  *done = true;
  return DRAPE_OK;
}

// Term arrays for n pairs and partial buffers for n candidates.
int ensure_terms(SelfContact* sc, int n) {
  if (n > sc->tcap) {
    const int cap = std::max(n, 2 * sc->tcap);
    TRY_RC(grow(&sc->d_crow, cap));
    TRY_RC(grow(&sc->d_cdat, cap));
    TRY_RC(grow(&sc->d_ccoef, cap));
    TRY_RC(grow(&sc->d_cfirst, cap));
    TRY_RC(grow(&sc->d_dist, cap));
    for (uint32_t** p : {&sc->d_ikey, &sc->d_ival, &sc->d_ikey_s, &sc->d_ival_s}) TRY_RC(grow(p, 4 * (size_t)cap));
    sc->tcap = cap;
  }
  return DRAPE_OK;
}

int ensure_parts(SelfContact* sc, int n) {
  const int parts = blocks_for(n);
  if (parts > sc->parts) {
    TRY_RC(grow(&sc->d_epart, 2 * (size_t)parts));
    TRY_RC(grow(&sc->d_apart, parts));
    TRY_RC(grow(&sc->d_aidx, parts));
    sc->parts = parts;
  }
  return DRAPE_OK;
}

}  // namespace

void self_contact_set_kappa(Context& c, double kappa) {
  if (c.self_contact) c.self_contact->kappa = kappa;
}

int self_contact_evaluate(Context& c) {
  if (!c.rules.self_contact) return DRAPE_OK;
  TRY_RC(setup(c));
  SelfContact* sc = c.self_contact;
  hipStream_t s = c.stream;
  sc->step_ready = false;
  TRY_RC(broad_phase(c, nullptr, sc->eval));
  const int n = sc->eval.npt + sc->eval.nee;
  sc->terms = true;
  if (n == 0) return DRAPE_OK;
  TRY_RC(ensure_terms(sc, n));
  k_pair_terms<<<blocks_for(n), 256, 0, s>>>(sc->eval.d_cand, n, sc->eval.npt, c.pos, sc->d_row_of, sc->dhat,
                                             sc->kappa, sc->d_crow, sc->d_cdat, sc->d_ccoef, sc->d_cfirst,
                                             sc->d_dist, sc->d_ikey, sc->d_ival);
  HIP_TRY(hipGetLastError());
  TRY_RC(sort_pairs(sc, sc->d_ikey, sc->d_ikey_s, sc->d_ival, sc->d_ival_s, 4 * n, s));
  k_row_ptr<<<blocks_for(c.R + 1), 256, 0, s>>>(sc->d_ikey_s, 4u * (uint32_t)n, c.R, sc->d_iptr);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

static bool has_terms(const Context& c) {
  const SelfContact* sc = c.self_contact;
  return c.rules.self_contact && sc && sc->terms && sc->eval.npt + sc->eval.nee > 0;
}

int self_contact_assemble(Context& c) {
  if (!has_terms(c)) return DRAPE_OK;
  SelfContact* sc = c.self_contact;
  k_contact_assemble<<<blocks_for(c.R), 256, 0, c.stream>>>(sc->d_iptr, sc->d_ival_s, sc->d_cdat, sc->d_ccoef,
                                                             sc->d_cfirst, c.diag_slot, c.R, c.sell_val, c.b);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

void self_contact_view(const Context& c, const uint32_t** cptr, const uint32_t** cinc, const int4** crow,
                       const float4** cdat, const float4** ccoef) {
  if (!has_terms(c)) return;
  const SelfContact* sc = c.self_contact;
  *cptr = sc->d_iptr;
  *cinc = sc->d_ival_s;
  *crow = sc->d_crow;
  *cdat = sc->d_cdat;
  *ccoef = sc->d_ccoef;
}

void self_contact_moved(Context& c) {
  if (c.self_contact) c.self_contact->step_ready = false;
}

static int step_set(Context& c) {
  SelfContact* sc = c.self_contact;
  if (!sc->step_ready) {
    bool done = false;
    TRY_RC(swept_phase(c, sc->step, &done));
    if (!done) TRY_RC(broad_phase(c, c.dir, sc->step));
    TRY_RC(ensure_parts(sc, sc->step.npt + sc->step.nee));
    sc->step_ready = true;
  }
  return DRAPE_OK;
}

int self_contact_step_bound(Context& c, float* alpha, int32_t* row) {
  *alpha = 1.0f;
  *row = -1;
  if (!c.rules.self_contact) return DRAPE_OK;
  TRY_RC(setup(c));
  SelfContact* sc = c.self_contact;
  hipStream_t s = c.stream;
  sc->step_ready = false;  // a new direction: a new line-search set
  TRY_RC(step_set(c));
  const int n = sc->step.npt + sc->step.nee;
  if (n == 0) return DRAPE_OK;
  const int groups = blocks_for(n);
  k_contact_ccd<<<groups, 256, 0, s>>>(sc->step.d_cand, n, sc->step.npt, c.pos, c.dir, sc->d_row_of, sc->d_apart,
                                       sc->d_aidx);
  k_min_final<<<1, 256, 0, s>>>(sc->d_apart, sc->d_aidx, groups, sc->step.d_cand, sc->d_row_of, sc->d_eout);
  HIP_TRY(hipGetLastError());
  double out[2];
  HIP_TRY(hipMemcpyAsync(out, sc->d_eout, sizeof(out), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c.read_bytes += sizeof(out);
  *alpha = (float)out[0];
  *row = (int32_t)out[1];
  return DRAPE_OK;
}

int self_contact_energy_change(Context& c, float alpha, double* sum, double* abs_sum) {
  *sum = 0.0;
  *abs_sum = 0.0;
  if (!c.rules.self_contact) return DRAPE_OK;
  TRY_RC(setup(c));
  SelfContact* sc = c.self_contact;
  hipStream_t s = c.stream;
  TRY_RC(step_set(c));
  const int n = sc->step.npt + sc->step.nee;
  if (n == 0) return DRAPE_OK;
  const int groups = blocks_for(n);
  k_contact_energy<<<groups, 256, 0, s>>>(sc->step.d_cand, n, sc->step.npt, c.pos, c.dir, sc->d_row_of, alpha,
                                          sc->dhat, sc->kappa, sc->d_epart);
  k_sum_pairs64<<<1, 256, 0, s>>>(sc->d_epart, groups, sc->d_eout);
  HIP_TRY(hipGetLastError());
  double out[2];
  HIP_TRY(hipMemcpyAsync(out, sc->d_eout, sizeof(out), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c.read_bytes += sizeof(out);
  *sum = out[0];
  *abs_sum = out[1];
  return DRAPE_OK;
}

int self_contact_pairs(Context& c, int32_t* n_pt, int32_t* n_ee, int32_t* vertices, double* distances) {
  const SelfContact* sc = c.self_contact;
  const bool have = c.rules.self_contact && sc && sc->terms;
  const int npt = have ? sc->eval.npt : 0, nee = have ? sc->eval.nee : 0;
  if (n_pt) *n_pt = npt;
  if (n_ee) *n_ee = nee;
  const int n = npt + nee;
  if (n == 0) return DRAPE_OK;
  if (vertices)
    HIP_TRY(hipMemcpyAsync(vertices, sc->eval.d_cand, (size_t)n * sizeof(int4), hipMemcpyDeviceToHost, c.stream));
  if (distances)
    HIP_TRY(hipMemcpyAsync(distances, sc->d_dist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

}  // namespace drape
