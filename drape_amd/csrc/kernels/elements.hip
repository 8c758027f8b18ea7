// Element, row and assembly kernels of the Newton step, plus the line-search and CCD kernels.
//
// Element math follows the reference's element definitions (membrane.rs:15-140 with
// membrane_element.wgsl:20-103; hinge.rs:13-106 with hinge_element.wgsl:17-65; drape.rs:304-322 and
// contact.rs:14-150 with inertia.wgsl:16-52) in FP32, one thread per element; outputs go to the
// reference's local layout so that the kernel-check gates compare entry by entry. The barrier terms are
// evaluated in FP64 (log(d/dhat) and 1/d lose FP32 accuracy as d -> 0); the line-search energy sums are
// FP64 fixed-order reductions.
#include <cstring>

#include "../internal.h"
#include "barrier.hpp"
#include "reduce.hpp"

namespace drape {

__device__ inline float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 xyz(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ inline float comp(float3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// ---- membrane (CST, orthotropic StVK) -------------------------------------------------------
// Projects the 2x2 stress onto its positive part (membrane.rs:116-140).
__device__ inline void clamp_negative_stress(float& a, float& c, float& b) {
  const float mean = 0.5f * (a + c);
  const float radius = sqrtf(0.25f * (a - c) * (a - c) + b * b);
  const float high = mean + radius, low = mean - radius;
  if (low >= 0.0f) return;
  if (high <= 0.0f) { a = c = b = 0.0f; return; }
  float f0 = b, f1 = high - a, s0 = high - c, s1 = b;
  const bool first = hypotf(f0, f1) >= hypotf(s0, s1);
  float p0 = first ? f0 : s0, p1 = first ? f1 : s1;
  const float length = hypotf(p0, p1);
  float v0, v1;
  if (length == 0.0f) { v0 = a >= c ? 1.0f : 0.0f; v1 = a >= c ? 0.0f : 1.0f; }
  else { v0 = p0 / length; v1 = p1 / length; }
  a = high * v0 * v0;
  c = high * v1 * v1;
  b = high * v0 * v1;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2: workgroup w runs on XCD w % 8.
// This bijection of [0, groups) gives each XCD one contiguous run of logical workgroups, so rows that share
// elements (a vertex and its neighbours one grid row away) gather their local blocks through the same L2, and
// the elements of one run of grid rows write their blocks through one L2.
constexpr int XCDS = 8;
__device__ inline int xcd_group(int groups) {
  const int w = blockIdx.x, x = w % XCDS, k = w / XCDS;
  const int q = groups / XCDS, r = groups % XCDS;
  return x * q + (x < r ? x : r) + k;
}

// The 9 floats of element-local Hessian block `record` (9e + 3k + l, 9T + 16h + 4k + l or 9T + 16H + row):
// at hess + 9 record in the reference's layout (rec_pos null), else at its slot in the block-sorted order
// (rec_pos[record]; nothing for a block that belongs to no BSR block).
__device__ inline float* block_out(float* hess, const uint32_t* __restrict__ rec_pos, size_t record) {
  if (!rec_pos) return hess + 9 * record;
  const uint32_t slot = rec_pos[record];
  return slot == NO_SLOT ? nullptr : hess + 9 * (size_t)slot;
}

__global__ void __launch_bounds__(256) k_membrane(const int4* __restrict__ tri, const float4* __restrict__ geom0,
                                                  const float4* __restrict__ geom1, const float4* __restrict__ pos,
                                                  const float* __restrict__ mat, int T, int exact,
                                                  float* __restrict__ grad, float* __restrict__ hess,
                                                  const uint32_t* __restrict__ rec_pos) {
  const int e = xcd_group(gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= T) return;
  const int4 v = tri[e];
  const float4 ga = geom0[e], gb = geom1[e];
  const float gx[3] = {ga.x, ga.z, gb.x}, gy[3] = {ga.y, ga.w, gb.y};
  const float area = gb.z;
  const float wale = mat[0], course = mat[1], coupling = mat[2], shear = mat[3];
  const float3 x0 = xyz(pos[v.x]), x1 = xyz(pos[v.y]), x2 = xyz(pos[v.z]);
  const float3 e1 = x1 - x0, e2 = x2 - x0;
  const float3 fw = e1 * gx[1] + e2 * gx[2];
  const float3 fc = e1 * gy[1] + e2 * gy[2];
  const float eww = 0.5f * (dot3(fw, fw) - 1.0f), ecc = 0.5f * (dot3(fc, fc) - 1.0f), ewc = 0.5f * dot3(fw, fc);
  const float sww = wale * eww + coupling * ecc, scc = course * ecc + coupling * eww, swc = 2.0f * shear * ewc;
  const float3 pw = fw * sww + fc * swc, pc = fw * swc + fc * scc;
  float* g = grad + 9 * (size_t)e;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float3 gk = (pw * gx[k] + pc * gy[k]) * area;
    g[3 * k] = gk.x; g[3 * k + 1] = gk.y; g[3 * k + 2] = gk.z;
  }
  float a = sww, c = scc, b = swc;
  if (!exact) clamp_negative_stress(a, c, b);
  float3 w[3], cc[3], sv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    w[k] = fw * gx[k];
    cc[k] = fc * gy[k];
    sv[k] = (fw * gy[k] + fc * gx[k]) * 0.5f;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const float geo = gx[k] * (a * gx[l] + b * gy[l]) + gy[k] * (b * gx[l] + c * gy[l]);
      float* blk = block_out(hess, rec_pos, 9 * (size_t)e + 3 * k + l);
      if (!blk) continue;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          float m = r == q ? geo : 0.0f;
          m += comp(w[k], r) * comp(w[l], q) * wale;
          m += comp(cc[k], r) * comp(cc[l], q) * course;
          m += comp(w[k], r) * comp(cc[l], q) * coupling;
          m += comp(cc[k], r) * comp(w[l], q) * coupling;
          m += comp(sv[k], r) * comp(sv[l], q) * (4.0f * shear);
          blk[3 * r + q] = m * area;
        }
    }
}

// ---- hinge (dihedral angle) -----------------------------------------------------------------
struct D {
  float v, t;
};
__device__ inline D operator+(D a, D b) { return {a.v + b.v, a.t + b.t}; }
__device__ inline D operator-(D a, D b) { return {a.v - b.v, a.t - b.t}; }
__device__ inline D operator*(D a, D b) { return {a.v * b.v, a.t * b.v + a.v * b.t}; }
__device__ inline D operator/(D a, D b) { return {a.v / b.v, (a.t * b.v - a.v * b.t) / (b.v * b.v)}; }
__device__ inline D operator-(D a) { return {-a.v, -a.t}; }
__device__ inline D dsqrt(D a) {
  const float s = sqrtf(a.v);
  return {s, a.t / (2.0f * s)};
}
__device__ inline float dsqrt(float a) { return sqrtf(a); }
template <class T> __device__ inline T cst(float v);
template <> __device__ inline float cst<float>(float v) { return v; }
template <> __device__ inline D cst<D>(float v) { return {v, 0.0f}; }

template <class T> struct V3 {
  T x, y, z;
};
template <class T> __device__ inline V3<T> vsub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ inline V3<T> vadd(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ inline V3<T> vscale(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ inline T vdot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ inline V3<T> vcross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// d(theta)/dx_k for the four hinge vertices (Bridson et al. 2003 bending modes; hinge.rs:94-106),
// written once for floats and forward-mode duals.
template <class T> __device__ inline void angle_gradient(const V3<T> p[4], V3<T> out[4]) {
  const V3<T> x1 = p[0], x2 = p[1], x3 = p[2], x4 = p[3];
  const V3<T> edge = vsub(x4, x3);
  const T length = dsqrt(vdot(edge, edge));
  const V3<T> n1 = vcross(vsub(x1, x3), vsub(x1, x4));
  const V3<T> n2 = vcross(vsub(x2, x4), vsub(x2, x3));
  const V3<T> a1 = vscale(n1, cst<T>(1.0f) / vdot(n1, n1));
  const V3<T> a2 = vscale(n2, cst<T>(1.0f) / vdot(n2, n2));
  const T m1 = cst<T>(-1.0f);
  out[0] = vscale(vscale(a1, length), m1);
  out[1] = vscale(vscale(a2, length), m1);
  out[2] = vscale(vadd(vscale(a1, vdot(vsub(x1, x4), edge) / length), vscale(a2, vdot(vsub(x2, x4), edge) / length)), m1);
  out[3] = vscale(vadd(vscale(a1, -vdot(vsub(x1, x3), edge) / length), vscale(a2, -vdot(vsub(x2, x3), edge) / length)), m1);
}

__device__ inline float hinge_angle(float3 x1, float3 x2, float3 x3, float3 x4) {
  const float3 edge = x4 - x3;
  const float3 n1 = cross3(x1 - x3, x1 - x4);
  const float3 n2 = cross3(x2 - x4, x2 - x3);
  const float3 u1 = n1 * (1.0f / sqrtf(dot3(n1, n1))), u2 = n2 * (1.0f / sqrtf(dot3(n2, n2)));
  const float sine = dot3(cross3(u1, u2), edge * (1.0f / sqrtf(dot3(edge, edge))));
  return atan2f(sine, dot3(u1, u2));
}

// One wavefront of 64 hinges per workgroup. Each lane forms its hinge's 12 gradient and 144 Hessian floats
// in LDS (padded rows, so lanes hit distinct banks); the wavefront then writes its 64 consecutive records
// (9,216 floats, contiguous in the element-local layout) with coalesced stores.
constexpr int HINGE_WG = 64;
constexpr int HINGE_HSTRIDE = 145;  // 144 floats + 1: lane * 145 covers every bank
constexpr int HINGE_GSTRIDE = 13;   // 12 floats + 1

__global__ void __launch_bounds__(HINGE_WG) k_hinge(const int4* __restrict__ hin, const float* __restrict__ stiff,
                                                    const float4* __restrict__ pos, int H, int exact,
                                                    float* __restrict__ grad, float* __restrict__ hess,
                                                    const uint32_t* __restrict__ rec_pos, size_t first_record) {
  __shared__ float stage_h[HINGE_WG * HINGE_HSTRIDE];
  __shared__ float stage_g[HINGE_WG * HINGE_GSTRIDE];
  const int h0 = xcd_group(gridDim.x) * HINGE_WG;
  const int h = h0 + threadIdx.x;
  if (h < H) {
    const int4 v = hin[h];
    const float3 x[4] = {xyz(pos[v.x]), xyz(pos[v.y]), xyz(pos[v.z]), xyz(pos[v.w])};
    const float k = stiff[h];
    const float theta = hinge_angle(x[0], x[1], x[2], x[3]);
    V3<float> pf[4], gf[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) pf[i] = {x[i].x, x[i].y, x[i].z};
    angle_gradient<float>(pf, gf);
    const float factor = 2.0f * k;
    const float gfac = factor * theta;
    float* g = stage_g + threadIdx.x * HINGE_GSTRIDE;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g[3 * i] = gf[i].x * gfac; g[3 * i + 1] = gf[i].y * gfac; g[3 * i + 2] = gf[i].z * gfac;
    }
    float* out = stage_h + threadIdx.x * HINGE_HSTRIDE;
    // column (l, c) of the Hessian: Gauss-Newton part factor*g_k g_l^T, plus factor*theta*d(grad theta)/dx_lc
    for (int l = 0; l < 4; ++l)
      for (int c = 0; c < 3; ++c) {
        const float glc = c == 0 ? gf[l].x : (c == 1 ? gf[l].y : gf[l].z);
        float col[12];
        if (exact) {
          V3<D> pd[4], gd[4];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            pd[i] = {{pf[i].x, (i == l && c == 0) ? 1.0f : 0.0f}, {pf[i].y, (i == l && c == 1) ? 1.0f : 0.0f},
                     {pf[i].z, (i == l && c == 2) ? 1.0f : 0.0f}};
          angle_gradient<D>(pd, gd);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            col[3 * i] = gd[i].x.t; col[3 * i + 1] = gd[i].y.t; col[3 * i + 2] = gd[i].z.t;
          }
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float gkr = r == 0 ? gf[kk].x : (r == 1 ? gf[kk].y : gf[kk].z);
            const float value = exact ? factor * (gkr * glc + theta * col[3 * kk + r]) : gkr * glc * factor;
            out[9 * (4 * kk + l) + 3 * r + c] = value;
          }
      }
  }
  __syncthreads();
  const int count = H - h0 < HINGE_WG ? H - h0 : HINGE_WG;
  if (rec_pos) {
    // block-sorted: lane j writes the 9 floats of records j, j + 64, ... (16 per hinge) at their slots (writing
    // float t from lane t % 64, seven blocks per store instruction, was slower: 37 against 26 us at n129)
    for (int i = threadIdx.x; i < 16 * count; i += HINGE_WG) {
      const uint32_t slot = rec_pos[first_record + 16 * (size_t)h0 + i];
      if (slot == NO_SLOT) continue;
      const float* src = stage_h + (i / 16) * HINGE_HSTRIDE + 9 * (i % 16);
      float* dst = hess + 9 * (size_t)slot;
#pragma unroll
      for (int e = 0; e < 9; ++e) dst[e] = src[e];
    }
  } else {
    float* hout = hess + 144 * (size_t)h0;
    for (int i = threadIdx.x; i < 144 * count; i += HINGE_WG) hout[i] = stage_h[(i / 144) * HINGE_HSTRIDE + i % 144];
  }
  float* gout = grad + 12 * (size_t)h0;
  for (int i = threadIdx.x; i < 12 * count; i += HINGE_WG) gout[i] = stage_g[(i / 12) * HINGE_GSTRIDE + i % 12];
}

// ---- rows: inertia and vertex-collider barrier ----------------------------------------------

struct Proximity {
  double distance;
  double n[3];
  double curvature;
};
__device__ inline Proximity proximity(const float* col, int k, const double x[3]) {
  const float* c = col + 4 + 8 * k;
  Proximity p;
  if (c[0] == 1.0f) {
    p.n[0] = c[1]; p.n[1] = c[2]; p.n[2] = c[3];
    p.distance = p.n[0] * x[0] + p.n[1] * x[1] + p.n[2] * x[2] - (double)c[4];
    p.curvature = 0.0;
  } else {
    const double q[3] = {x[0] - (double)c[1], x[1] - (double)c[2], x[2] - (double)c[3]};
    const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    p.n[0] = q[0] / len; p.n[1] = q[1] / len; p.n[2] = q[2] / len;
    p.distance = len - (double)c[4];
    p.curvature = 1.0 / len;
  }
  return p;
}

__global__ void __launch_bounds__(256) k_rows(const float4* __restrict__ pos, const float4* __restrict__ xhat,
                                              const uint32_t* __restrict__ rvert, const float* __restrict__ rmass,
                                              const float* __restrict__ mu, const float* __restrict__ colliders,
                                              int R, float* __restrict__ grad, float* __restrict__ hess,
                                              const uint32_t* __restrict__ rec_pos, size_t first_record) {
  const int row = xcd_group(gridDim.x) * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const uint32_t v = rvert[row];
  const float m = rmass[row];
  const float4 xv = pos[v], xh = xhat[v];
  float g[3] = {m * (xv.x - xh.x), m * (xv.y - xh.y), m * (xv.z - xh.z)};
  float h[9];
  const float diag = m * (1.0f + mu[0]);
#pragma unroll
  for (int i = 0; i < 9; ++i) h[i] = (i % 4 == 0) ? diag : 0.0f;
  const int K = colliders ? (int)colliders[0] : 0;
  if (K > 0) {
    const double dhat = colliders[1], kappa = colliders[2];
    const double keep_curvature = 1.0 - (double)colliders[3];
    const double x[3] = {xv.x, xv.y, xv.z};
    double gd[3] = {0, 0, 0}, hd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < K; ++k) {
      const Proximity p = proximity(colliders, k, x);
      if (p.distance >= dhat) continue;
      const BarrierTerms b = barrier_terms(p.distance, dhat, kappa);
      for (int r = 0; r < 3; ++r) {
        gd[r] += b.first * p.n[r];
        for (int c = 0; c < 3; ++c) {
          const double nn = p.n[r] * p.n[c];
          hd[3 * r + c] += b.second * nn + b.first * p.curvature * keep_curvature * ((r == c ? 1.0 : 0.0) - nn);
        }
      }
    }
    for (int r = 0; r < 3; ++r) g[r] = (float)((double)g[r] + gd[r]);
    for (int i = 0; i < 9; ++i) h[i] = (float)((double)h[i] + hd[i]);
  }
  float* go = grad + 3 * (size_t)row;
  go[0] = g[0]; go[1] = g[1]; go[2] = g[2];
  float* ho = rec_pos ? block_out(hess, rec_pos, first_record + row) : hess + 9 * (size_t)row;
  if (!ho) return;
#pragma unroll
  for (int i = 0; i < 9; ++i) ho[i] = h[i];
}


// ---- assembly (assemble_matrix.wgsl / assemble_force.wgsl, one thread per block or row) -----
// vals of CSR block b = sum of its local blocks in list order (mass/contact, membranes, hinges),
// written into the sliced layout: entry e of block b at sell_val[block_dst[b] + 64 e].
__global__ void __launch_bounds__(256) k_assemble_blocks(const uint32_t* __restrict__ bc_ptr,
                                                         const uint32_t* __restrict__ bc_idx,
                                                         const float* __restrict__ lhess,
                                                         const uint32_t* __restrict__ block_dst, int B,
                                                         float* __restrict__ sell_val) {
  const int b = xcd_group(gridDim.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t end = bc_ptr[b + 1];
  for (uint32_t j = bc_ptr[b]; j < end; ++j) {
    const float* s = lhess + 9 * (size_t)bc_idx[j];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = acc[e] + s[e];
  }
  float* d = sell_val + block_dst[b];
#pragma unroll
  for (int e = 0; e < 9; ++e) d[64 * e] = acc[e];
}

// The same sums from the block-sorted records: block b's row, membrane and hinge blocks are three contiguous runs
// of their regions (kind_ptr), summed in that order, which is the list order, so the sum and its order are those
// of k_assemble_blocks.
__device__ inline void add_run(float* acc, const float* __restrict__ sorted, uint32_t first, uint32_t last) {
  for (uint32_t j = first; j < last; ++j) {
    const float* s = sorted + 9 * (size_t)j;
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = acc[e] + s[e];
  }
}

__global__ void __launch_bounds__(256) k_assemble_sorted(const uint32_t* __restrict__ kind_ptr,
                                                         const float* __restrict__ sorted, uint32_t hinge0,
                                                         uint32_t row0, const uint32_t* __restrict__ block_dst, int B,
                                                         float* __restrict__ sell_val) {
  const int b = xcd_group(gridDim.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t* k0 = kind_ptr + 3 * (size_t)b;
  add_run(acc, sorted, row0 + k0[0], row0 + k0[3]);
  add_run(acc, sorted, k0[1], k0[4]);
  add_run(acc, sorted, hinge0 + k0[2], hinge0 + k0[5]);
  float* d = sell_val + block_dst[b];
#pragma unroll
  for (int e = 0; e < 9; ++e) d[64 * e] = acc[e];
}

__global__ void __launch_bounds__(256) k_assemble_rhs(const uint32_t* __restrict__ fc_ptr,
                                                      const uint32_t* __restrict__ fc_idx,
                                                      const float* __restrict__ lgrad, int R, float4* __restrict__ b) {
  const int row = xcd_group(gridDim.x) * blockDim.x + threadIdx.x;
  if (row >= R) return;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  const uint32_t end = fc_ptr[row + 1];
  for (uint32_t j = fc_ptr[row]; j < end; ++j) {
    const float* s = lgrad + 3 * (size_t)fc_idx[j];
    a0 = a0 + s[0]; a1 = a1 + s[1]; a2 = a2 + s[2];
  }
  b[row] = make_float4(-a0, -a1, -a2, 0.0f);
}

// ---- FP64 sums over rows ---------------------------------------------------------------------
// mode 0: sum b.b, mode 1: sum b.x. Per-workgroup partials, then one workgroup in fixed order.
__global__ void __launch_bounds__(256) k_row_sums(const float4* __restrict__ a, const float4* __restrict__ c, int R,
                                                  double* __restrict__ part) {
  __shared__ double2 lds[256];
  const int row = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (row < R) {
    const float4 u = a[row], w = c[row];
    s = (double)u.x * w.x + (double)u.y * w.y + (double)u.z * w.z;
  }
  const double2 t = tree256d2(s, 0.0, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t.x;
}

__global__ void __launch_bounds__(256) k_final_sums(const double* __restrict__ part, int count, int stride,
                                                    double* __restrict__ out) {
  __shared__ double2 lds[256];
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < count; j += 256) {
    a += part[stride * j];
    if (stride == 2) b += part[2 * j + 1];
  }
  const double2 t = tree256d2(a, b, lds);
  if (threadIdx.x == 0) {
    out[0] = t.x;
    out[1] = t.y;
  }
}

// ---- line search: energy change W(x + alpha d) - W(x) --------------------------------------
// membrane_energy_change.wgsl / hinge_energy_change.wgsl / inertia_energy_change.wgsl in one launch:
// workgroups [0, gt) membranes, [gt, gt + gh) hinges, then rows. FP64 partial sums of (d, |d|).
__global__ void __launch_bounds__(256) k_energy_change(const int4* __restrict__ tri, const float4* __restrict__ geom0,
                                                       const float4* __restrict__ geom1, const int4* __restrict__ hin,
                                                       const float* __restrict__ stiff, const float4* __restrict__ pos,
                                                       const float4* __restrict__ xhat, const float4* __restrict__ dir,
                                                       const uint32_t* __restrict__ rvert,
                                                       const float* __restrict__ rmass, const float* __restrict__ mat,
                                                       const float* __restrict__ colliders, float alpha, int T, int H,
                                                       int R, int gt, int gh, double* __restrict__ part) {
  __shared__ double2 lds[256];
  const int wg = blockIdx.x;
  double delta = 0.0;
  if (wg < gt) {
    const int e = wg * 256 + threadIdx.x;
    if (e < T) {
      const int4 v = tri[e];
      const float4 ga = geom0[e], gb = geom1[e];
      const float3 x0 = xyz(pos[v.x]), x1 = xyz(pos[v.y]), x2 = xyz(pos[v.z]);
      const float3 d0 = xyz(dir[v.x]), d1 = xyz(dir[v.y]), d2 = xyz(dir[v.z]);
      const float3 fw = (x1 - x0) * ga.z + (x2 - x0) * gb.x, fc = (x1 - x0) * ga.w + (x2 - x0) * gb.y;
      const float3 dw = ((d1 - d0) * ga.z + (d2 - d0) * gb.x) * alpha, dc = ((d1 - d0) * ga.w + (d2 - d0) * gb.y) * alpha;
      const float c0 = dot3(fw, dw) + 0.5f * dot3(dw, dw), c1 = dot3(fc, dc) + 0.5f * dot3(dc, dc);
      const float c2 = 0.5f * (dot3(fw, dc) + dot3(dw, fc) + dot3(dw, dc));
      const float eww = 0.5f * (dot3(fw, fw) - 1.0f), ecc = 0.5f * (dot3(fc, fc) - 1.0f), ewc = 0.5f * dot3(fw, fc);
      const float wale = mat[0], course = mat[1], coupling = mat[2], shear = mat[3];
      const float sww = wale * eww + coupling * ecc, scc = course * ecc + coupling * eww, swc = 2.0f * shear * ewc;
      const float first = sww * c0 + scc * c1 + 2.0f * swc * c2;
      const float w = 0.5f * wale * c0 * c0 + 0.5f * course * c1 * c1 + coupling * c0 * c1 + 2.0f * shear * c2 * c2;
      delta = (double)(gb.z * (first + w));
    }
  } else if (wg < gt + gh) {
    const int h = (wg - gt) * 256 + threadIdx.x;
    if (h < H) {
      const int4 v = hin[h];
      const float3 p1 = xyz(pos[v.x]), p2 = xyz(pos[v.y]), p3 = xyz(pos[v.z]), p4 = xyz(pos[v.w]);
      const float before = hinge_angle(p1, p2, p3, p4);
      const float after = hinge_angle(p1 + xyz(dir[v.x]) * alpha, p2 + xyz(dir[v.y]) * alpha,
                                      p3 + xyz(dir[v.z]) * alpha, p4 + xyz(dir[v.w]) * alpha);
      delta = (double)(stiff[h] * (after - before) * (after + before));
    }
  } else {
    const int row = (wg - gt - gh) * 256 + threadIdx.x;
    if (row < R) {
      const uint32_t v = rvert[row];
      const float3 x = xyz(pos[v]);
      const float3 off = x - xyz(xhat[v]);
      const float3 moved = xyz(dir[v]) * alpha;
      const float inertia = rmass[row] * (dot3(off, moved) + 0.5f * dot3(moved, moved));
      double change = 0.0;
      const int K = colliders ? (int)colliders[0] : 0;
      if (K > 0) {
        const double dhat = colliders[1], kappa = colliders[2];
        const double xa[3] = {x.x, x.y, x.z};
        const double xb[3] = {x.x + moved.x, x.y + moved.y, x.z + moved.z};
        for (int k = 0; k < K; ++k) {
          const double before = barrier_terms(proximity(colliders, k, xa).distance, dhat, kappa).value;
          const double after = barrier_terms(proximity(colliders, k, xb).distance, dhat, kappa).value;
          change += after - before;
        }
      }
      delta = (double)inertia + change;
    }
  }
  const double2 t = tree256d2(delta, fabs(delta), lds);
  if (threadIdx.x == 0) {
    part[2 * wg] = t.x;
    part[2 * wg + 1] = t.y;
  }
}

// ---- CCD bound and minimum distance (ccd.wgsl + argmin.wgsl) --------------------------------
__device__ inline float sphere_step(float3 q, float3 p, float radius, float d0) {
  const float target = radius + 0.1f * d0;
  const float a = dot3(p, p), b = 2.0f * dot3(q, p);
  const float c = (1.0f - 0.1f) * d0 * (sqrtf(dot3(q, q)) + target);
  const float disc = b * b - 4.0f * a * c;
  if (b >= 0.0f || disc < 0.0f) return 1.0f;
  return fminf(1.0f, 2.0f * c / (-b + sqrtf(disc)));
}

__global__ void __launch_bounds__(256) k_ccd(const float4* __restrict__ pos, const float4* __restrict__ dir,
                                             const uint32_t* __restrict__ rvert, const float* __restrict__ colliders,
                                             int R, float* __restrict__ fpart, int32_t* __restrict__ ipart) {
  __shared__ float sa[256], sd[256];
  __shared__ int ia[256], id[256];
  const int row = blockIdx.x * 256 + threadIdx.x;
  float alpha = 1.0f, nearest = 3.0e38f;
  if (row < R) {
    const uint32_t v = rvert[row];
    const float3 x = xyz(pos[v]), p = xyz(dir[v]);
    const int K = (int)colliders[0];
    for (int k = 0; k < K; ++k) {
      const float* c = colliders + 4 + 8 * k;
      const float3 a = make_float3(c[1], c[2], c[3]);
      float d0, step;
      if (c[0] == 1.0f) {
        d0 = dot3(a, x) - c[4];
        const float rate = dot3(a, p);
        step = rate >= 0.0f ? 1.0f : fminf(1.0f, (1.0f - 0.1f) * d0 / -rate);
      } else {
        const float3 q = x - a;
        d0 = sqrtf(dot3(q, q)) - c[4];
        step = sphere_step(q, p, c[4], d0);
      }
      alpha = fminf(alpha, step);
      nearest = fminf(nearest, d0);
    }
  }
  sa[threadIdx.x] = alpha; sd[threadIdx.x] = nearest;
  ia[threadIdx.x] = row; id[threadIdx.x] = row;
  __syncthreads();
  // arg-min, ties to the lower row (argmin.wgsl:34-61)
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int t = threadIdx.x, u = t + s;
      if (sa[u] < sa[t] || (sa[u] == sa[t] && ia[u] < ia[t])) { sa[t] = sa[u]; ia[t] = ia[u]; }
      if (sd[u] < sd[t] || (sd[u] == sd[t] && id[u] < id[t])) { sd[t] = sd[u]; id[t] = id[u]; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    fpart[2 * blockIdx.x] = sa[0]; fpart[2 * blockIdx.x + 1] = sd[0];
    ipart[2 * blockIdx.x] = ia[0]; ipart[2 * blockIdx.x + 1] = id[0];
  }
}

__global__ void __launch_bounds__(256) k_ccd_final(const float* __restrict__ fpart, const int32_t* __restrict__ ipart,
                                                   int count, double* __restrict__ out) {
  __shared__ float sa[256], sd[256];
  __shared__ int ia[256], id[256];
  float a = 3.0e38f, d = 3.0e38f;
  int ai = 0x7fffffff, di = 0x7fffffff;
  for (int j = threadIdx.x; j < count; j += 256) {
    const float fa = fpart[2 * j], fd = fpart[2 * j + 1];
    const int ja = ipart[2 * j], jd = ipart[2 * j + 1];
    if (fa < a || (fa == a && ja < ai)) { a = fa; ai = ja; }
    if (fd < d || (fd == d && jd < di)) { d = fd; di = jd; }
  }
  sa[threadIdx.x] = a; sd[threadIdx.x] = d; ia[threadIdx.x] = ai; id[threadIdx.x] = di;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int t = threadIdx.x, u = t + s;
      if (sa[u] < sa[t] || (sa[u] == sa[t] && ia[u] < ia[t])) { sa[t] = sa[u]; ia[t] = ia[u]; }
      if (sd[u] < sd[t] || (sd[u] == sd[t] && id[u] < id[t])) { sd[t] = sd[u]; id[t] = id[u]; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sa[0]; out[1] = ia[0]; out[2] = sd[0]; out[3] = id[0];
  }
}

// ---- vertex updates --------------------------------------------------------------------------
__global__ void k_scatter_direction(const float4* __restrict__ rows, const uint32_t* __restrict__ rvert, int R,
                                    float4* __restrict__ dir) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const float4 v = rows[row];
  dir[rvert[row]] = make_float4(v.x, v.y, v.z, 0.0f);
}

// positions[v] += alpha * direction[v] (newton.ts:97-101: computed in double, stored as f32).
__global__ void k_advance(float4* __restrict__ pos, const float4* __restrict__ dir, const uint32_t* __restrict__ rvert,
                          int R, float alpha) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const uint32_t v = rvert[row];
  float4 p = pos[v];
  const float4 d = dir[v];
  p.x = (float)((double)p.x + (double)alpha * (double)d.x);
  p.y = (float)((double)p.y + (double)alpha * (double)d.y);
  p.z = (float)((double)p.z + (double)alpha * (double)d.z);
  pos[v] = p;
}

// Per-workgroup maxima over rows (order-independent): out[4g..] = max|a_row| (a = PCG solution),
// max|pos - org| and max|pos| over the row vertices, and the count of non-finite position components.
__global__ void __launch_bounds__(256) k_row_maxes(const float4* __restrict__ a, const float4* __restrict__ pos,
                                                   const float4* __restrict__ org, const uint32_t* __restrict__ rvert,
                                                   int R, float* __restrict__ part) {
  __shared__ float l0[256], l1[256], l2[256], l3[256];
  const int row = blockIdx.x * 256 + threadIdx.x;
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, bad = 0.0f;
  if (row < R) {
    const float4 v = a[row];
    m0 = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
    const uint32_t vx = rvert[row];
    const float4 p = pos[vx], o = org[vx];
    m1 = fmaxf(fabsf(p.x - o.x), fmaxf(fabsf(p.y - o.y), fabsf(p.z - o.z)));
    m2 = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
    bad = (finite32(p.x) && finite32(p.y) && finite32(p.z)) ? 0.0f : 1.0f;
  }
  l0[threadIdx.x] = m0; l1[threadIdx.x] = m1; l2[threadIdx.x] = m2; l3[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int t = threadIdx.x;
      l0[t] = fmaxf(l0[t], l0[t + s]);
      l1[t] = fmaxf(l1[t], l1[t + s]);
      l2[t] = fmaxf(l2[t], l2[t + s]);
      l3[t] = l3[t] + l3[t + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = l0[0]; part[4 * blockIdx.x + 1] = l1[0];
    part[4 * blockIdx.x + 2] = l2[0]; part[4 * blockIdx.x + 3] = l3[0];
  }
}

__global__ void __launch_bounds__(256) k_row_maxes_final(const float* __restrict__ part, int count,
                                                         double* __restrict__ out) {
  __shared__ float l0[256], l1[256], l2[256], l3[256];
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, bad = 0.0f;
  for (int j = threadIdx.x; j < count; j += 256) {
    m0 = fmaxf(m0, part[4 * j]); m1 = fmaxf(m1, part[4 * j + 1]); m2 = fmaxf(m2, part[4 * j + 2]);
    bad = bad + part[4 * j + 3];
  }
  l0[threadIdx.x] = m0; l1[threadIdx.x] = m1; l2[threadIdx.x] = m2; l3[threadIdx.x] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int t = threadIdx.x;
      l0[t] = fmaxf(l0[t], l0[t + s]);
      l1[t] = fmaxf(l1[t], l1[t + s]);
      l2[t] = fmaxf(l2[t], l2[t + s]);
      l3[t] = l3[t] + l3[t + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[4] = l0[0]; out[5] = l1[0]; out[6] = l2[0]; out[7] = l3[0];
  }
}

// dir[v] = x_hat[v] - pos[v] for every vertex (drape-e4.ts:323-330 newtonStart, f32 difference).
__global__ void k_towards(const float4* __restrict__ xhat, const float4* __restrict__ pos, int V, float4* __restrict__ dir) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const float4 a = xhat[v], b = pos[v];
  dir[v] = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, 0.0f);
}

// pos[v] = f32(pos[v] + alpha dir[v]) for every vertex (double arithmetic as the reference host does).
__global__ void k_step_all(float4* __restrict__ pos, const float4* __restrict__ dir, int V, float alpha) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  float4 p = pos[v];
  const float4 d = dir[v];
  p.x = (float)((double)p.x + (double)alpha * (double)d.x);
  p.y = (float)((double)p.y + (double)alpha * (double)d.y);
  p.z = (float)((double)p.z + (double)alpha * (double)d.z);
  pos[v] = p;
}

// x_hat = x + h^2 g on row vertices, row_mass = m/h^2 (drape-e7.ts:56-72, f32 as the reference host).
__global__ void k_ptc_prepare(const float4* __restrict__ pos, float4* __restrict__ xhat, float* __restrict__ rmass,
                              const float* __restrict__ unit_mass, const uint32_t* __restrict__ rvert, int V, int R,
                              float fall, double h2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < V) xhat[i] = pos[i];
  if (i < R) rmass[i] = (float)((double)unit_mass[i] / h2);
}
__global__ void k_ptc_fall(const float4* __restrict__ pos, float4* __restrict__ xhat, const uint32_t* __restrict__ rvert,
                           int R, float fall) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const uint32_t v = rvert[row];
  float4 h = pos[v];
  h.z = (float)((double)h.z + (double)fall);
  xhat[v] = h;
}

// drape.rs:869 prediction (x_hat = x + h v + h^2 g on the free vertices, x elsewhere), in FP64 then rounded.
__global__ void k_dyn_predict(const float4* __restrict__ pos, const float4* __restrict__ vel,
                              const uint32_t* __restrict__ rvert, int V, int R, double h, double g,
                              float4* __restrict__ xhat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < V) xhat[i] = pos[i];
}
__global__ void k_dyn_predict_rows(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                   const uint32_t* __restrict__ rvert, int R, double h, double g,
                                   float4* __restrict__ xhat) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const uint32_t v = rvert[row];
  const float4 x = pos[v], u = vel[v];
  xhat[v] = make_float4((float)((double)x.x + h * u.x), (float)((double)x.y + h * u.y),
                        (float)((double)x.z + h * u.z - h * h * g), 0.0f);
}
// v = (x_new - x_old) / h for every vertex
__global__ void k_dyn_velocity(const float4* __restrict__ pos, const float4* __restrict__ org, int V, double h,
                               float4* __restrict__ vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const float4 a = pos[i], b = org[i];
  vel[i] = make_float4((float)(((double)a.x - b.x) / h), (float)(((double)a.y - b.y) / h),
                       (float)(((double)a.z - b.z) / h), 0.0f);
}

int launch_dyn_predict(Context& c, const float4* vel, double h) {
  k_dyn_predict<<<blocks_for(c.V), 256, 0, c.stream>>>(c.pos, vel, c.rvert, c.V, c.R, h, c.problem.gravity, c.xhat);
  k_dyn_predict_rows<<<blocks_for(c.R), 256, 0, c.stream>>>(c.pos, vel, c.rvert, c.R, h, c.problem.gravity, c.xhat);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int launch_dyn_velocity(Context& c, float4* vel, double h) {
  k_dyn_velocity<<<blocks_for(c.V), 256, 0, c.stream>>>(c.pos, c.org, c.V, h, vel);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

__global__ void k_set_float(float* dst, float v) {
  if (threadIdx.x == 0) dst[0] = v;
}

// ---- launchers --------------------------------------------------------------------------------
int launch_set_float(Context& c, float* dst, float v) {
  k_set_float<<<1, 64, 0, c.stream>>>(dst, v);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int launch_evaluate(Context& c, bool exact, bool reference) {
  hipStream_t s = c.stream;
  const bool sorted = !reference && c.rec_pos;
  const uint32_t* rp = sorted ? c.rec_pos : nullptr;
  // the reference layout places each kernel's blocks after the previous kernel's; the sorted one has slots
  float* hh = sorted ? c.lhess : c.lhess + 81 * (size_t)c.T;
  float* hr = sorted ? c.lhess : c.lhess + 81 * (size_t)c.T + 144 * (size_t)c.H;
  if (c.T > 0)
    k_membrane<<<blocks_for(c.T), 256, 0, s>>>(c.tri, c.geom0, c.geom1, c.pos, c.material, c.T, exact, c.lgrad, c.lhess,
                                               rp);
  if (c.H > 0)
    k_hinge<<<blocks_for(c.H, HINGE_WG), HINGE_WG, 0, s>>>(c.hin, c.hstiff, c.pos, c.H, exact, c.lgrad + 9 * (size_t)c.T,
                                                           hh, rp, 9 * (size_t)c.T);
  k_rows<<<blocks_for(c.R), 256, 0, s>>>(c.pos, c.xhat, c.rvert, c.rmass, c.mu, c.K > 0 ? c.colliders : nullptr, c.R,
                                         c.lgrad + 9 * (size_t)c.T + 12 * (size_t)c.H, hr, rp,
                                         9 * (size_t)c.T + 16 * (size_t)c.H);
  HIP_TRY(hipGetLastError());
  c.hess_sorted = sorted;
  return self_contact_evaluate(c);
}

int launch_assemble(Context& c) {
  hipStream_t s = c.stream;
  ++c.system_version;
  TRY_RC(two_level_before_assembly(c));
  if (c.hess_sorted)
    k_assemble_sorted<<<blocks_for(c.B), 256, 0, s>>>(c.kind_ptr, c.lhess, c.hinge0, c.row0, c.block_dst, c.B,
                                                      c.sell_val);
  else
    k_assemble_blocks<<<blocks_for(c.B), 256, 0, s>>>(c.bc_ptr, c.bc_idx, c.lhess, c.block_dst, c.B, c.sell_val);
  k_assemble_rhs<<<blocks_for(c.R), 256, 0, s>>>(c.fc_ptr, c.fc_idx, c.lgrad, c.R, c.b);
  HIP_TRY(hipGetLastError());
  return self_contact_assemble(c);
}

static int row_sum(Context& c, const float4* a, const float4* w, double* out) {
  hipStream_t s = c.stream;
  k_row_sums<<<c.G, 256, 0, s>>>(a, w, c.R, c.dpart);
  k_final_sums<<<1, 256, 0, s>>>(c.dpart, c.G, 1, c.dout);
  HIP_TRY(hipGetLastError());
  TRY_RC(read_back(c, c.dout_host, c.dout, 2 * sizeof(double)));
  *out = c.dout_host[0];
  return DRAPE_OK;
}

int launch_rhs_norm(Context& c, double* norm2) { return row_sum(c, c.b, c.b, norm2); }

int launch_slope(Context& c, double* slope) {
  double v = 0.0;
  int rc = row_sum(c, c.b, c.x, &v);
  *slope = -v;
  return rc;
}

int launch_scatter_direction(Context& c, const float4* rows_src) {
  k_scatter_direction<<<blocks_for(c.R), 256, 0, c.stream>>>(rows_src, c.rvert, c.R, c.dir);
  HIP_TRY(hipGetLastError());
  self_contact_moved(c);
  return DRAPE_OK;
}

int launch_advance(Context& c, float alpha) {
  k_advance<<<blocks_for(c.R), 256, 0, c.stream>>>(c.pos, c.dir, c.rvert, c.R, alpha);
  HIP_TRY(hipGetLastError());
  self_contact_moved(c);
  return DRAPE_OK;
}

int launch_energy_change(Context& c, float alpha, double* sum, double* abs_sum) {
  hipStream_t s = c.stream;
  const int gt = blocks_for(c.T), gh = blocks_for(c.H), gr = blocks_for(c.R);
  k_energy_change<<<gt + gh + gr, 256, 0, s>>>(c.tri, c.geom0, c.geom1, c.hin, c.hstiff, c.pos, c.xhat, c.dir, c.rvert,
                                               c.rmass, c.material, c.K > 0 ? c.colliders : nullptr, alpha, c.T, c.H,
                                               c.R, gt, gh, c.dpart);
  k_final_sums<<<1, 256, 0, s>>>(c.dpart, gt + gh + gr, 2, c.dout);
  HIP_TRY(hipGetLastError());
  TRY_RC(read_back(c, c.dout_host, c.dout, 2 * sizeof(double)));
  *sum = c.dout_host[0];
  *abs_sum = c.dout_host[1];
  if (c.rules.self_contact) {
    double cs = 0.0, ca = 0.0;
    TRY_RC(self_contact_energy_change(c, alpha, &cs, &ca));
    *sum += cs;
    *abs_sum += ca;
  }
  return DRAPE_OK;
}

int launch_max_step(Context& c, float* alpha, int32_t* row, float* min_distance, int32_t* min_row) {
  if (c.K == 0) {
    *alpha = 1.0f; *row = -1; *min_distance = INFINITY; *min_row = -1;
    return DRAPE_OK;
  }
  hipStream_t s = c.stream;
  k_ccd<<<c.G, 256, 0, s>>>(c.pos, c.dir, c.rvert, c.colliders, c.R, c.fpart, c.ipart);
  k_ccd_final<<<1, 256, 0, s>>>(c.fpart, c.ipart, c.G, c.dout);
  HIP_TRY(hipGetLastError());
  TRY_RC(read_back(c, c.dout_host, c.dout, 4 * sizeof(double)));
  *alpha = (float)c.dout_host[0];
  *row = (int32_t)c.dout_host[1];
  *min_distance = (float)c.dout_host[2];
  *min_row = (int32_t)c.dout_host[3];
  if (c.rules.self_contact) {
    float a = 1.0f;
    int32_t r = -1;
    TRY_RC(self_contact_step_bound(c, &a, &r));
    if (a < *alpha) {
      *alpha = a;
      *row = r;
    }
  }
  return DRAPE_OK;
}

int launch_min_distance(Context& c, float* min_distance) {
  *min_distance = INFINITY;
  if (c.K == 0) return DRAPE_OK;
  hipStream_t s = c.stream;
  k_ccd<<<c.G, 256, 0, s>>>(c.pos, c.dir, c.rvert, c.colliders, c.R, c.fpart, c.ipart);
  k_ccd_final<<<1, 256, 0, s>>>(c.fpart, c.ipart, c.G, c.dout);
  HIP_TRY(hipGetLastError());
  TRY_RC(read_back(c, c.dout_host, c.dout, 4 * sizeof(double)));
  *min_distance = (float)c.dout_host[2];
  return DRAPE_OK;
}

int launch_row_maxes(Context& c, const float4* org, float out[4]) {
  hipStream_t s = c.stream;
  k_row_maxes<<<c.G, 256, 0, s>>>(c.x, c.pos, org, c.rvert, c.R, c.fpart);
  k_row_maxes_final<<<1, 256, 0, s>>>(c.fpart, c.G, c.dout);
  HIP_TRY(hipGetLastError());
  TRY_RC(read_back(c, c.dout_host, c.dout, 8 * sizeof(double)));
  for (int i = 0; i < 4; ++i) out[i] = (float)c.dout_host[4 + i];
  return DRAPE_OK;
}

int launch_towards(Context& c) {
  k_towards<<<blocks_for(c.V), 256, 0, c.stream>>>(c.xhat, c.pos, c.V, c.dir);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int launch_step_all(Context& c, float alpha) {
  k_step_all<<<blocks_for(c.V), 256, 0, c.stream>>>(c.pos, c.dir, c.V, alpha);
  HIP_TRY(hipGetLastError());
  self_contact_moved(c);
  return DRAPE_OK;
}

int launch_ptc_prepare(Context& c, const float4* base, double h) {
  note_shift(c, c.shift_mu, h);
  const float h2f = (float)(h * h);
  const float fall = h2f * (float)(-c.problem.gravity);
  const int n = c.V > c.R ? c.V : c.R;
  k_ptc_prepare<<<blocks_for(n), 256, 0, c.stream>>>(base, c.xhat, c.rmass, c.unit_mass_dev, c.rvert, c.V, c.R, fall, h * h);
  k_ptc_fall<<<blocks_for(c.R), 256, 0, c.stream>>>(base, c.xhat, c.rvert, c.R, fall);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

}  // namespace drape
