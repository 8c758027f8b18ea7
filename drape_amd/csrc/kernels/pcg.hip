// Block-Jacobi FP32 PCG with residual replacement on a sliced 3x3 BSR matrix (SELL-64).
//
// Semantics follow the reference's GPU solver and its CPU emulation (solver.ts:109-191, reduce.wgsl,
// update_dot.wgsl, residual.wgsl; pcg.rs:60-281): per-row 3-term dot products, fixed 256-row workgroup
// trees and a strided final pass. Built with -ffp-contract=off, the iterates are bit-identical to the
// oracle's f32 emulation.
//
// MI355X layout: one thread per block row, one wavefront per 64-row slice. A slice stores, for each
// block position k < its width, 64 column indices and 9 planes of 64 values, so every load of the SpMV
// is a 256-byte coalesced wavefront access; the gathered p vector is float4 per row (one 16-byte load).
// The scalar reductions that the reference runs in separate single-workgroup passes are folded into the
// consumers: every workgroup re-reduces the <= ceil(R/256) partials in the same fixed order, so one
// iteration is three launches (spmv+dot, update+dots, direction).
#include "../internal.h"
#include "coarse.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace drape {

// Group tree of per-row values held by physical position: the value goes to its logical slot first.
__device__ inline float tree_logical(float v, uint32_t row, float* lds) {
  const uint32_t t = row == 0xffffffffu ? threadIdx.x : row - blockIdx.x * 256;
  lds[t] = v;
  __syncthreads();
  const float own = lds[threadIdx.x];
  __syncthreads();
  return tree256(own, lds);
}
__device__ inline float2 tree_logical2(float a, float b, uint32_t row, float2* lds) {
  const uint32_t t = row == 0xffffffffu ? threadIdx.x : row - blockIdx.x * 256;
  lds[t] = make_float2(a, b);
  __syncthreads();
  const float2 own = lds[threadIdx.x];
  __syncthreads();
  return tree256x2(own.x, own.y, lds);
}


// jacobi_invert + cg_init: D^-1, x = 0, r = b, z = D^-1 r, p = z, partial (r.z, r.r).
__global__ void __launch_bounds__(256) k_pcg_setup(SellView A, const uint32_t* __restrict__ diag_slot,
                                                   const float4* __restrict__ b, float* __restrict__ dinv,
                                                   float4* x, float4* r, float4* z, float4* p,
                                                   float* __restrict__ partials, int R) {
  __shared__ float2 lds[256];
  const int row = blockIdx.x * 256 + threadIdx.x;
  float rz = 0.0f, rr = 0.0f;
  if (row < R) {
    const size_t d = diag_slot[row];
    float m[9], inv[9];
    for (int e = 0; e < 9; ++e) m[e] = A.val[d + 64 * e];
    invert3(m, inv);
    for (int e = 0; e < 9; ++e) dinv[9 * row + e] = inv[e];
    const float4 bb = b[row];
    x[row] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r[row] = bb;
    const float3 zz = precondition(dinv, row, bb.x, bb.y, bb.z);
    const float4 z4 = make_float4(zz.x, zz.y, zz.z, 0.0f);
    z[row] = z4;
    p[row] = z4;
    rz = bb.x * zz.x + bb.y * zz.y + bb.z * zz.z;
    rr = bb.x * bb.x + bb.y * bb.y + bb.z * bb.z;
  }
  const float2 sum = tree256x2(rz, rr, lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sum.x;
    partials[2 * blockIdx.x + 1] = sum.y;
  }
}

// reduce.wgsl MODE 0 (start).
__global__ void __launch_bounds__(256) k_pcg_start(const float* __restrict__ partials, int G, PcgState* st,
                                                   float tol2, float above2, int max_iterations, float diverge2) {
  __shared__ float2 lds[256];
  const float2 s = reduce_partials2(partials, G, lds);
  if (threadIdx.x == 0) {
    st->rz[0] = s.x;
    st->rz[1] = 0.0f;
    st->rr = s.y;
    st->rr0 = s.y;
    st->tol2 = tol2;
    st->above2 = above2;
    st->diverge2 = diverge2;
    st->max_iterations = max_iterations;
    st->iteration = 0;
    st->status = s.y <= 0.0f ? 1 : 0;
    st->alpha = st->beta = st->pq = 0.0f;
    st->true_rr = 0.0f;
  }
}

// spmv_dot.wgsl: q = A p, partial p.q (the first iteration of a chunk, and the timing probe).
__global__ void __launch_bounds__(256) k_pcg_spmv(SellView A, const float4* __restrict__ p, float4* __restrict__ q,
                                                  float* __restrict__ partial_pq, const PcgState* st, int R,
                                                  int ignore_status) {
  __shared__ float lds[256];
  if (!ignore_status && st->status != 0) return;
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  float pq = 0.0f;
  if (row != 0xffffffffu) {
    float3 a = row_times<false>(A, phys, p, nullptr, 0.0f);
    if (A.cptr) a = add3(a, contact_times<false>(A, row, p, nullptr, 0.0f));
    q[row] = make_float4(a.x, a.y, a.z, 0.0f);
    const float4 pr = p[row];
    pq = pr.x * a.x + pr.y * a.y + pr.z * a.z;
  }
  const float s = tree_logical(pq, row, lds);
  if (threadIdx.x == 0) partial_pq[blockIdx.x] = s;
}

// direction.wgsl of iteration it-1 fused with spmv_dot.wgsl of iteration it: beta and the convergence
// test from the (r.z, r.r) partials of the last update; then p_it = z + beta p_{it-1} for the own row
// (written to the other p buffer) and, for the gathered columns, the same value formed on the fly;
// q = A p_it and partial p.q.
__global__ void __launch_bounds__(256) k_pcg_dir_spmv(SellView A, const float* __restrict__ partial_rz, int G,
                                                      PcgState* st, const float4* __restrict__ z,
                                                      const float4* __restrict__ p_old, float4* __restrict__ p_new,
                                                      float4* __restrict__ q, float* __restrict__ partial_pq, int R,
                                                      int it) {
  __shared__ float lds[256];
  if (st->status != 0) return;
  const float rz_old = st->rz[(it - 1) & 1];
  const float2 s = wave_reduce_partials2(partial_rz, G);
  const float rz_new = s.x, rr = s.y;
  const float beta = rz_new / rz_old;
  int status = 0;
  if (rr <= st->tol2 * st->rr0) status = 1;
  else if (it >= st->max_iterations) status = 2;
  else if (rr > st->diverge2 * st->rr0) status = 3;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[it & 1] = rz_new;
    st->rr = rr;
    st->beta = beta;
    st->iteration = it;
    if (status != 0) st->status = status;
  }
  if (status != 0) return;
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  float pq = 0.0f;
  if (row != 0xffffffffu) {
    float3 a = row_times<true>(A, phys, p_old, z, beta);
    if (A.cptr) a = add3(a, contact_times<true>(A, row, p_old, z, beta));
    q[row] = make_float4(a.x, a.y, a.z, 0.0f);
    const float4 zr = z[row], po = p_old[row];
    float4 pr;
    pr.x = zr.x + beta * po.x;
    pr.y = zr.y + beta * po.y;
    pr.z = zr.z + beta * po.z;
    pr.w = 0.0f;
    p_new[row] = pr;
    pq = pr.x * a.x + pr.y * a.y + pr.z * a.z;
  }
  const float t = tree_logical(pq, row, lds);
  if (threadIdx.x == 0) partial_pq[blockIdx.x] = t;
}

// reduce.wgsl MODE 1 (alpha, breakdown) folded into update_dot.wgsl: x += alpha p, r -= alpha q,
// z = D^-1 r, partial (r.z, r.r). The row's operands are loaded before the pq reduction. (The two-level
// iteration folds the same update into its restriction and prolongation launches.)
__global__ void __launch_bounds__(256) k_pcg_update(const float* __restrict__ partial_pq, int G, PcgState* st,
                                                    const float4* __restrict__ p, const float4* __restrict__ q,
                                                    const float* __restrict__ dinv, float4* x, float4* r, float4* z,
                                                    float* __restrict__ partial_rz, int R, int it) {
  __shared__ float2 lds2[256];
  if (st->status != 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  const bool live = row < R;
  float4 pr = make_float4(0, 0, 0, 0), qr = pr, xr = pr, rv = pr;
  float d[9];
  if (live) {
    pr = p[row];
    qr = q[row];
    xr = x[row];
    rv = r[row];
#pragma unroll
    for (int e = 0; e < 9; ++e) d[e] = dinv[9 * row + e];
  }
  const float rzv = st->rz[it & 1];
  const float pq = wave_reduce_partials(partial_pq, G);
  const float alpha = rzv / pq;
  if (!(pq > 0.0f && finite32(alpha))) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->status = 3;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq;
    st->alpha = alpha;
  }
  float rz = 0.0f, rr = 0.0f;
  if (live) {
    xr.x = xr.x + alpha * pr.x;
    xr.y = xr.y + alpha * pr.y;
    xr.z = xr.z + alpha * pr.z;
    rv.x = rv.x - alpha * qr.x;
    rv.y = rv.y - alpha * qr.y;
    rv.z = rv.z - alpha * qr.z;
    x[row] = xr;
    r[row] = rv;
    const float z0 = d[0] * rv.x + d[1] * rv.y + d[2] * rv.z;
    const float z1 = d[3] * rv.x + d[4] * rv.y + d[5] * rv.z;
    const float z2 = d[6] * rv.x + d[7] * rv.y + d[8] * rv.z;
    z[row] = make_float4(z0, z1, z2, 0.0f);
    rz = rv.x * z0 + rv.y * z1 + rv.z * z2;
    rr = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  }
  const float2 sum = tree256x2(rz, rr, lds2);
  if (threadIdx.x == 0) {
    partial_rz[2 * blockIdx.x] = sum.x;
    partial_rz[2 * blockIdx.x + 1] = sum.y;
  }
}

// reduce.wgsl MODE 2 (beta, convergence, iteration limit) folded into direction.wgsl:
// p_{it+1} = z + beta p_it, from p_old into p_new.
__global__ void __launch_bounds__(256) k_pcg_direction(const float* __restrict__ partial_rz, int G, PcgState* st,
                                                       const float4* __restrict__ z, const float4* __restrict__ p_old,
                                                       float4* __restrict__ p_new, int R, int it) {
  if (st->status != 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  float4 zr = make_float4(0, 0, 0, 0), po = zr;
  if (row < R) {
    zr = z[row];
    po = p_old[row];
  }
  const float rz_old = st->rz[it & 1];
  const float2 s = wave_reduce_partials2(partial_rz, G);
  const float rz_new = s.x, rr = s.y;
  const float beta = rz_new / rz_old;
  int status = 0;
  if (rr <= st->tol2 * st->rr0) status = 1;
  else if (it + 1 >= st->max_iterations) status = 2;
  else if (rr > st->diverge2 * st->rr0) status = 3;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[(it + 1) & 1] = rz_new;
    st->rr = rr;
    st->beta = beta;
    st->iteration = it + 1;
    if (status != 0) st->status = status;
  }
  if (status != 0) return;
  if (row < R) {
    float4 pr;
    pr.x = zr.x + beta * po.x;
    pr.y = zr.y + beta * po.y;
    pr.z = zr.z + beta * po.z;
    pr.w = 0.0f;
    p_new[row] = pr;
  }
}

// residual.wgsl: s = b - A x; replace (r = s, z = D^-1 s, partial (r.z, r.r)) or measure (partial s.s).
__global__ void __launch_bounds__(256) k_pcg_residual(SellView A, const float4* __restrict__ b,
                                                      const float4* __restrict__ x, const float* __restrict__ dinv,
                                                      float4* r, float4* z, float* __restrict__ partials,
                                                      const PcgState* st, int R, int replace) {
  __shared__ float2 lds[256];
  // replace only while rr > above2 rr0 (pcg.rs:293-296; a NaN rr does not replace)
  if (replace && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  float a = 0.0f, c = 0.0f;
  if (row != 0xffffffffu) {
    float3 ax = row_times<false>(A, phys, x, nullptr, 0.0f);
    if (A.cptr) ax = add3(ax, contact_times<false>(A, row, x, nullptr, 0.0f));
    const float4 bb = b[row];
    const float s0 = bb.x - ax.x, s1 = bb.y - ax.y, s2 = bb.z - ax.z;
    const float ss = s0 * s0 + s1 * s1 + s2 * s2;
    if (replace) {
      const float3 zz = precondition(dinv, row, s0, s1, s2);
      r[row] = make_float4(s0, s1, s2, 0.0f);
      z[row] = make_float4(zz.x, zz.y, zz.z, 0.0f);
      a = s0 * zz.x + s1 * zz.y + s2 * zz.z;
      c = ss;
    } else {
      a = ss;
    }
  }
  const float2 sum = tree_logical2(a, c, row, lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sum.x;
    partials[2 * blockIdx.x + 1] = sum.y;
  }
}

// reduce.wgsl MODE 4 (replaced r.z, r.r) and MODE 3 (true residual).
__global__ void __launch_bounds__(256) k_pcg_residual_final(const float* __restrict__ partials, int G, PcgState* st,
                                                            int replace, int it) {
  __shared__ float2 lds[256];
  const int status = st->status;
  const float rr = st->rr, gate = st->above2 * st->rr0;
  __syncthreads();
  if (replace && (status != 0 || !(rr > gate))) return;
  const float2 s = reduce_partials2(partials, G, lds);
  if (threadIdx.x == 0) {
    if (replace) {
      st->rz[(it + 1) & 1] = s.x;
      st->rr = s.y;
    } else {
      st->true_rr = s.x;
    }
  }
}

// ---- two-level V-cycle (symmetric multiplicative two-level preconditioner) --------------------------
// z1 = w D^-1 r;  r1 = r - A z1;  y = B^T B P^T r1;  z2 = z1 + P y;  z = z2 + w D^-1 (r - A z2).
// Guards as in the coarse correction: 0 runs while the solve runs, 1 with a residual replacement, 2 always.
__device__ inline bool vc_skip(const PcgState* st, int guard) {
  return (guard == 0 && st->status != 0) || (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0)));
}

// z1 = w D^-1 r (setup and replacement)
__global__ void __launch_bounds__(256) k_vc_smooth(const float* __restrict__ dinv, const float4* __restrict__ r,
                                                   float4* __restrict__ z1, float omega, int R, const PcgState* st,
                                                   int guard) {
  if (vc_skip(st, guard)) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= R) return;
  const float4 rv = r[row];
  const float3 d = precondition(dinv, row, rv.x, rv.y, rv.z);
  z1[row] = make_float4(omega * d.x, omega * d.y, omega * d.z, 0.0f);
}

// One iteration's update (reduce.wgsl MODE 1 with update_dot.wgsl): alpha from the p.q partials, breakdown,
// x += alpha p, r -= alpha q; then z1 = w D^-1 r.
__global__ void __launch_bounds__(256) k_vc_update(const float* __restrict__ partial_pq, int G, PcgState* st,
                                                   const float4* __restrict__ p, const float4* __restrict__ q,
                                                   const float* __restrict__ dinv, float4* x, float4* r,
                                                   float4* __restrict__ z1, float omega, int R, int it) {
  if (st->status != 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  const bool live = row < R;
  float4 pr = make_float4(0, 0, 0, 0), qr = pr, xr = pr, rv = pr;
  if (live) {
    pr = p[row];
    qr = q[row];
    xr = x[row];
    rv = r[row];
  }
  const float rzv = st->rz[it & 1];
  const float pq = wave_reduce_partials(partial_pq, G);
  const float alpha = rzv / pq;
  if (!(pq > 0.0f && finite32(alpha))) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->status = 3;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq;
    st->alpha = alpha;
  }
  if (!live) return;
  xr.x = xr.x + alpha * pr.x;
  xr.y = xr.y + alpha * pr.y;
  xr.z = xr.z + alpha * pr.z;
  rv.x = rv.x - alpha * qr.x;
  rv.y = rv.y - alpha * qr.y;
  rv.z = rv.z - alpha * qr.z;
  x[row] = xr;
  r[row] = rv;
  const float3 d = precondition(dinv, row, rv.x, rv.y, rv.z);
  z1[row] = make_float4(omega * d.x, omega * d.y, omega * d.z, 0.0f);
}

// r1 = r - A z1
__global__ void __launch_bounds__(256) k_vc_residual1(SellView A, const float4* __restrict__ z1,
                                                      const float4* __restrict__ r, float4* __restrict__ r1,
                                                      const PcgState* st, int guard) {
  if (vc_skip(st, guard)) return;
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  if (row == 0xffffffffu) return;
  const float3 a = row_times<false>(A, phys, z1, nullptr, 0.0f);
  const float4 rv = r[row];
  r1[row] = make_float4(rv.x - a.x, rv.y - a.y, rv.z - a.z, 0.0f);
}

// z2 of a row: z1 + (P y)_row over its parent slots
__device__ inline float4 vc_z2(const float4* __restrict__ z1, const CoarseView& cv, uint32_t row) {
  float4 v = z1[row];
  coarse_prolong_add(cv.ppos, cv.w, cv.lever, cv.yc, (int)row, v.x, v.y, v.z);
  return v;
}

// A z2 with z2 formed for the gathered columns, then z = z2 + w D^-1 (r - A z2), partial (r.z, r.r) by
// logical row; with guard 2 also p = z.
__global__ void __launch_bounds__(256) k_vc_post(SellView A, const float4* __restrict__ z1, CoarseView cv,
                                                 const float* __restrict__ dinv, const float4* __restrict__ r,
                                                 float4* __restrict__ z, float4* __restrict__ p, float omega,
                                                 float* __restrict__ partials, const PcgState* st, int guard) {
  __shared__ float2 lds[256];
  if (vc_skip(st, guard)) return;
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  float rz = 0.0f, rr = 0.0f;
  if (row != 0xffffffffu) {
    const int sl = phys >> 6, l = phys & 63;
    const uint32_t base = A.slice_ptr[sl], n = A.cnt[phys];
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t q = base + k;
      const uint32_t col = A.col[q * 64 + l];
      const float* vp = A.val + (size_t)q * 9 * 64 + l;
      const float4 w = vc_z2(z1, cv, col);
      a0 = a0 + (vp[0] * w.x + vp[64] * w.y + vp[128] * w.z);
      a1 = a1 + (vp[192] * w.x + vp[256] * w.y + vp[320] * w.z);
      a2 = a2 + (vp[384] * w.x + vp[448] * w.y + vp[512] * w.z);
    }
    const float4 rv = r[row];
    const float4 own = vc_z2(z1, cv, row);
    const float3 d = precondition(dinv, row, rv.x - a0, rv.y - a1, rv.z - a2);
    const float4 zz = make_float4(own.x + omega * d.x, own.y + omega * d.y, own.z + omega * d.z, 0.0f);
    z[row] = zz;
    if (guard == 2) p[row] = zz;
    rz = rv.x * zz.x + rv.y * zz.y + rv.z * zz.z;
    rr = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  }
  const float2 sum = tree_logical2(rz, rr, row, lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sum.x;
    partials[2 * blockIdx.x + 1] = sum.y;
  }
}

// The system's view, with the self-contact coupling when self-contact terms are active. The V-cycle's
// smoothing residuals use the view without it (the mesh operator with the contact diagonal blocks): a valid
// symmetric preconditioner of the full operator, which PCG's own products apply.
static bool streamed(const Context& c) { return c.sell_slots * 64 * 40 > ((size_t)128 << 20); }
static SellView view_of(Context& c) {
  SellView v{c.slice_ptr, c.sell_col, c.sell_val, c.row_cnt, c.perm};
  v.stream = streamed(c);
  self_contact_view(c, &v.cptr, &v.cinc, &v.crow, &v.cdat, &v.ccoef);
  return v;
}
static SellView mesh_view_of(Context& c) {
  SellView v{c.slice_ptr, c.sell_col, c.sell_val, c.row_cnt, c.perm};
  v.stream = streamed(c);
  return v;
}

// z = M^-1 r by the V-cycle for the current r (after the caller's z1 = w D^-1 r): guard as above.
static int vcycle_rest(Context& c, float* partials, int guard) {
  hipStream_t s = c.stream;
  const float omega = (float)c.rules.smoother;
  SellView A = mesh_view_of(c);
  k_vc_residual1<<<c.G, 256, 0, s>>>(A, c.z1, c.r, c.r1, c.state, guard);
  TRY_RC(two_level_restrict_solve(c, c.r1, guard));
  k_vc_post<<<c.G, 256, 0, s>>>(A, c.z1, two_level_view(c), c.dinv, c.r, c.z, c.p, omega, partials, c.state, guard);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

static int vcycle_apply(Context& c, float* partials, int guard) {
  k_vc_smooth<<<c.G, 256, 0, c.stream>>>(c.dinv, c.r, c.z1, (float)c.rules.smoother, c.R, c.state, guard);
  return vcycle_rest(c, partials, guard);
}

// One PCG solve; `coarse` adds the two-level coarse correction to block-Jacobi (z = D^-1 r + P A_c^-1 P^T r).
static int pcg_run(Context& c, double tolerance, int max_iterations, drape_pcg_stats* stats, bool coarse) {
  const int G = c.G, R = c.R;
  float* part_setup = c.partials;
  float* part_pq = c.partials + 2 * G;
  float* part_rz = c.partials + 3 * G;
  float* part_res = c.partials + 5 * G;
  SellView A = view_of(c);
  const float tol2 = (float)(tolerance * tolerance);
  const float above2 = (float)(c.rules.replace_above * c.rules.replace_above);
  hipStream_t s = c.stream;
  HIP_TRY(hipEventRecord(c.ev0, s));
  k_pcg_setup<<<G, 256, 0, s>>>(A, c.diag_slot, c.b, c.dinv, c.x, c.r, c.z, c.p, part_setup, R);
  const bool vcycle = coarse && c.rules.smoother > 0.0;
  if (vcycle) TRY_RC(vcycle_apply(c, part_setup, 2));
  else if (coarse) TRY_RC(two_level_apply(c, part_setup, 2));
  // a two-level solve whose residual grows 1e4-fold has met curvature FP32 cannot resolve: a breakdown
  const float diverge2 = coarse ? 1e8f : INFINITY;
  k_pcg_start<<<1, 256, 0, s>>>(part_setup, G, c.state, tol2, above2, max_iterations, diverge2);
  HIP_TRY(hipGetLastError());
  // p ping-pongs between c.p (even iterations) and c.p2 (odd). A chunk opens with a plain SpMV (p is
  // current), fuses each later direction update into its SpMV, and closes with an explicit direction
  // update so that a residual replacement sees the updated p (pcg.rs:277-288 order).
  const int replace_every = c.rules.replace_every;
  const int chunk = replace_every > 0 ? replace_every : CHUNK_PLAIN;
  const int chunk_limit = (max_iterations + chunk - 1) / chunk + 1;
  const int max_launch = chunk_limit * chunk;
  // Iterations go out in polled groups: after each group the status is copied to pinned memory behind an
  // event, and the host waits for a group's status only once the next group is queued, so the device never
  // idles on the host and at most one group runs past the stop (its kernels return on the status).
  const int group = coarse ? 5 : 10;
  float4* pbuf[2] = {c.p, c.p2};
  int chunks = 0, it = 0;
  bool stopped = false;
  int pending[2], npending = 0, slot = 0;
  for (;;) {
    if (it < max_launch) {
      for (int q = 0; q < group && it < max_launch; ++q, ++it) {
        float4* pit = pbuf[it & 1];
        if (it % chunk == 0) {
          ++chunks;
          k_pcg_spmv<<<G, 256, 0, s>>>(A, pit, c.q, part_pq, c.state, R, 0);
        } else {
          k_pcg_dir_spmv<<<G, 256, 0, s>>>(A, part_rz, G, c.state, c.z, pbuf[(it - 1) & 1], pit, c.q, part_pq, R, it);
        }
        if (vcycle) {
          k_vc_update<<<G, 256, 0, s>>>(part_pq, G, c.state, pit, c.q, c.dinv, c.x, c.r, c.z1,
                                        (float)c.rules.smoother, R, it);
          TRY_RC(vcycle_rest(c, part_rz, 0));
        } else if (coarse) {
          TRY_RC(two_level_iterate(c, part_pq, pit, part_rz, it));
        } else {
          k_pcg_update<<<G, 256, 0, s>>>(part_pq, G, c.state, pit, c.q, c.dinv, c.x, c.r, c.z, part_rz, R, it);
        }
        if ((it + 1) % chunk == 0) {
          // direction update after the chunk's last iteration it: p_{it+1} = z + beta p_it into pbuf[(it+1) & 1]
          k_pcg_direction<<<G, 256, 0, s>>>(part_rz, G, c.state, c.z, pbuf[it & 1], pbuf[(it + 1) & 1], R, it);
          if (replace_every > 0) {
            k_pcg_residual<<<G, 256, 0, s>>>(A, c.b, c.x, c.dinv, c.r, c.z, part_res, c.state, R, 1);
            if (vcycle) TRY_RC(vcycle_apply(c, part_res, 1));
            else if (coarse) TRY_RC(two_level_apply(c, part_res, 1));
            k_pcg_residual_final<<<1, 256, 0, s>>>(part_res, G, c.state, 1, it);
          }
        }
      }
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(c.poll_host + slot, &c.state->status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipEventRecord(c.ev_poll[slot], s));
      pending[npending++] = slot;
      slot ^= 1;
      if (npending < 2 && it < max_launch) continue;
    }
    if (npending == 0) break;
    if (coarse) TRY_RC(two_level_launch_deferred(c));  // a refactorisation beside this solve, now the solve is queued
    const int oldest = pending[0];
    HIP_TRY(hipEventSynchronize(c.ev_poll[oldest]));
    c.read_bytes += (int64_t)sizeof(int32_t);
    pending[0] = pending[1];
    --npending;
    if (c.poll_host[oldest] != 0) {
      stopped = true;
      break;
    }
  }
  if (!stopped) {
    set_error("PCG never raised a stop status: the iteration counter or reduction is broken");
    return DRAPE_E_STATE;
  }
  k_pcg_residual<<<G, 256, 0, s>>>(A, c.b, c.x, c.dinv, c.r, c.z, part_res, c.state, R, 0);
  k_pcg_residual_final<<<1, 256, 0, s>>>(part_res, G, c.state, 0, 0);
  HIP_TRY(hipEventRecord(c.ev1, s));
  TRY_RC(read_back(c, c.state_host, c.state, sizeof(PcgState)));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  c.last_solve_ms = ms;
  const PcgState& h = *c.state_host;
  stats->status = h.status;
  stats->iterations = h.iteration;
  stats->recursive_relative_residual = std::sqrt((double)h.rr / (double)h.rr0);
  stats->true_relative_residual = std::sqrt((double)h.true_rr / (double)h.rr0);
  stats->chunks = chunks;
  stats->refinements = 0;
  stats->fp64_relative_residual = NAN;
  return DRAPE_OK;
}

// Truncated CG: a solve that met p^T A p <= 0 returns its last iterate before the negative curvature
// (status 4) when that iterate reduced the residual; the caller then steps along a descent direction.
static bool truncate_breakdown(Context& c, drape_pcg_stats* stats) {
  if (stats->iterations > 0 && stats->recursive_relative_residual < 1.0) {
    stats->status = 4;
    c.truncations++;
    return true;
  }
  return false;
}

// The preconditioner the rules select. A two-level solve is redone with block-Jacobi when the coarse operator
// is not positive definite; on a breakdown it is truncated (rules.fallback = 2), redone with block-Jacobi (1)
// or reported (0). A block-Jacobi breakdown is truncated under rules.block_jacobi_truncate, else reported.
// Truncations, reported breakdowns and block-Jacobi re-solves are counted apart.
static int pcg_solve_once(Context& c, double tolerance, int max_iterations, drape_pcg_stats* stats) {
  if (c.rules.preconditioner == DRAPE_PRECOND_TWO_LEVEL && c.two_level) {
    bool ok = false;
    TRY_RC(two_level_prepare(c, &ok));
    if (ok) {
      TRY_RC(pcg_run(c, tolerance, max_iterations, stats, true));
      if (stats->status != 3) {
        two_level_observe(c, stats->iterations, stats->recursive_relative_residual);
        return DRAPE_OK;
      }
      two_level_observe(c, -1, 0.0);
      if (c.rules.fallback == 2 && truncate_breakdown(c, stats)) return DRAPE_OK;
      if (c.rules.fallback != 1) {
        c.breakdowns++;
        return DRAPE_OK;
      }
    }
    c.fallbacks++;
  }
  TRY_RC(pcg_run(c, tolerance, max_iterations, stats, false));
  if (stats->status == 3) {
    if (c.rules.block_jacobi_truncate && truncate_breakdown(c, stats)) return DRAPE_OK;
    c.breakdowns++;
  }
  return DRAPE_OK;
}

// ---- FP64 iterative refinement around the FP32 PCG ------------------------------------------------
// r = b - A x with the FP32 matrix values and FP64 x and sums; partial |r|^2 per 256-row group by logical row.
__global__ void __launch_bounds__(256) k_residual64(SellView A, const double* __restrict__ x,
                                                   const float4* __restrict__ b, double* __restrict__ r,
                                                   double* __restrict__ part) {
  __shared__ double lds[256];
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  double rr = 0.0;
  if (row != 0xffffffffu) {
    const int sl = phys >> 6, l = phys & 63;
    const uint32_t base = A.slice_ptr[sl], n = A.cnt[phys];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t q = base + k;
      const uint32_t col = A.col[q * 64 + l];
      const float* v = A.val + (size_t)q * 9 * 64 + l;
      const double x0 = x[3 * col], x1 = x[3 * col + 1], x2 = x[3 * col + 2];
      a0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2;
      a1 += (double)v[192] * x0 + (double)v[256] * x1 + (double)v[320] * x2;
      a2 += (double)v[384] * x0 + (double)v[448] * x1 + (double)v[512] * x2;
    }
    if (A.cptr) {
      double ct[3];
      contact_times64(A, row, x, ct);
      a0 += ct[0];
      a1 += ct[1];
      a2 += ct[2];
    }
    const float4 bb = b[row];
    const double r0 = (double)bb.x - a0, r1 = (double)bb.y - a1, r2 = (double)bb.z - a2;
    r[3 * row] = r0;
    r[3 * row + 1] = r1;
    r[3 * row + 2] = r2;
    rr = r0 * r0 + r1 * r1 + r2 * r2;
  }
  // logical slot order, then a fixed tree
  const uint32_t t = row == 0xffffffffu ? threadIdx.x : row - blockIdx.x * 256;
  lds[t] = rr;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) lds[threadIdx.x] += lds[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

__global__ void __launch_bounds__(256) k_sum64(const double* __restrict__ part, int count, double* __restrict__ out) {
  __shared__ double lds[256];
  double a = 0.0;
  for (int j = threadIdx.x; j < count; j += 256) a += part[j];
  lds[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) lds[threadIdx.x] += lds[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

// b_pcg = f32(r); and after the inner solve x64 += x32
__global__ void k_refine_rhs(const double* __restrict__ r, float4* __restrict__ b, int R) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row < R) b[row] = make_float4((float)r[3 * row], (float)r[3 * row + 1], (float)r[3 * row + 2], 0.0f);
}
__global__ void k_refine_add(const float4* __restrict__ d, double* __restrict__ x, int R) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row < R) {
    const float4 v = d[row];
    x[3 * row] += (double)v.x;
    x[3 * row + 1] += (double)v.y;
    x[3 * row + 2] += (double)v.z;
  }
}
// x32 = f32(x64); b = saved b
__global__ void k_refine_finish(const double* __restrict__ x64, float4* __restrict__ x, const float4* __restrict__ saved,
                                float4* __restrict__ b, int R) {
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row < R) {
    x[row] = make_float4((float)x64[3 * row], (float)x64[3 * row + 1], (float)x64[3 * row + 2], 0.0f);
    b[row] = saved[row];
  }
}

constexpr int MAX_REFINEMENTS = 8;
constexpr double INNER_FLOOR = 1e-3;  // the FP32 solve is asked for no more than three decades at a time

// FP32 PCG inside an FP64 true-residual loop: r = b - A x in FP64, solve A d = r in FP32 to
// max(tol |b| / |r|, 1e-3), x += d in FP64, until |b - A x| <= tol |b| (at most 8 passes). The solution
// left in x (FP32) and every statistic refer to the FP64 iterate.
static int pcg_refine(Context& c, double tolerance, int max_iterations, drape_pcg_stats* stats) {
  const int R = c.R, G = c.G;
  const size_t Rp = (size_t)G * WG;
  double* x64 = c.refine;
  double* r64 = c.refine + 3 * Rp;
  hipStream_t s = c.stream;
  SellView A = view_of(c);
  HIP_TRY(hipMemcpyAsync(c.b_saved, c.b, (size_t)R * sizeof(float4), hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemsetAsync(x64, 0, 3 * Rp * sizeof(double), s));
  double bb = 0.0, rel = 1.0;
  int total = 0, passes = 0, status = 0;
  drape_pcg_stats inner{};
  for (int pass = 0;; ++pass) {
    k_residual64<<<G, 256, 0, s>>>(A, x64, c.b_saved, r64, c.dpart);
    k_sum64<<<1, 256, 0, s>>>(c.dpart, G, c.dout);
    HIP_TRY(hipGetLastError());
    TRY_RC(read_back(c, c.dout_host, c.dout, sizeof(double)));
    const double rr = c.dout_host[0];
    if (pass == 0) bb = rr;
    rel = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    if (!(rel > tolerance)) {
      status = 1;
      break;
    }
    if (pass == MAX_REFINEMENTS || total >= max_iterations) {
      status = 2;
      break;
    }
    k_refine_rhs<<<blocks_for(R), 256, 0, s>>>(r64, c.b, R);
    HIP_TRY(hipGetLastError());
    const double inner_tol = std::max(tolerance / rel, INNER_FLOOR);
    TRY_RC(pcg_solve_once(c, inner_tol, max_iterations - total, &inner));
    total += inner.iterations;
    if (inner.status == 3) {  // the correction is discarded; the Newton loop sees the breakdown
      status = 3;
      break;
    }
    k_refine_add<<<blocks_for(R), 256, 0, s>>>(c.x, x64, R);
    HIP_TRY(hipGetLastError());
    ++passes;
    if (inner.status == 4) {  // truncated at negative curvature: keep the descent step, stop refining
      status = 4;
      break;
    }
  }
  k_refine_finish<<<blocks_for(R), 256, 0, s>>>(x64, c.x, c.b_saved, c.b, R);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  stats->status = status;
  stats->iterations = total;
  stats->recursive_relative_residual = rel;
  stats->true_relative_residual = rel;
  stats->chunks = inner.chunks;
  stats->refinements = passes;
  stats->fp64_relative_residual = rel;
  return DRAPE_OK;
}

int pcg_solve(Context& c, double tolerance, int max_iterations, drape_pcg_stats* stats) {
  if (c.rules.refine_fp64) return pcg_refine(c, tolerance, max_iterations, stats);
  return pcg_solve_once(c, tolerance, max_iterations, stats);
}

// The SpMV alone, `launches` times back to back between two events on the context's stream (p -> q);
// returns the average time per launch in ms.
int spmv_timing(Context& c, int launches, double* avg_ms) {
  SellView A = view_of(c);
  hipStream_t s = c.stream;
  float* part_pq = c.partials + 2 * c.G;
  k_pcg_spmv<<<c.G, 256, 0, s>>>(A, c.p, c.q, part_pq, c.state, c.R, 1);
  HIP_TRY(hipEventRecord(c.ev_spmv0, s));
  for (int i = 0; i < launches; ++i) k_pcg_spmv<<<c.G, 256, 0, s>>>(A, c.p, c.q, part_pq, c.state, c.R, 1);
  HIP_TRY(hipEventRecord(c.ev_spmv1, s));
  HIP_TRY(hipEventSynchronize(c.ev_spmv1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c.ev_spmv0, c.ev_spmv1));
  *avg_ms = ms / launches;
  return DRAPE_OK;
}

// The fused direction update + SpMV of a PCG iteration alone, `launches` times back to back between two events
// on the context's stream (z, p -> p2, q), against a state of its own that never stops; average ms per launch.
int dir_spmv_timing(Context& c, int launches, double* avg_ms) {
  SellView A = view_of(c);
  hipStream_t s = c.stream;
  float* part_pq = c.partials + 2 * c.G;
  float* part_rz = c.partials + 3 * c.G;
  PcgState* st = nullptr;
  HIP_TRY(hipMalloc((void**)&st, sizeof(PcgState)));
  PcgState init{};
  init.rz[0] = init.rz[1] = 1.0f;
  init.rr0 = 1.0f;
  init.tol2 = -1.0f;                      // rr <= tol2 rr0 never holds
  init.max_iterations = 0x7fffffff;
  init.diverge2 = INFINITY;
  int rc = DRAPE_OK;
  auto run = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(st, &init, sizeof init, hipMemcpyHostToDevice, s));
    k_pcg_dir_spmv<<<c.G, 256, 0, s>>>(A, part_rz, c.G, st, c.z, c.p, c.p2, c.q, part_pq, c.R, 1);
    HIP_TRY(hipEventRecord(c.ev_spmv0, s));
    for (int i = 0; i < launches; ++i)
      k_pcg_dir_spmv<<<c.G, 256, 0, s>>>(A, part_rz, c.G, st, c.z, c.p, c.p2, c.q, part_pq, c.R, 1);
    HIP_TRY(hipEventRecord(c.ev_spmv1, s));
    HIP_TRY(hipEventSynchronize(c.ev_spmv1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c.ev_spmv0, c.ev_spmv1));
    *avg_ms = ms / launches;
    return DRAPE_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(s);
  (void)hipFree(st);
  return rc;
}

// A plain copy of n float4 (16 bytes each), grid-stride: the known-size kernel that scales hardware
// byte counters.
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int copy_timing(size_t bytes, int launches, double* avg_ms) {
  const size_t n = bytes / 16;
  if (n == 0 || launches <= 0) {
    set_error("copy timing needs at least 16 bytes and one launch");
    return DRAPE_E_INVALID;
  }
  float4 *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DRAPE_OK;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && rc == DRAPE_OK) {
      set_error(std::string("copy timing: ") + hipGetErrorString(e));
      rc = DRAPE_E_DEVICE;
    }
    return rc == DRAPE_OK;
  };
  if (check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && check(hipEventCreate(&e0)) &&
      check(hipEventCreate(&e1)) && check(hipMalloc((void**)&src, n * 16)) && check(hipMalloc((void**)&dst, n * 16)) &&
      check(hipMemsetAsync(src, 0, n * 16, s))) {
    const unsigned grid = 1024u * 8u;  // eight workgroups per CU slot of the 256-CU part, grid-stride beyond
    k_copy<<<grid, 256, 0, s>>>(src, dst, n);
    check(hipGetLastError());
    if (check(hipEventRecord(e0, s))) {
      for (int i = 0; i < launches; ++i) k_copy<<<grid, 256, 0, s>>>(src, dst, n);
      float ms = 0.0f;
      if (check(hipEventRecord(e1, s)) && check(hipEventSynchronize(e1)) && check(hipEventElapsedTime(&ms, e0, e1)))
        *avg_ms = ms / launches;
    }
  }
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

}  // namespace drape
