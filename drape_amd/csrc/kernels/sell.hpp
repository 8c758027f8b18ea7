// Sliced 3x3 BSR (SELL-64) access shared by the solver and training kernels: the layout view, the per-row
// product in the reference's column order, block-Jacobi application and the 3x3 inverse.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace drape {

struct SellView {
  const uint32_t* slice_ptr;  // per physical slice: first block slot
  const uint32_t* col;        // [slots * 64] logical column of (slot, lane)
  const float* val;           // [slots * 9 * 64] entry e of (slot q, lane l) at (9q + e) * 64 + l
  const uint32_t* cnt;        // blocks of the row at each physical position (0 past the last row)
  const uint32_t* perm;       // physical position -> logical row (same 256-row group); ~0u past the last row
  // the matrix exceeds what the caches hold between products: its loads are non-temporal, so that the
  // gathered vector stays cached (at n1025 the SpMV takes 86 us against 96; at n129, where the matrix stays
  // in the Infinity Cache between products, ordinary loads are faster)
  bool stream = false;
  // self-contact coupling outside the block pattern (null: none). Pair c has rows crow[c] (-1: pinned vertex)
  // and Hessian b'' (c_k n)(c_l n)^T in block (k, l), cdat[c] = (n, b'') and ccoef[c] = c_0..c_3; the rows'
  // pair slots are cinc[cptr[row] .. cptr[row + 1]) = 4 pair + k.
  const uint32_t* cptr = nullptr;
  const uint32_t* cinc = nullptr;
  const int4* crow = nullptr;
  const float4* cdat = nullptr;
  const float4* ccoef = nullptr;
};

// Blocks of a row processed per chunk: every load of a chunk is issued before the first use, so a row
// costs one column-load and one gather latency per chunk instead of one per block. 16 covers a whole row of
// the sheets (at most 11 blocks): at n129 the SpMV takes 4.41 us against 4.67 us with chunks of 8, and the
// HBM-bound n1025 SpMV is unchanged.
constexpr int CHUNK = 16;

// One chunk of a row's blocks in registers: column, the nine values and the gathered vector entries (with the
// fused direction update also z at the column). Every load of a chunk is issued before the first use.
template <bool DIR>
struct RowChunk {
  uint32_t c[CHUNK];
  float v[CHUNK][9];
  float4 pc[CHUNK], zc[CHUNK];
};

template <bool DIR, bool STREAM>
__device__ inline void chunk_load(const SellView& A, uint32_t base, int l, uint32_t n, uint32_t k0,
                                  const float4* __restrict__ p, const float4* __restrict__ z, RowChunk<DIR>& rc) {
#pragma unroll
  for (int j = 0; j < CHUNK; ++j) {
    if (k0 + j < n) {
      const uint32_t q = base + k0 + j;
      const float* vp = A.val + (size_t)q * 9 * 64 + l;
      if (STREAM) {  // read once per product: keep the caches for the gathered vector
        rc.c[j] = __builtin_nontemporal_load(A.col + q * 64 + l);
#pragma unroll
        for (int e = 0; e < 9; ++e) rc.v[j][e] = __builtin_nontemporal_load(vp + 64 * e);
      } else {
        rc.c[j] = A.col[q * 64 + l];
#pragma unroll
        for (int e = 0; e < 9; ++e) rc.v[j][e] = vp[64 * e];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CHUNK; ++j) {
    if (k0 + j < n) {
      rc.pc[j] = p[rc.c[j]];
      if (DIR) rc.zc[j] = z[rc.c[j]];
    }
  }
}

// acc[r] = acc[r] + (v[3r] p0 + v[3r+1] p1 + v[3r+2] p2) over the chunk's blocks in column order (the
// reference's per-row order, pcg.rs:359-369): contributions are formed independently, the sum stays
// sequential. With DIR the gathered value is z[c] + beta p[c].
template <bool DIR>
__device__ inline void chunk_add(const RowChunk<DIR>& rc, uint32_t n, uint32_t k0, float beta, float& a0, float& a1,
                                 float& a2) {
#pragma unroll
  for (int j = 0; j < CHUNK; ++j) {
    if (k0 + j < n) {
      float4 w = rc.pc[j];
      if (DIR) {
        w.x = rc.zc[j].x + beta * w.x;
        w.y = rc.zc[j].y + beta * w.y;
        w.z = rc.zc[j].z + beta * w.z;
      }
      a0 = a0 + (rc.v[j][0] * w.x + rc.v[j][1] * w.y + rc.v[j][2] * w.z);
      a1 = a1 + (rc.v[j][3] * w.x + rc.v[j][4] * w.y + rc.v[j][5] * w.z);
      a2 = a2 + (rc.v[j][6] * w.x + rc.v[j][7] * w.y + rc.v[j][8] * w.z);
    }
  }
}

// The row product, chunk by chunk: one column-load and one gather latency per chunk instead of one per block.
template <bool DIR, bool STREAM>
__device__ inline float3 row_times_(const SellView& A, int phys, const float4* __restrict__ p,
                                    const float4* __restrict__ z, float beta) {
  const int s = phys >> 6, l = phys & 63;
  const uint32_t base = A.slice_ptr[s];
  const uint32_t n = A.cnt[phys];
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (uint32_t k0 = 0; k0 < n; k0 += CHUNK) {
    RowChunk<DIR> rc;
    chunk_load<DIR, STREAM>(A, base, l, n, k0, p, z, rc);
    chunk_add<DIR>(rc, n, k0, beta, a0, a1, a2);
  }
  return make_float3(a0, a1, a2);
}

// The row product with the matrix loads chosen once per launch (uniform branch, two straight-line bodies).
template <bool DIR>
__device__ inline float3 row_times(const SellView& A, int phys, const float4* __restrict__ p,
                                   const float4* __restrict__ z, float beta) {
  return A.stream ? row_times_<DIR, true>(A, phys, p, z, beta) : row_times_<DIR, false>(A, phys, p, z, beta);
}

__device__ inline float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }

// The self-contact coupling of a logical row: for each contact pair it belongs to (slot k), in the row's
// list order, b'' c_k n (sum over the pair's other free vertices l of c_l n.w_l), w = p or, with DIR,
// z + beta p. The diagonal blocks b'' c_k^2 n n^T are part of the assembled matrix.
__device__ inline float coef_of(const float4& c, int k) { return k == 0 ? c.x : k == 1 ? c.y : k == 2 ? c.z : c.w; }
__device__ inline int row_of(const int4& r, int k) { return k == 0 ? r.x : k == 1 ? r.y : k == 2 ? r.z : r.w; }

template <bool DIR>
__device__ inline float3 contact_times(const SellView& A, uint32_t row, const float4* __restrict__ p,
                                       const float4* __restrict__ z, float beta) {
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  const uint32_t t1 = A.cptr[row + 1];
  for (uint32_t t = A.cptr[row]; t < t1; ++t) {
    const uint32_t e = A.cinc[t];
    const uint32_t pair = e >> 2;
    const int k = (int)(e & 3u);
    const int4 rows = A.crow[pair];
    const float4 d = A.cdat[pair], c = A.ccoef[pair];
    float s = 0.0f;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int rl = row_of(rows, l);
      if (l == k || rl < 0) continue;
      float4 w = p[rl];
      if (DIR) {
        const float4 zz = z[rl];
        w.x = zz.x + beta * w.x;
        w.y = zz.y + beta * w.y;
        w.z = zz.z + beta * w.z;
      }
      s = s + coef_of(c, l) * (d.x * w.x + d.y * w.y + d.z * w.z);
    }
    const float f = d.w * coef_of(c, k) * s;
    a0 = a0 + f * d.x;
    a1 = a1 + f * d.y;
    a2 = a2 + f * d.z;
  }
  return make_float3(a0, a1, a2);
}

// The same in FP64 over an FP64 vector x [3R] (the refinement's true residual).
__device__ inline void contact_times64(const SellView& A, uint32_t row, const double* __restrict__ x, double out[3]) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const uint32_t t1 = A.cptr[row + 1];
  for (uint32_t t = A.cptr[row]; t < t1; ++t) {
    const uint32_t e = A.cinc[t];
    const uint32_t pair = e >> 2;
    const int k = (int)(e & 3u);
    const int4 rows = A.crow[pair];
    const float4 d = A.cdat[pair], c = A.ccoef[pair];
    double s = 0.0;
    for (int l = 0; l < 4; ++l) {
      const int rl = row_of(rows, l);
      if (l == k || rl < 0) continue;
      s += (double)coef_of(c, l) * ((double)d.x * x[3 * rl] + (double)d.y * x[3 * rl + 1] + (double)d.z * x[3 * rl + 2]);
    }
    const double f = (double)d.w * (double)coef_of(c, k) * s;
    a0 += f * d.x;
    a1 += f * d.y;
    a2 += f * d.z;
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

__device__ inline float3 precondition(const float* __restrict__ dinv, int row, float a, float b, float c) {
  const float* d = dinv + 9 * row;
  return make_float3(d[0] * a + d[1] * b + d[2] * c, d[3] * a + d[4] * b + d[5] * c, d[6] * a + d[7] * b + d[8] * c);
}

// pcg.rs:268-281 invert3 (adjugate / determinant), same operation order.
__device__ inline void invert3(const float m[9], float* out) {
  float a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  float c00 = e * i - f * h, c01 = c * h - b * i, c02 = b * f - c * e;
  float c10 = f * g - d * i, c11 = a * i - c * g, c12 = c * d - a * f;
  float c20 = d * h - e * g, c21 = b * g - a * h, c22 = a * e - b * d;
  float det = a * c00 + b * c10 + c * c20;
  out[0] = c00 / det; out[1] = c01 / det; out[2] = c02 / det;
  out[3] = c10 / det; out[4] = c11 / det; out[5] = c12 / det;
  out[6] = c20 / det; out[7] = c21 / det; out[8] = c22 / det;
}

}  // namespace drape
