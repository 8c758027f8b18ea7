// Training of the learned prolongation on the device (no reference counterpart): the network's forward pass
// with its activations kept, the two-level projection loss over a set of error vectors with its gradient in
// the prolongation weights, and the backward pass through the softmax head, the three message-passing layers
// and the input layer. Every sum is taken in a fixed order (no atomics): a gradient is reproducible bit for
// bit.
//
// Loss of one assembled system A over error vectors e_1..e_n, weights w of the network at theta:
//   L = (1/n) sum_e |e - P c_e|_A^2 / |e|_A^2,   c_e = A_c^-1 P^T A e,   A_c = P^T A P.
// Galerkin orthogonality, P^T A (e - P c_e) = 0, removes the derivative through c_e:
//   dL/dw_ik = -(2/n) sum_e (A eps_e)_i . c_e[parent(i, k)] / |e|_A^2,   eps_e = e - P c_e.
// With rotation unknowns, slot block [w I | -[v]x] and lever v = w sigma d: the lever's gradient is
//   g_v = -(2/n) sum_e (A eps_e)_i x c_r / |e|_A^2   (c_r the parent's rotation unknowns),
// which adds g_v . v / w to dL/dw (through v) and gives dL/dsigma = g_v . v / sigma for the rotation head.
// A_c^-1 is the preconditioner's own factorisation (FP64 banded Cholesky, applied as B^T B with the FP32
// inverse factor B = L^-1, also where the preconditioner applies the FP16 copy); the products with A are the
// solver's SELL-64 rows.
//
// Error vectors: given by the caller, or made on the device from uniform random starts in [-1, 1) (SplitMix64
// of the seed and the entry index, Steele, Lea and Flood 2014) by damped block-Jacobi sweeps
// e <- e - omega D^-1 A e, which leave the errors the smoother cannot reduce: those the coarse space is for.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../internal.h"
#include "reduce.hpp"
#include "sell.hpp"

namespace drape {

namespace {

constexpr int F_IN = GNN_FEATURES, H = GNN_WIDTH;
constexpr int OUTER_ROWS = 128;                 // rows per chunk of the outer-product reductions
constexpr int LAYER_STRIDE = 2 * H * H + H;     // Ws, Wn, b of one layer
constexpr int OFF_LAYERS = H * F_IN + H;        // after W0, b0
constexpr int OFF_HEAD = OFF_LAYERS + GNN_LAYERS * LAYER_STRIDE;  // Wo [4 x H], bo [4]
constexpr int OFF_ROT = OFF_HEAD + 4 * H + 4;                       // the rotation head: Wr [4 x H], br [4]

}  // namespace

struct GnnTrain {
  int count = 0;               // error vectors held
  int capacity = 0;            // vectors allocated
  int chunks = 0;              // OUTER_ROWS-row chunks of the outer-product reductions
  float4* d_E = nullptr;       // [capacity][R] error vectors (w = 0)
  float4* d_q = nullptr;       // [R] A e, then A eps
  float4* d_eps = nullptr;     // [R] e - P c
  float4* d_tmp = nullptr;     // [R] sweep buffer
  float* d_dinv = nullptr;     // [9R] block-Jacobi inverse of the system the vectors were made on
  float* d_hs = nullptr;       // [4][32R] h0..h3
  float* d_ms = nullptr;       // [3][32R] messages of the three layers
  float* d_gw = nullptr;       // [4R] dL/dw
  float* d_gs = nullptr;       // [4R] dL/dsigma (rotation scales)
  float* d_go = nullptr;       // [4R] dL/do (rotation head outputs)
  float* d_gz = nullptr;       // [4R] dL/dz (softmax inputs)
  float* d_gh = nullptr;       // [32R] dL/dh
  float* d_gu = nullptr;       // [32R] dL/du (pre-activation)
  float* d_t = nullptr;        // [32R] gu Wn
  uint32_t* d_tpos = nullptr;  // [B] CSR position of the transposed block
  float* d_partial = nullptr;  // [2G] per-group dot partials: |e|_A^2 | |eps|_A^2
  float* d_opart = nullptr;    // [chunks][32 x 33] outer-product partials
  float* d_grad = nullptr;     // [params]
  double* d_sums = nullptr;    // [capacity][2] |e|_A^2, |eps|_A^2
};

void gnn_train_destroy(GnnTrain* g) {
  for (void* p : {(void*)g->d_E, (void*)g->d_q, (void*)g->d_eps, (void*)g->d_tmp, (void*)g->d_dinv, (void*)g->d_hs,
                  (void*)g->d_ms, (void*)g->d_gw, (void*)g->d_gs, (void*)g->d_go, (void*)g->d_gz, (void*)g->d_gh,
                  (void*)g->d_gu, (void*)g->d_t,
                  (void*)g->d_tpos, (void*)g->d_partial, (void*)g->d_opart, (void*)g->d_grad, (void*)g->d_sums})
    if (p) (void)hipFree(p);
  delete g;
}

namespace {

SellView view_of(const Context& c) { return {c.slice_ptr, c.sell_col, c.sell_val, c.row_cnt, c.perm}; }

// ---- error vectors ------------------------------------------------------------------------------

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// entry (v, i, d) of the start: 24 random bits mapped to [-1, 1) exactly
__global__ void __launch_bounds__(256) k_random_start(float4* __restrict__ E, int R, int count, uint64_t seed) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)R * count) return;
  float u[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint64_t h = splitmix64(seed ^ splitmix64(3 * t + d));
    u[d] = (float)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
  }
  E[t] = make_float4(u[0], u[1], u[2], 0.0f);
}

__global__ void __launch_bounds__(256) k_block_inverse(const float* __restrict__ sell_val,
                                                       const uint32_t* __restrict__ diag_slot, int R,
                                                       float* __restrict__ dinv) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  float m[9];
  const float* d = sell_val + diag_slot[i];
#pragma unroll
  for (int e = 0; e < 9; ++e) m[e] = d[64 * e];
  invert3(m, dinv + 9 * i);
}

// out = e - omega D^-1 A e, one thread per physical row position
__global__ void __launch_bounds__(256) k_sweep(SellView A, const float* __restrict__ dinv,
                                               const float4* __restrict__ e, float4* __restrict__ out, float omega) {
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  if (row == 0xffffffffu) return;
  const float3 a = row_times<false>(A, phys, e, nullptr, 0.0f);
  const float3 d = precondition(dinv, row, a.x, a.y, a.z);
  const float4 v = e[row];
  out[row] = make_float4(v.x - omega * d.x, v.y - omega * d.y, v.z - omega * d.z, 0.0f);
}

// ---- loss and weight gradient ---------------------------------------------------------------------

// q = A x and the group partial of x.q (physical order within the group, fixed tree)
__global__ void __launch_bounds__(256) k_spmv_dot(SellView A, const float4* __restrict__ x, float4* __restrict__ q,
                                                  float* __restrict__ partial) {
  __shared__ float lds[256];
  const int phys = blockIdx.x * 256 + threadIdx.x;
  const uint32_t row = A.perm[phys];
  float dot = 0.0f;
  if (row != 0xffffffffu) {
    const float3 a = row_times<false>(A, phys, x, nullptr, 0.0f);
    q[row] = make_float4(a.x, a.y, a.z, 0.0f);
    const float4 v = x[row];
    dot = v.x * a.x + v.y * a.y + v.z * a.z;
  }
  const float s = tree256(dot, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// eps = e - P c (ppos: the position of each parent slot's node in c, -1 when empty; lever: the slots' rotation
// levers, null without rotation unknowns)
__global__ void __launch_bounds__(256) k_eps(const int32_t* __restrict__ ppos, const float* __restrict__ w,
                                             const float4* __restrict__ lever, const float* __restrict__ yc,
                                             const float4* __restrict__ e, float4* __restrict__ eps, int R) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  float4 v = e[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int P = ppos[4 * i + k];
    if (P >= 0) {
      const float ww = w[4 * i + k];
      v.x = v.x - ww * yc[P];
      v.y = v.y - ww * yc[P + 1];
      v.z = v.z - ww * yc[P + 2];
      if (lever) {
        const float4 lv = lever[4 * i + k];
        const float a = yc[P + 3], b = yc[P + 4], c = yc[P + 5];
        v.x = v.x - (b * lv.z - c * lv.y);
        v.y = v.y - (c * lv.x - a * lv.z);
        v.z = v.z - (a * lv.y - b * lv.x);
      }
    }
  }
  eps[i] = v;
}

// out[0] = sum of partial[0..G), out[1] = sum of partial[G..2G), in FP64: lane sums in index order, then a
// fixed tree.
__global__ void __launch_bounds__(256) k_sum_pair(const float* __restrict__ partial, int G, double* __restrict__ out) {
  __shared__ double lds[2][256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int j = t; j < G; j += 256) {
    a += (double)partial[j];
    b += (double)partial[G + j];
  }
  lds[0][t] = a;
  lds[1][t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      lds[0][t] += lds[0][t + s];
      lds[1][t] += lds[1][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = lds[0][0];
    out[1] = lds[1][0];
  }
}

// gw[i, k] (+)= -2 / (n |e|_A^2) (A eps)_i . c[parent(i, k)]  (the first vector writes, the others add); with
// levers also the lever terms: g_v = -2 / (n |e|_A^2) (A eps)_i x c_r, gw += g_v . v / w, gs (+)= g_v . v / sigma
// (sigma: the rotation scales, 1 when null)
__global__ void __launch_bounds__(256) k_weight_grad(const int32_t* __restrict__ ppos, const float* __restrict__ yc,
                                                     const float4* __restrict__ aeps, const double* __restrict__ sums,
                                                     int n, int first, float* __restrict__ gw, int R,
                                                     const float* __restrict__ w, const float4* __restrict__ lever,
                                                     const float* __restrict__ sigma, float* __restrict__ gs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const float scale = (float)(-2.0 / ((double)n * sums[0]));
  const float4 a = aeps[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int P = ppos[4 * i + k];
    float v = 0.0f, vs = 0.0f;
    if (P >= 0) {
      v = scale * (a.x * yc[P] + a.y * yc[P + 1] + a.z * yc[P + 2]);
      if (lever) {
        const float4 lv = lever[4 * i + k];
        const float cx = yc[P + 3], cy = yc[P + 4], cz = yc[P + 5];
        const float gx = scale * (a.y * cz - a.z * cy), gy = scale * (a.z * cx - a.x * cz),
                    gz = scale * (a.x * cy - a.y * cx);
        const float gv = gx * lv.x + gy * lv.y + gz * lv.z;
        v = v + gv / w[4 * i + k];
        vs = gv / (sigma ? sigma[4 * i + k] : 1.0f);
      }
    }
    gw[4 * i + k] = first ? v : gw[4 * i + k] + v;
    if (gs) gs[4 * i + k] = first ? vs : gs[4 * i + k] + vs;
  }
}

// The rotation head's backward: o = br + Wr h (recomputed as the forward head does), sigma = exp(clamp(o, -3, 3)),
// go = gs sigma inside the clamp, 0 outside.
__global__ void __launch_bounds__(256) k_rotation_back(const float* __restrict__ rot, const float* __restrict__ h,
                                                       const float* __restrict__ gs, int R, float* __restrict__ go) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const float* hi = h + (size_t)H * i;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float o = rot[4 * H + k];
    for (int j = 0; j < H; ++j) o = o + rot[H * k + j] * hi[j];
    const float sg = expf(fminf(fmaxf(o, -3.0f), 3.0f));
    go[4 * i + k] = (o > -3.0f && o < 3.0f) ? gs[4 * i + k] * sg : 0.0f;
  }
}

// ---- backward ---------------------------------------------------------------------------------------

// w = mass softmax(z), z_k = log wbil_k + (Wo h + bo)_k over the valid slots: recomputes the softmax as
// the forward head does, then dz_k = mass p_k (gw_k - sum_j gw_j p_j) on the valid slots, 0 elsewhere.
__global__ void __launch_bounds__(256) k_softmax_back(const float* __restrict__ head, const float* __restrict__ h,
                                                      const float* __restrict__ wbil, const float* __restrict__ gw,
                                                      int R, float* __restrict__ gz) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const float* hi = h + (size_t)H * i;
  float z[4], mass = 0.0f, zmax = -INFINITY;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float wb = wbil[4 * i + k];
    float d = head[4 * H + k];
    for (int j = 0; j < H; ++j) d = d + head[H * k + j] * hi[j];
    z[k] = wb > 0.0f ? logf(wb) + d : -INFINITY;
    mass = mass + wb;
    zmax = fmaxf(zmax, z[k]);
  }
  float e[4], sum = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    e[k] = wbil[4 * i + k] > 0.0f ? expf(z[k] - zmax) : 0.0f;
    sum = sum + e[k];
  }
  float p[4], gp = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    p[k] = sum > 0.0f ? e[k] / sum : 0.0f;
    gp = gp + gw[4 * i + k] * p[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    gz[4 * i + k] = wbil[4 * i + k] > 0.0f ? mass * p[k] * (gw[4 * i + k] - gp) : 0.0f;
}

// Chunk partials of the outer-product sums over rows: partial[chunk][o][k] = sum over the chunk's rows of
// a[i][o] b[i][k] for k < BW, and of a[i][o] for k = BW (the bias). Rows in index order.
template <int AW, int BW>
__global__ void __launch_bounds__(256) k_outer(const float* __restrict__ a, const float* __restrict__ b, int R,
                                               float* __restrict__ partial) {
  __shared__ float sa[OUTER_ROWS * AW];
  __shared__ float sb[OUTER_ROWS * BW];
  const int r0 = blockIdx.x * OUTER_ROWS;
  const int rows = min(OUTER_ROWS, R - r0);
  for (int j = threadIdx.x; j < OUTER_ROWS * AW; j += 256)
    sa[j] = j < rows * AW ? a[(size_t)r0 * AW + j] : 0.0f;
  for (int j = threadIdx.x; j < OUTER_ROWS * BW; j += 256)
    sb[j] = j < rows * BW ? b[(size_t)r0 * BW + j] : 0.0f;
  __syncthreads();
  constexpr int OUT = AW * (BW + 1);
  float* dst = partial + (size_t)blockIdx.x * OUT;
  for (int idx = threadIdx.x; idx < OUT; idx += 256) {
    const int o = idx / (BW + 1), k = idx % (BW + 1);
    float acc = 0.0f;
    if (k < BW) {
      for (int i = 0; i < rows; ++i) acc = acc + sa[i * AW + o] * sb[i * BW + k];
    } else {
      for (int i = 0; i < rows; ++i) acc = acc + sa[i * AW + o];
    }
    dst[idx] = acc;
  }
}

// W[o][k] = sum over the chunks in order of partial[chunk][o][k] (FP64 accumulation), bias[o] likewise from
// column BW (skipped when bias is null).
__global__ void __launch_bounds__(256) k_outer_sum(const float* __restrict__ partial, int chunks, int AW, int BW,
                                                   float* __restrict__ W, float* __restrict__ bias) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  const int OUT = AW * (BW + 1);
  if (idx >= OUT) return;
  double acc = 0.0;
  for (int c = 0; c < chunks; ++c) acc += (double)partial[(size_t)c * OUT + idx];
  const int o = idx / (BW + 1), k = idx % (BW + 1);
  if (k < BW)
    W[o * BW + k] = (float)acc;
  else if (bias)
    bias[o] = (float)acc;
}

// gh[i][j] = sum_k gz[i][k] Wo[k][j] (+ sum_k go[i][k] Wr[k][j] with the rotation head)
__global__ void __launch_bounds__(256) k_head_back(const float* __restrict__ Wo, const float* __restrict__ gz,
                                                   const float* __restrict__ Wr, const float* __restrict__ go, int R,
                                                   float* __restrict__ gh) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int i = t / H, j = t % H;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) acc = acc + gz[4 * i + k] * Wo[H * k + j];
  if (go) {
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = acc + go[4 * i + k] * Wr[H * k + j];
  }
  gh[t] = acc;
}

// gu = gh where the layer's output is positive (relu), else 0
__global__ void __launch_bounds__(256) k_relu_back(const float* __restrict__ gh, const float* __restrict__ h, int n,
                                                   float* __restrict__ gu) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  gu[t] = h[t] > 0.0f ? gh[t] : 0.0f;
}

// t[i][k] = sum_o gu[i][o] Wn[o][k]
__global__ void __launch_bounds__(256) k_times_wn(const float* __restrict__ gu, const float* __restrict__ Wn, int R,
                                                  float* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int i = t / H, k = t % H;
  const float* g = gu + (size_t)H * i;
  float acc = 0.0f;
#pragma unroll 8
  for (int o = 0; o < H; ++o) acc = acc + g[o] * Wn[H * o + k];
  out[t] = acc;
}

// gh[j][k] = sum_o gu[j][o] Ws[o][k] + sum over the blocks b of row j of c~[transpose(b)] t[col(b)][k]: the
// message m_i = sum_j c~_ij h_j transposed through the symmetric pattern, in CSR order.
__global__ void __launch_bounds__(256) k_layer_back(const float* __restrict__ gu, const float* __restrict__ Ws,
                                                    const float* __restrict__ tw, const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, const float* __restrict__ cw,
                                                    const uint32_t* __restrict__ tpos, int R, float* __restrict__ gh) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int j = t / H, k = t % H;
  const float* g = gu + (size_t)H * j;
  float acc = 0.0f;
#pragma unroll 8
  for (int o = 0; o < H; ++o) acc = acc + g[o] * Ws[H * o + k];
  float msg = 0.0f;
  for (uint32_t b = row_ptr[j]; b < row_ptr[j + 1]; ++b) msg = msg + cw[tpos[b]] * tw[(size_t)H * col[b] + k];
  gh[t] = acc + msg;
}

template <typename T>
int alloc_n(T** p, size_t n) {
  HIP_TRY(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
  return DRAPE_OK;
}

// The buffers that do not depend on the number of vectors, and the transpose map of the pattern.
int ensure_state(Context& c) {
  if (c.gnn_train) return DRAPE_OK;
  GnnTrain* g = new GnnTrain();
  c.gnn_train = g;
  const size_t R = (size_t)c.R, Rp = (size_t)c.G * WG, plane = (size_t)H * R;
  g->chunks = (c.R + OUTER_ROWS - 1) / OUTER_ROWS;
  TRY_RC(alloc_n(&g->d_q, Rp));
  TRY_RC(alloc_n(&g->d_eps, Rp));
  TRY_RC(alloc_n(&g->d_tmp, Rp));
  TRY_RC(alloc_n(&g->d_dinv, 9 * R));
  TRY_RC(alloc_n(&g->d_hs, (GNN_LAYERS + 1) * plane));
  TRY_RC(alloc_n(&g->d_ms, GNN_LAYERS * plane));
  TRY_RC(alloc_n(&g->d_gw, 4 * R));
  TRY_RC(alloc_n(&g->d_gs, 4 * R));
  TRY_RC(alloc_n(&g->d_go, 4 * R));
  TRY_RC(alloc_n(&g->d_gz, 4 * R));
  TRY_RC(alloc_n(&g->d_gh, plane));
  TRY_RC(alloc_n(&g->d_gu, plane));
  TRY_RC(alloc_n(&g->d_t, plane));
  TRY_RC(alloc_n(&g->d_partial, 2 * (size_t)c.G));
  TRY_RC(alloc_n(&g->d_opart, (size_t)g->chunks * H * (H + 1)));
  TRY_RC(alloc_n(&g->d_grad, (size_t)gnn_param_count(true)));
  // the pattern is symmetric: block (i, j) at b has its transpose (j, i) in row j, columns sorted
  const std::vector<uint32_t>& rp = c.row_ptr_host;
  const std::vector<uint32_t>& ci = c.col_idx_host;
  std::vector<uint32_t> tpos(ci.size());
  for (int i = 0; i < c.R; ++i)
    for (uint32_t b = rp[i]; b < rp[i + 1]; ++b) {
      const uint32_t j = ci[b];
      const auto first = ci.begin() + rp[j], last = ci.begin() + rp[j + 1];
      const auto it = std::lower_bound(first, last, (uint32_t)i);
      if (it == last || *it != (uint32_t)i) {
        set_error("prolongation training: the block pattern is not symmetric with sorted columns");
        return DRAPE_E_INVALID;
      }
      tpos[b] = (uint32_t)(it - ci.begin());
    }
  TRY_RC(alloc_n(&g->d_tpos, tpos.size()));
  HIP_TRY(hipMemcpyAsync(g->d_tpos, tpos.data(), tpos.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

int ensure_vectors(Context& c, int count) {
  GnnTrain* g = c.gnn_train;
  if (count > g->capacity) {
    if (g->d_E) HIP_TRY(hipFree(g->d_E));
    if (g->d_sums) HIP_TRY(hipFree(g->d_sums));
    g->d_E = nullptr;
    g->d_sums = nullptr;
    g->capacity = 0;
    TRY_RC(alloc_n(&g->d_E, (size_t)count * c.R));
    TRY_RC(alloc_n(&g->d_sums, 2 * (size_t)count));
    g->capacity = count;
  }
  g->count = count;
  return DRAPE_OK;
}

int check_context(Context& c) {
  if (c.linear || !c.two_level) {
    set_error("prolongation training needs a mesh context with a two-level layout");
    return DRAPE_E_INVALID;
  }
  return ensure_state(c);
}

}  // namespace

int gnn_train_set_vectors(Context& c, const float* vectors, int count) {
  if (!vectors || count < 1) {
    set_error("training vectors: null array or count < 1");
    return DRAPE_E_INVALID;
  }
  TRY_RC(check_context(c));
  TRY_RC(ensure_vectors(c, count));
  const size_t R = (size_t)c.R;
  std::vector<float4> v((size_t)count * R);
  for (size_t e = 0; e < (size_t)count; ++e)
    for (size_t i = 0; i < R; ++i) {
      const float* s = vectors + e * 3 * R + 3 * i;
      v[e * R + i] = make_float4(s[0], s[1], s[2], 0.0f);
    }
  HIP_TRY(hipMemcpyAsync(c.gnn_train->d_E, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

int gnn_train_make_vectors(Context& c, int count, int sweeps, float omega, uint64_t seed, float* out) {
  if (count < 1 || sweeps < 0) {
    set_error("training vectors: count < 1 or sweeps < 0");
    return DRAPE_E_INVALID;
  }
  TRY_RC(check_context(c));
  TRY_RC(ensure_vectors(c, count));
  GnnTrain* g = c.gnn_train;
  hipStream_t s = c.stream;
  const int R = c.R;
  const SellView A = view_of(c);
  k_random_start<<<(int)(((size_t)R * count + 255) / 256), 256, 0, s>>>(g->d_E, R, count, seed);
  k_block_inverse<<<blocks_for(R), 256, 0, s>>>(c.sell_val, c.diag_slot, R, g->d_dinv);
  for (int e = 0; e < count; ++e) {
    float4* E = g->d_E + (size_t)e * R;
    for (int k = 0; k < sweeps; ++k) {
      k_sweep<<<c.G, 256, 0, s>>>(A, g->d_dinv, E, g->d_tmp, omega);
      HIP_TRY(hipMemcpyAsync(E, g->d_tmp, (size_t)R * sizeof(float4), hipMemcpyDeviceToDevice, s));
    }
  }
  HIP_TRY(hipGetLastError());
  if (out) {
    std::vector<float4> v((size_t)count * R);
    HIP_TRY(hipMemcpyAsync(v.data(), g->d_E, v.size() * sizeof(float4), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t t = 0; t < v.size(); ++t) {
      out[3 * t] = v[t].x;
      out[3 * t + 1] = v[t].y;
      out[3 * t + 2] = v[t].z;
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

int gnn_train_gradient(Context& c, const float* params, int count, double* loss, float* grad) {
  if (!params || !loss || !grad) {
    set_error("prolongation gradient: null argument");
    return DRAPE_E_INVALID;
  }
  TRY_RC(check_context(c));
  GnnTrain* g = c.gnn_train;
  if (g->count < 1) {
    set_error("prolongation gradient: no training vectors (set or make them first)");
    return DRAPE_E_INVALID;
  }
  TRY_RC(two_level_set_model(c, params, count));  // checks the parameter count
  hipStream_t s = c.stream;
  const int R = c.R, n = g->count;
  const size_t plane = (size_t)H * R;
  // forward with the activations kept; its weights become the factorisation's
  GnnArgs a = two_level_gnn_args(c);
  TRY_RC(gnn_forward_keep(a, g->d_hs, g->d_ms));
  bool ok = false;
  TRY_RC(two_level_factor_weights(c, &ok));
  if (!ok) {
    set_error("prolongation gradient: the coarse operator at these parameters is not positive definite");
    return DRAPE_E_INVALID;
  }
  // loss terms and dL/dw
  const SellView A = view_of(c);
  const CoarseView cv = two_level_view(c);
  for (int e = 0; e < n; ++e) {
    const float4* E = g->d_E + (size_t)e * R;
    k_spmv_dot<<<c.G, 256, 0, s>>>(A, E, g->d_q, g->d_partial);
    TRY_RC(two_level_restrict_solve(c, g->d_q, 2, true));
    k_eps<<<blocks_for(R), 256, 0, s>>>(cv.ppos, cv.w, cv.lever, cv.yc, E, g->d_eps, R);
    k_spmv_dot<<<c.G, 256, 0, s>>>(A, g->d_eps, g->d_q, g->d_partial + c.G);
    k_sum_pair<<<1, 256, 0, s>>>(g->d_partial, c.G, g->d_sums + 2 * e);
    k_weight_grad<<<blocks_for(R), 256, 0, s>>>(cv.ppos, cv.yc, g->d_q, g->d_sums + 2 * e, n, e == 0, g->d_gw, R,
                                                cv.w, cv.lever, a.sigma, cv.lever ? g->d_gs : nullptr);
  }
  // backward
  const float* theta = a.theta;
  float* dg = g->d_grad;
  const int sum_blocks = blocks_for(H * (H + 1));
  const float* h_top = g->d_hs + GNN_LAYERS * plane;
  k_softmax_back<<<blocks_for(R), 256, 0, s>>>(theta + OFF_HEAD, h_top, a.wbil, g->d_gw, R, g->d_gz);
  k_outer<4, H><<<g->chunks, 256, 0, s>>>(g->d_gz, h_top, R, g->d_opart);
  k_outer_sum<<<sum_blocks, 256, 0, s>>>(g->d_opart, g->chunks, 4, H, dg + OFF_HEAD, dg + OFF_HEAD + 4 * H);
  const bool rot_head = a.sigma != nullptr;  // the network has a rotation head and the coarse space rotations
  if (rot_head) {
    const float* rot = theta + OFF_ROT;
    k_rotation_back<<<blocks_for(R), 256, 0, s>>>(rot, h_top, g->d_gs, R, g->d_go);
    k_outer<4, H><<<g->chunks, 256, 0, s>>>(g->d_go, h_top, R, g->d_opart);
    k_outer_sum<<<sum_blocks, 256, 0, s>>>(g->d_opart, g->chunks, 4, H, dg + OFF_ROT, dg + OFF_ROT + 4 * H);
  } else if (count == gnn_param_count(true)) {
    HIP_TRY(hipMemsetAsync(dg + OFF_ROT, 0, (4 * H + 4) * sizeof(float), s));  // the head is not used
  }
  k_head_back<<<blocks_for(R * H), 256, 0, s>>>(theta + OFF_HEAD, g->d_gz, rot_head ? theta + OFF_ROT : nullptr,
                                                rot_head ? g->d_go : nullptr, R, g->d_gh);
  for (int l = GNN_LAYERS - 1; l >= 0; --l) {
    const int base = OFF_LAYERS + l * LAYER_STRIDE;
    const float* h_in = g->d_hs + l * plane;
    k_relu_back<<<blocks_for(R * H), 256, 0, s>>>(g->d_gh, g->d_hs + (l + 1) * plane, R * H, g->d_gu);
    k_outer<H, H><<<g->chunks, 256, 0, s>>>(g->d_gu, h_in, R, g->d_opart);
    k_outer_sum<<<sum_blocks, 256, 0, s>>>(g->d_opart, g->chunks, H, H, dg + base, dg + base + 2 * H * H);
    k_outer<H, H><<<g->chunks, 256, 0, s>>>(g->d_gu, g->d_ms + l * plane, R, g->d_opart);
    k_outer_sum<<<sum_blocks, 256, 0, s>>>(g->d_opart, g->chunks, H, H, dg + base + H * H, nullptr);
    k_times_wn<<<blocks_for(R * H), 256, 0, s>>>(g->d_gu, theta + base + H * H, R, g->d_t);
    k_layer_back<<<blocks_for(R * H), 256, 0, s>>>(g->d_gu, theta + base, g->d_t, c.csr_row_ptr, c.csr_col, a.edge,
                                                   g->d_tpos, R, g->d_gh);
  }
  k_relu_back<<<blocks_for(R * H), 256, 0, s>>>(g->d_gh, g->d_hs, R * H, g->d_gu);
  k_outer<H, F_IN><<<g->chunks, 256, 0, s>>>(g->d_gu, a.features, R, g->d_opart);
  k_outer_sum<<<sum_blocks, 256, 0, s>>>(g->d_opart, g->chunks, H, F_IN, dg, dg + H * F_IN);
  HIP_TRY(hipGetLastError());
  std::vector<double> sums(2 * (size_t)n);
  HIP_TRY(hipMemcpyAsync(sums.data(), g->d_sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(grad, dg, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double L = 0.0;
  for (int e = 0; e < n; ++e) L += sums[2 * e + 1] / sums[2 * e];
  *loss = L / n;
  return DRAPE_OK;
}

}  // namespace drape
