// Inference of the learned prolongation: a three-layer message-passing network over the fine rows that turns
// the assembled system into the parent-slot weights of the two-level preconditioner (no reference
// counterpart). Per factorisation, on the device, deterministic (every sum in a fixed order), FP32:
//   features f_i (13): the four bilinear parent weights, the position (fx, fy) inside the coarse cell, the
//     diagonal block over s_i = trace(A_ii)/3 minus the identity (xx, yy, zz, symmetric xy, xz, yz), and the
//     coupling strength sum_j |A_ij|_F / (3 s_i) over the off-diagonal blocks (CSR order);
//   edge weights c_ij = |A_ij|_F / (3 sqrt(s_i s_j)), normalised per row by (sum_j c_ij + 1e-6);
//   h0 = relu(W0 f + b0); three layers h <- relu(Ws h_i + Wn m_i + b), m_i = sum_j c~_ij h_j;
//   delta = Wo h + bo; w_ik = mass_i softmax_k(log w_bilinear_ik + delta_ik) over the non-empty slots;
//   with the rotation head (coarse nodes with rotation unknowns), the rotation scale of each slot
//   sigma_ik = exp(clamp(Wr h + br, -3, 3))_k, which multiplies the slot's geometric lever, so a zero rotation
//   head keeps the levers of the current shape.
// Parameters: one flat array, W0 [32 x 13], b0 [32], then per layer Ws [32 x 32], Wn [32 x 32], b [32], then
// Wo [4 x 32], bo [4], and optionally Wr [4 x 32], br [4], row-major.
#include <cmath>

#include "../internal.h"

namespace drape {

namespace {

constexpr int F_IN = GNN_FEATURES, H = GNN_WIDTH;

// s_i = trace(A_ii) / 3
__global__ void __launch_bounds__(256) k_gnn_scale(const float* __restrict__ sell_val,
                                                   const uint32_t* __restrict__ diag_slot, int R,
                                                   float* __restrict__ s) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const float* d = sell_val + diag_slot[i];
  s[i] = (d[0] + d[64 * 4] + d[64 * 8]) / 3.0f;
}

// features, edge weights (normalised) and nothing else; one thread per row
__global__ void __launch_bounds__(256) k_gnn_features(const float* __restrict__ sell_val,
                                                      const uint32_t* __restrict__ block_dst,
                                                      const uint32_t* __restrict__ diag_slot,
                                                      const uint32_t* __restrict__ row_ptr,
                                                      const uint32_t* __restrict__ col, const float* __restrict__ s,
                                                      const float2* __restrict__ fxy,
                                                      const float* __restrict__ wbil, int R,
                                                      float* __restrict__ f, float* __restrict__ cw) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= R) return;
  const float si = s[i];
  const float* d = sell_val + diag_slot[i];
  float dn[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) dn[e] = d[64 * e] / si;
  float strength = 0.0f, norm = 0.0f;
  const uint32_t b0 = row_ptr[i], b1 = row_ptr[i + 1];
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t j = col[b];
    float c = 0.0f;
    if ((int)j != i) {
      const float* a = sell_val + block_dst[b];
      float n2 = 0.0f;
#pragma unroll
      for (int e = 0; e < 9; ++e) n2 = n2 + a[64 * e] * a[64 * e];
      const float nb = sqrtf(n2);
      strength = strength + nb;
      c = nb / (3.0f * sqrtf(si * s[j]));
      norm = norm + c;
    }
    cw[b] = c;
  }
  const float inv = 1.0f / (norm + 1e-6f);
  for (uint32_t b = b0; b < b1; ++b) cw[b] = cw[b] * inv;
  float* fi = f + (size_t)F_IN * i;
#pragma unroll
  for (int k = 0; k < 4; ++k) fi[k] = wbil[4 * i + k];
  const float2 p = fxy[i];
  fi[4] = p.x;
  fi[5] = p.y;
  fi[6] = dn[0] - 1.0f;
  fi[7] = dn[4] - 1.0f;
  fi[8] = dn[8] - 1.0f;
  fi[9] = 0.5f * (dn[1] + dn[3]);
  fi[10] = 0.5f * (dn[2] + dn[6]);
  fi[11] = 0.5f * (dn[5] + dn[7]);
  fi[12] = strength / (3.0f * si);
}

// h0 = relu(W0 f + b0): one thread per (row, output)
__global__ void __launch_bounds__(256) k_gnn_input(const float* __restrict__ f, const float* __restrict__ theta,
                                                   int R, float* __restrict__ h) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int i = t / H, o = t % H;
  const float* w = theta + F_IN * o;
  const float* fi = f + (size_t)F_IN * i;
  float acc = theta[H * F_IN + o];
#pragma unroll
  for (int k = 0; k < F_IN; ++k) acc = acc + w[k] * fi[k];
  h[t] = fmaxf(acc, 0.0f);
}

// m_i = sum_j c~_ij h_j over the row's off-diagonal blocks in CSR order: one thread per (row, feature)
__global__ void __launch_bounds__(256) k_gnn_message(const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col, const float* __restrict__ cw,
                                                     const float* __restrict__ h, int R, float* __restrict__ m) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int i = t / H, k = t % H;
  float acc = 0.0f;
  for (uint32_t b = row_ptr[i]; b < row_ptr[i + 1]; ++b) acc = acc + cw[b] * h[(size_t)H * col[b] + k];
  m[t] = acc;
}

// h_new = relu(Ws h + Wn m + b): one thread per (row, output); weights of the layer at `layer`
__global__ void __launch_bounds__(256) k_gnn_dense(const float* __restrict__ layer, const float* __restrict__ h,
                                                   const float* __restrict__ m, int R, float* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= R * H) return;
  const int i = t / H, o = t % H;
  const float* ws = layer + H * o;
  const float* wn = layer + H * H + H * o;
  const float* hi = h + (size_t)H * i;
  const float* mi = m + (size_t)H * i;
  float a = 0.0f, b = 0.0f;
#pragma unroll 8
  for (int k = 0; k < H; ++k) {
    a = a + ws[k] * hi[k];
    b = b + wn[k] * mi[k];
  }
  out[t] = fmaxf(a + b + layer[2 * H * H + o], 0.0f);
}

// delta = Wo h + bo, then the softmax over the non-empty slots scaled by the bilinear mass; with `sigma`, the
// rotation head's scales from Wr, br (after bo)
__global__ void __launch_bounds__(256) k_gnn_head(const float* __restrict__ head, const float* __restrict__ h,
                                                  const float* __restrict__ wbil, int R, float* __restrict__ w,
                                                  float* __restrict__ sigma) {
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
#pragma unroll
  for (int k = 0; k < 4; ++k) w[4 * i + k] = mass * (e[k] / sum);
  if (sigma) {
    const float* rot = head + 4 * H + 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float o = rot[4 * H + k];
      for (int j = 0; j < H; ++j) o = o + rot[H * k + j] * hi[j];
      sigma[4 * i + k] = expf(fminf(fmaxf(o, -3.0f), 3.0f));
    }
  }
}

}  // namespace

int gnn_param_count(bool rotation_head) {
  return H * F_IN + H + 3 * (2 * H * H + H) + 4 * H + 4 + (rotation_head ? 4 * H + 4 : 0);
}

int gnn_features(const GnnArgs& a) {
  hipStream_t s = a.stream;
  const int R = a.R;
  k_gnn_scale<<<blocks_for(R), 256, 0, s>>>(a.sell_val, a.diag_slot, R, a.scale);
  k_gnn_features<<<blocks_for(R), 256, 0, s>>>(a.sell_val, a.block_dst, a.diag_slot, a.row_ptr, a.col, a.scale, a.fxy,
                                               a.wbil, R, a.features, a.edge);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int gnn_infer(const GnnArgs& a) {
  hipStream_t s = a.stream;
  const int R = a.R;
  TRY_RC(gnn_features(a));
  k_gnn_input<<<blocks_for(R * H), 256, 0, s>>>(a.features, a.theta, R, a.h0);
  const float* layer = a.theta + H * F_IN + H;
  float* h = a.h0;
  float* out = a.h1;
  for (int l = 0; l < GNN_LAYERS; ++l) {
    k_gnn_message<<<blocks_for(R * H), 256, 0, s>>>(a.row_ptr, a.col, a.edge, h, R, a.msg);
    k_gnn_dense<<<blocks_for(R * H), 256, 0, s>>>(layer, h, a.msg, R, out);
    layer += 2 * H * H + H;
    float* tmp = h;
    h = out;
    out = tmp;
  }
  k_gnn_head<<<blocks_for(R), 256, 0, s>>>(layer, h, a.wbil, R, a.w, a.sigma);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int gnn_forward_keep(const GnnArgs& a, float* hs, float* ms) {
  hipStream_t s = a.stream;
  const int R = a.R;
  const size_t plane = (size_t)H * R;
  TRY_RC(gnn_features(a));
  k_gnn_input<<<blocks_for(R * H), 256, 0, s>>>(a.features, a.theta, R, hs);
  const float* layer = a.theta + H * F_IN + H;
  for (int l = 0; l < GNN_LAYERS; ++l) {
    k_gnn_message<<<blocks_for(R * H), 256, 0, s>>>(a.row_ptr, a.col, a.edge, hs + l * plane, R, ms + l * plane);
    k_gnn_dense<<<blocks_for(R * H), 256, 0, s>>>(layer, hs + l * plane, ms + l * plane, R, hs + (l + 1) * plane);
    layer += 2 * H * H + H;
  }
  k_gnn_head<<<blocks_for(R), 256, 0, s>>>(layer, hs + GNN_LAYERS * plane, a.wbil, R, a.w, a.sigma);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

}  // namespace drape
