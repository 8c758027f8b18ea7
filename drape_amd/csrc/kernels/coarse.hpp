// Parent-slot arithmetic of the two-level prolongation P, shared by every kernel that applies P or P^T.
//
// Fine row i has four parent slots k (the corners of its coarse cell); slot k of row i holds the coarse node
// at position ppos[4 i + k] of the coarse vector (-1: empty), its weight w_ik and, with rotation unknowns,
// its lever v_ik (float4, w = 0). The slot's 3 x 3 (translations only) or 3 x 6 block of P is
//     [ w_ik I | S_ik ],   S_ik = -[v_ik]x   (column a = e_a x v_ik, a rotation of the coarse node),
// so P y adds w y_t + y_r x v and P^T r adds (w r, v x r) to the node's (translation, rotation) unknowns.
// The translation arithmetic is the same, operation for operation, with and without the rotation terms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drape {

// z += sum over the slots of `row` of w y_t (+ y_r x v when lever is not null)
__device__ inline void coarse_prolong_add(const int32_t* __restrict__ ppos, const float* __restrict__ w,
                                          const float4* __restrict__ lever, const float* __restrict__ yc, int row,
                                          float& z0, float& z1, float& z2) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int P = ppos[4 * row + k];
    if (P >= 0) {
      const float ww = w[4 * row + k];
      z0 = z0 + ww * yc[P];
      z1 = z1 + ww * yc[P + 1];
      z2 = z2 + ww * yc[P + 2];
      if (lever) {
        const float4 v = lever[4 * row + k];
        const float a = yc[P + 3], b = yc[P + 4], c = yc[P + 5];
        z0 = z0 + (b * v.z - c * v.y);
        z1 = z1 + (c * v.x - a * v.z);
        z2 = z2 + (a * v.y - b * v.x);
      }
    }
  }
}

// The operands of coarse_prolong_add for one row, loaded ahead of their use (the slots' nodes, weights and
// levers, and the coarse solution at the nodes), and their application in coarse_prolong_add's order.
struct SlotOperands {
  int P[4];
  float w[4];
  float4 lv[4];
  float y[4][6];
};

__device__ inline void coarse_prolong_load(const int32_t* __restrict__ ppos, const float* __restrict__ w,
                                           const float4* __restrict__ lever, const float* __restrict__ yc, int row,
                                           SlotOperands& o) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o.P[k] = ppos[4 * row + k];
    o.w[k] = w[4 * row + k];
    if (lever) o.lv[k] = lever[4 * row + k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (o.P[k] >= 0) {
#pragma unroll
      for (int e = 0; e < 3; ++e) o.y[k][e] = yc[o.P[k] + e];
      if (lever)
#pragma unroll
        for (int e = 3; e < 6; ++e) o.y[k][e] = yc[o.P[k] + e];
    }
}

__device__ inline void coarse_prolong_apply(const SlotOperands& o, bool rotations, float& z0, float& z1, float& z2) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (o.P[k] >= 0) {
      const float ww = o.w[k];
      z0 = z0 + ww * o.y[k][0];
      z1 = z1 + ww * o.y[k][1];
      z2 = z2 + ww * o.y[k][2];
      if (rotations) {
        const float4 v = o.lv[k];
        const float a = o.y[k][3], b = o.y[k][4], c = o.y[k][5];
        z0 = z0 + (b * v.z - c * v.y);
        z1 = z1 + (c * v.x - a * v.z);
        z2 = z2 + (a * v.y - b * v.x);
      }
    }
  }
}

// The rotation part of one slot's restriction: v x r
__device__ inline float3 lever_cross(const float4 v, float rx, float ry, float rz) {
  return make_float3(v.y * rz - v.z * ry, v.z * rx - v.x * rz, v.x * ry - v.y * rx);
}

}  // namespace drape
