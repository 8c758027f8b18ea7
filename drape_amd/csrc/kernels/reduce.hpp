// Deterministic workgroup reductions for 256-thread workgroups (4 wavefronts of 64).
//
// tree256: lane[t] += lane[t + s] for s = 128, 64 (LDS), then 32..1 inside wavefront 0 (DPP/permute
// via __shfl_down). The addition order is exactly the fixed 256-wide tree of the reference's WGSL
// reductions and of its CPU emulation (pcg.rs:446-476), so FP32 results are reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace drape {

// Returns the tree sum in thread 0 (other threads: unspecified). `lds` holds 256 floats.
__device__ inline float tree256(float v, float* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  if (t < 128) lds[t] = lds[t] + lds[t + 128];
  __syncthreads();
  float s = 0.0f;
  if (t < 64) {
    s = lds[t] + lds[t + 64];
    s = s + __shfl_down(s, 32, 64);
    s = s + __shfl_down(s, 16, 64);
    s = s + __shfl_down(s, 8, 64);
    s = s + __shfl_down(s, 4, 64);
    s = s + __shfl_down(s, 2, 64);
    s = s + __shfl_down(s, 1, 64);
  }
  __syncthreads();
  return s;
}

__device__ inline float2 tree256x2(float a, float b, float2* lds) {
  const int t = threadIdx.x;
  lds[t] = make_float2(a, b);
  __syncthreads();
  if (t < 128) {
    float2 u = lds[t], w = lds[t + 128];
    lds[t] = make_float2(u.x + w.x, u.y + w.y);
  }
  __syncthreads();
  float sa = 0.0f, sb = 0.0f;
  if (t < 64) {
    float2 u = lds[t], w = lds[t + 64];
    sa = u.x + w.x;
    sb = u.y + w.y;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      sa = sa + __shfl_down(sa, s, 64);
      sb = sb + __shfl_down(sb, s, 64);
    }
  }
  __syncthreads();
  return make_float2(sa, sb);
}

// The final pass of the reference's reductions: lane t sums partials t, t+256, ... in order starting
// from 0, then the tree. The result is broadcast to every thread.
__device__ inline float reduce_partials(const float* partials, int count, float* lds) {
  float a = 0.0f;
  for (int j = threadIdx.x; j < count; j += 256) a = a + partials[j];
  float s = tree256(a, lds);
  if (threadIdx.x == 0) lds[0] = s;
  __syncthreads();
  float out = lds[0];
  __syncthreads();
  return out;
}

__device__ inline float2 reduce_partials2(const float* partials, int count, float2* lds) {
  float a = 0.0f, b = 0.0f;
  for (int j = threadIdx.x; j < count; j += 256) {
    a = a + partials[2 * j];
    b = b + partials[2 * j + 1];
  }
  float2 s = tree256x2(a, b, lds);
  if (threadIdx.x == 0) lds[0] = s;
  __syncthreads();
  float2 out = lds[0];
  __syncthreads();
  return out;
}

// FP64 variants (fixed order; used for the Newton loop's sums).
__device__ inline double2 tree256d2(double a, double b, double2* lds) {
  const int t = threadIdx.x;
  lds[t] = make_double2(a, b);
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      double2 u = lds[t], w = lds[t + s];
      lds[t] = make_double2(u.x + w.x, u.y + w.y);
    }
    __syncthreads();
  }
  double2 out = lds[0];
  __syncthreads();
  return out;
}

__device__ inline bool finite32(float v) { return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u; }

// The reference's final reduction pass (reduce.wgsl:33-53, pcg.rs:447-457) inside one wavefront: lane l
// stands for the four tree lanes l, l+64, l+128, l+192 (each the strided sum of its partials), combines
// them as the 128- and 64-strides of the tree do, and finishes with shuffles. No LDS, no barrier: every
// wavefront obtains the identical result.
__device__ inline float wave_reduce_partials(const float* __restrict__ part, int count) {
  const int l = threadIdx.x & 63;
  float v[4];
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    float a = 0.0f;
    for (int j = l + 64 * qd; j < count; j += 256) a = a + part[j];
    v[qd] = a;
  }
  float s = (v[0] + v[2]) + (v[1] + v[3]);
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) s = s + __shfl_down(s, st, 64);
  return __shfl(s, 0, 64);
}

// wave_reduce_partials / wave_reduce_partials2 in two phases, so that a kernel can issue these loads first and
// its own before the sums need them (loads complete in issue order): start loads the first entry of each of
// the lane's four strides, finish forms the same sums in the same order (the rest of each stride loaded there)
// and the same tree.
__device__ inline void partials_start(const float* __restrict__ part, int count, float raw[4]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    const int j = l + 64 * qd;
    raw[qd] = j < count ? part[j] : 0.0f;
  }
}

__device__ inline float partials_finish(const float* __restrict__ part, int count, const float raw[4]) {
  const int l = threadIdx.x & 63;
  float v[4];
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    float a = 0.0f;
    int j = l + 64 * qd;
    if (j < count) {
      a = a + raw[qd];
      for (j += 256; j < count; j += 256) a = a + part[j];
    }
    v[qd] = a;
  }
  float s = (v[0] + v[2]) + (v[1] + v[3]);
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) s = s + __shfl_down(s, st, 64);
  return __shfl(s, 0, 64);
}

__device__ inline void partials2_start(const float* __restrict__ part, int count, float ra[4], float rb[4]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    const int j = l + 64 * qd;
    ra[qd] = j < count ? part[2 * j] : 0.0f;
    rb[qd] = j < count ? part[2 * j + 1] : 0.0f;
  }
}

__device__ inline float2 partials2_finish(const float* __restrict__ part, int count, const float ra[4],
                                          const float rb[4]) {
  const int l = threadIdx.x & 63;
  float va[4], vb[4];
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    float a = 0.0f, b = 0.0f;
    int j = l + 64 * qd;
    if (j < count) {
      a = a + ra[qd];
      b = b + rb[qd];
      for (j += 256; j < count; j += 256) {
        a = a + part[2 * j];
        b = b + part[2 * j + 1];
      }
    }
    va[qd] = a;
    vb[qd] = b;
  }
  float sa = (va[0] + va[2]) + (va[1] + va[3]);
  float sb = (vb[0] + vb[2]) + (vb[1] + vb[3]);
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) {
    sa = sa + __shfl_down(sa, st, 64);
    sb = sb + __shfl_down(sb, st, 64);
  }
  return make_float2(__shfl(sa, 0, 64), __shfl(sb, 0, 64));
}

__device__ inline float2 wave_reduce_partials2(const float* __restrict__ part, int count) {
  const int l = threadIdx.x & 63;
  float va[4], vb[4];
#pragma unroll
  for (int qd = 0; qd < 4; ++qd) {
    float a = 0.0f, b = 0.0f;
    for (int j = l + 64 * qd; j < count; j += 256) {
      a = a + part[2 * j];
      b = b + part[2 * j + 1];
    }
    va[qd] = a;
    vb[qd] = b;
  }
  float sa = (va[0] + va[2]) + (va[1] + va[3]);
  float sb = (vb[0] + vb[2]) + (vb[1] + vb[3]);
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) {
    sa = sa + __shfl_down(sa, st, 64);
    sb = sb + __shfl_down(sb, st, 64);
  }
  return make_float2(__shfl(sa, 0, 64), __shfl(sb, 0, 64));
}

}  // namespace drape
