// Two-level preconditioner for the Newton systems: block-Jacobi plus a Galerkin coarse correction,
//     M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T A P,
// additive and symmetric, so PCG keeps its SPD preconditioner whenever A_c is SPD. There is no reference
// counterpart: the reference's PCG is block-Jacobi only (solver.ts:109-191) and names a hierarchical
// preconditioner as its next step. The coarse nodes are every fourth vertex of the sheet grid in each
// direction (1/16 of the vertices); P = W (x) I3 with four parent slots per fine row whose weights are
// bilinear by default or set by the learned prolongation network.
//
// Device work per factorisation: the Galerkin product (one thread per coarse 3x3 block, contributions in a
// fixed order, FP64 accumulation), the FP64 Cholesky factor A_c = L L^T and its inverse B = L^-1 (rocSOLVER
// potrf/trtri: 3Nc <= ~3300 unknowns), and FP32 copies of B and B^T. The coarse solve is applied as
// B^T (B r_c): a Gram product, so the FP32 operator stays positive semidefinite whatever the rounding of B
// (an FP32 copy of A_c^-1 itself turns indefinite once cond(A_c) nears 1e7). Per PCG iteration: restriction
// (one wavefront per coarse node), two triangular GEMVs (one wavefront per row, float4 loads, fixed shuffle
// reduction) and the prolongation fused with the block-Jacobi term and the (r.z, r.r) partial sums.
// Everything is deterministic. The factorisation is kept across Newton systems while mu and the step size
// are unchanged and PCG stays within twice the iterations per decade of the first solve after it: any
// fixed SPD operator is a valid preconditioner, and a stale Galerkin operator of a nearby system is one.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <vector>

#include "../internal.h"
#include "reduce.hpp"

namespace drape {

struct TwoLevel {
  int nc = 0;        // coarse grid columns
  int Nc = 0;        // coarse rows (free coarse nodes)
  int N3 = 0;        // 3 Nc
  int N3p = 0;       // padded row stride of the FP32 inverse (multiple of 64 floats)
  int npairs = 0;    // coarse block pairs
  uint64_t factored = ~0ull;  // system version of the current factorisation (~0: none)
  uint64_t factored_shift = 0;  // mu / step-size version of the current factorisation
  bool ok = false;            // the last factorisation succeeded
  bool refresh = false;       // PCG slowed down: refactor at the next new system
  double base_rate = 0.0;     // iterations per decade of the first solve after the factorisation
  int solves = 0;             // solves since the factorisation
  std::vector<int32_t> parent;   // [4R] coarse row of each parent slot, -1 when empty
  std::vector<float> bilinear;   // [4R] default weights
  int32_t* d_parent = nullptr;
  float* d_w = nullptr;          // [4R] current weights
  uint32_t* d_child_ptr = nullptr;  // [Nc+1]
  uint32_t* d_child = nullptr;      // 4 i + k of each child slot, in fine-row order
  uint32_t* d_pair_ptr = nullptr;   // [npairs+1]
  uint32_t* d_pair_key = nullptr;   // [npairs] I * Nc + J
  uint4* d_pair_terms = nullptr;    // (block b, slot ik, slot jl, 0)
  double* d_Ac = nullptr;           // [N3 * N3] column-major; L, then L^-1 in the lower triangle
  float* d_B = nullptr;             // [N3 * N3p] row-major B = L^-1 (zero above the diagonal)
  float* d_BT = nullptr;            // [N3 * N3p] row-major B^T (zero below the diagonal)
  float* d_t = nullptr;             // [N3p] B r_c
  float* d_rc = nullptr;            // [N3p]
  float* d_yc = nullptr;            // [N3p]
  int* d_info = nullptr;
  int* h_info = nullptr;
  rocblas_handle handle = nullptr;
  int64_t factorisations = 0, failures = 0;
  double factor_ms = 0.0;     // wall time of all factorisations (Galerkin product, potrf, trtri, copies)
};

void two_level_destroy(TwoLevel* t) {
  if (!t) return;
  for (void* p : {(void*)t->d_parent, (void*)t->d_w, (void*)t->d_child_ptr, (void*)t->d_child, (void*)t->d_pair_ptr,
                  (void*)t->d_pair_key, (void*)t->d_pair_terms, (void*)t->d_Ac, (void*)t->d_B, (void*)t->d_BT,
                  (void*)t->d_t, (void*)t->d_rc, (void*)t->d_yc, (void*)t->d_info})
    if (p) (void)hipFree(p);
  if (t->h_info) (void)hipHostFree(t->h_info);
  if (t->handle) rocblas_destroy_handle(t->handle);
  delete t;
}

// ---- kernels ----------------------------------------------------------------------------------
// A_c(I, J) = sum over fine blocks b = (i, j) and parent slots k of i, l of j with parent I, J of
// w_ik w_jl A_b, in the fixed order of the pair's term list.
__global__ void __launch_bounds__(256) k_galerkin(const uint32_t* __restrict__ pair_ptr,
                                                  const uint32_t* __restrict__ pair_key,
                                                  const uint4* __restrict__ terms, const float* __restrict__ w,
                                                  const float* __restrict__ sell_val,
                                                  const uint32_t* __restrict__ block_dst, int npairs, int Nc,
                                                  double* __restrict__ Ac) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t t = pair_ptr[p]; t < pair_ptr[p + 1]; ++t) {
    const uint4 e = terms[t];
    const double ww = (double)w[e.y] * (double)w[e.z];
    const float* a = sell_val + block_dst[e.x];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] += ww * (double)a[64 * q];
  }
  const uint32_t key = pair_key[p];
  const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
  const size_t N3 = 3 * (size_t)Nc;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Ac[(3 * (size_t)I + r) * N3 + 3 * (size_t)J + c] = acc[3 * r + c];
}

// B = L^-1 from trtri (lower triangle of the column-major A) as row-major FP32 B and B^T, row stride N3p.
__global__ void k_inverse_to_f32(const double* __restrict__ A, int N3, int N3p, float* __restrict__ B,
                                 float* __restrict__ BT) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)N3 * N3p) return;
  const int r = (int)(idx / N3p), c = (int)(idx % N3p);
  float lower = 0.0f, upper = 0.0f;
  if (c < N3) {
    if (c <= r) lower = (float)A[(size_t)c * N3 + r];  // B(r, c)
    if (c >= r) upper = (float)A[(size_t)r * N3 + c];  // B^T(r, c) = B(c, r)
  }
  B[idx] = lower;
  BT[idx] = upper;
}

__global__ void k_scale_diagonal(double* __restrict__ A, int n, double factor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(size_t)i * n + i] *= factor;
}

__global__ void k_identity(double* __restrict__ A, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(size_t)i * n + i] = 1.0;
}

// r_c[I] = sum over the children (i, k) of I of w_ik r_i: one wavefront per coarse node, lane j takes
// child j (a node has at most 49), then a fixed shuffle tree.
__global__ void __launch_bounds__(256) k_restrict(const uint32_t* __restrict__ child_ptr,
                                                  const uint32_t* __restrict__ child, const float* __restrict__ w,
                                                  const float4* __restrict__ r, int Nc, float* __restrict__ rc,
                                                  const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (I >= Nc) return;
  const uint32_t t0 = child_ptr[I], n = child_ptr[I + 1] - t0;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (uint32_t j = lane; j < n; j += 64) {
    const uint32_t s = child[t0 + j];
    const float ww = w[s];
    const float4 v = r[s >> 2];
    a0 = a0 + ww * v.x;
    a1 = a1 + ww * v.y;
    a2 = a2 + ww * v.z;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a0 = a0 + __shfl_down(a0, o, 64);
    a1 = a1 + __shfl_down(a1, o, 64);
    a2 = a2 + __shfl_down(a2, o, 64);
  }
  if (lane == 0) {
    rc[3 * I] = a0;
    rc[3 * I + 1] = a1;
    rc[3 * I + 2] = a2;
  }
}

// y = T x for a triangular T stored row-major in full (zeros outside the triangle): one wavefront per row
// over the float4 groups that meet the triangle, lane sums then a fixed shuffle tree.
template <bool LOWER>
__global__ void __launch_bounds__(256) k_tri_gemv(const float* __restrict__ T, const float* __restrict__ x,
                                                  int N3, int N3p, float* __restrict__ y, const PcgState* st,
                                                  int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= N3) return;
  const float4* m = reinterpret_cast<const float4*>(T + (size_t)row * N3p);
  const float4* v = reinterpret_cast<const float4*>(x);
  const int lo = LOWER ? 0 : row / 4, hi = LOWER ? row / 4 + 1 : N3p / 4;
  float acc = 0.0f;
  // eight independent loads in flight per lane before the sequential accumulation
  for (int j0 = lo + lane; j0 < hi; j0 += 8 * 64) {
    float4 a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = j0 + 64 * u;
      if (j < hi) {
        a[u] = m[j];
        b[u] = v[j];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + 64 * u < hi) acc = acc + (a[u].x * b[u].x + a[u].y * b[u].y + a[u].z * b[u].z + a[u].w * b[u].w);
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) acc = acc + __shfl_down(acc, s, 64);
  if (lane == 0) y[row] = acc;
}

// z = D^-1 r + sum_k w_ik y_c[parent(i, k)], partial (r.z, r.r); mode 2 (setup) also sets p = z.
__global__ void __launch_bounds__(256) k_prolong(const int32_t* __restrict__ parent, const float* __restrict__ w,
                                                 const float* __restrict__ yc, const float* __restrict__ dinv,
                                                 const float4* __restrict__ r, float4* __restrict__ z,
                                                 float4* __restrict__ p, float* __restrict__ partials, int R,
                                                 const PcgState* st, int guard) {
  __shared__ float2 lds[256];
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  float rz = 0.0f, rr = 0.0f;
  if (row < R) {
    const float4 rv = r[row];
    const float* d = dinv + 9 * row;
    float z0 = d[0] * rv.x + d[1] * rv.y + d[2] * rv.z;
    float z1 = d[3] * rv.x + d[4] * rv.y + d[5] * rv.z;
    float z2 = d[6] * rv.x + d[7] * rv.y + d[8] * rv.z;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int I = parent[4 * row + k];
      if (I >= 0) {
        const float ww = w[4 * row + k];
        z0 = z0 + ww * yc[3 * I];
        z1 = z1 + ww * yc[3 * I + 1];
        z2 = z2 + ww * yc[3 * I + 2];
      }
    }
    const float4 zz = make_float4(z0, z1, z2, 0.0f);
    z[row] = zz;
    if (guard == 2) p[row] = zz;
    rz = rv.x * z0 + rv.y * z1 + rv.z * z2;
    rr = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  }
  const float2 sum = tree256x2(rz, rr, lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sum.x;
    partials[2 * blockIdx.x + 1] = sum.y;
  }
}

// ---- host ------------------------------------------------------------------------------------
// Coarse layout of a grid sheet with `columns` vertices per side ((columns - 1) divisible by 4): parents,
// bilinear weights, children and the Galerkin term lists.
int two_level_create(Context& c, int columns) {
  if (columns < 9 || (columns - 1) % 4 != 0 || (size_t)columns * columns != (size_t)c.V) {
    set_error("two-level preconditioner: the sheet must be a grid with (columns - 1) divisible by 4");
    return DRAPE_E_INVALID;
  }
  TwoLevel* t = new TwoLevel();
  const int n = columns, nc = (n - 1) / 4 + 1;
  t->nc = nc;
  std::vector<int> row_of(c.V, -1);
  for (int row = 0; row < c.R; ++row) row_of[c.row_vertex[row]] = row;
  // coarse rows: free coarse vertices in vertex order
  std::vector<int> coarse_of(nc * nc, -1);
  int Nc = 0;
  for (int b = 0; b < nc; ++b)
    for (int a = 0; a < nc; ++a)
      if (row_of[(4 * b) * n + 4 * a] >= 0) coarse_of[b * nc + a] = Nc++;
  t->Nc = Nc;
  t->N3 = 3 * Nc;
  t->N3p = (t->N3 + 63) / 64 * 64;
  t->parent.assign(4 * (size_t)c.R, -1);
  t->bilinear.assign(4 * (size_t)c.R, 0.0f);
  for (int row = 0; row < c.R; ++row) {
    const int v = (int)c.row_vertex[row];
    const int x = v % n, y = v / n;
    const int a0 = std::min(x / 4, nc - 2), b0 = std::min(y / 4, nc - 2);
    const double fx = (x - 4.0 * a0) / 4.0, fy = (y - 4.0 * b0) / 4.0;
    const int ca[4] = {a0, a0 + 1, a0, a0 + 1}, cb[4] = {b0, b0, b0 + 1, b0 + 1};
    const double wt[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
    for (int k = 0; k < 4; ++k) {
      const int I = coarse_of[cb[k] * nc + ca[k]];
      if (I < 0 || wt[k] == 0.0) continue;
      t->parent[4 * row + k] = I;
      t->bilinear[4 * row + k] = (float)wt[k];
    }
  }
  // children of each coarse node, in fine-row order
  std::vector<std::vector<uint32_t>> kids(Nc);
  for (int row = 0; row < c.R; ++row)
    for (int k = 0; k < 4; ++k)
      if (t->parent[4 * row + k] >= 0) kids[t->parent[4 * row + k]].push_back((uint32_t)(4 * row + k));
  std::vector<uint32_t> child_ptr(1, 0), child;
  for (auto& l : kids) {
    child.insert(child.end(), l.begin(), l.end());
    child_ptr.push_back((uint32_t)child.size());
  }
  // Galerkin terms grouped by coarse block pair (I, J), each pair's terms in CSR block order
  std::map<uint64_t, std::vector<uint4>> pairs;
  for (int i = 0; i < c.R; ++i)
    for (uint32_t b = c.row_ptr_host[i]; b < c.row_ptr_host[i + 1]; ++b) {
      const int j = (int)c.col_idx_host[b];
      for (int k = 0; k < 4; ++k) {
        const int I = t->parent[4 * i + k];
        if (I < 0) continue;
        for (int l = 0; l < 4; ++l) {
          const int J = t->parent[4 * j + l];
          if (J < 0) continue;
          pairs[(uint64_t)I * Nc + J].push_back(make_uint4(b, 4 * i + k, 4 * j + l, 0));
        }
      }
    }
  std::vector<uint32_t> pair_ptr(1, 0), pair_key;
  std::vector<uint4> terms;
  for (auto& [key, list] : pairs) {
    pair_key.push_back((uint32_t)key);
    terms.insert(terms.end(), list.begin(), list.end());
    pair_ptr.push_back((uint32_t)terms.size());
  }
  t->npairs = (int)pair_key.size();
  hipStream_t s = c.stream;
  auto fail = [&](hipError_t e) {
    set_error(std::string("two-level setup: ") + hipGetErrorString(e));
    two_level_destroy(t);
    return DRAPE_E_DEVICE;
  };
  hipError_t e = hipSuccess;
#define TL_ALLOC(ptr, n) \
  if ((e = hipMalloc((void**)&(ptr), std::max<size_t>(1, (n)) * sizeof(*(ptr)))) != hipSuccess) return fail(e)
  TL_ALLOC(t->d_parent, t->parent.size());
  TL_ALLOC(t->d_w, t->bilinear.size());
  TL_ALLOC(t->d_child_ptr, child_ptr.size());
  TL_ALLOC(t->d_child, child.size());
  TL_ALLOC(t->d_pair_ptr, pair_ptr.size());
  TL_ALLOC(t->d_pair_key, pair_key.size());
  TL_ALLOC(t->d_pair_terms, terms.size());
  TL_ALLOC(t->d_Ac, (size_t)t->N3 * t->N3);
  TL_ALLOC(t->d_B, (size_t)t->N3 * t->N3p);
  TL_ALLOC(t->d_BT, (size_t)t->N3 * t->N3p);
  TL_ALLOC(t->d_t, (size_t)t->N3p);
  TL_ALLOC(t->d_rc, (size_t)t->N3p);
  TL_ALLOC(t->d_yc, (size_t)t->N3p);
  TL_ALLOC(t->d_info, 1);
#undef TL_ALLOC
  if ((e = hipHostMalloc((void**)&t->h_info, sizeof(int))) != hipSuccess) return fail(e);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  if ((e = up(t->d_parent, t->parent.data(), t->parent.size() * 4)) != hipSuccess ||
      (e = up(t->d_w, t->bilinear.data(), t->bilinear.size() * 4)) != hipSuccess ||
      (e = up(t->d_child_ptr, child_ptr.data(), child_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_child, child.data(), child.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_ptr, pair_ptr.data(), pair_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_key, pair_key.data(), pair_key.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_terms, terms.data(), terms.size() * sizeof(uint4))) != hipSuccess ||
      (e = hipMemsetAsync(t->d_rc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_yc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_t, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return fail(e);
  if (rocblas_create_handle(&t->handle) != rocblas_status_success ||
      rocblas_set_stream(t->handle, s) != rocblas_status_success) {
    set_error("two-level setup: rocblas handle");
    two_level_destroy(t);
    return DRAPE_E_DEVICE;
  }
  // load rocSOLVER's kernels for this size now rather than inside the first timed solve: factor the identity
  {
    const size_t n2 = (size_t)t->N3 * t->N3;
    if ((e = hipMemsetAsync(t->d_Ac, 0, n2 * sizeof(double), s)) != hipSuccess) return fail(e);
    k_identity<<<blocks_for(t->N3), 256, 0, s>>>(t->d_Ac, t->N3);
    if (rocsolver_dpotrf(t->handle, rocblas_fill_lower, t->N3, t->d_Ac, t->N3, t->d_info) != rocblas_status_success ||
        rocsolver_dtrtri(t->handle, rocblas_fill_lower, rocblas_diagonal_non_unit, t->N3, t->d_Ac, t->N3,
                         t->d_info) != rocblas_status_success) {
      set_error("two-level setup: rocsolver warm-up");
      two_level_destroy(t);
      return DRAPE_E_DEVICE;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  }
  if (c.two_level) two_level_destroy(c.two_level);
  c.two_level = t;
  return DRAPE_OK;
}

int two_level_set_weights(Context& c, const float* w) {
  TwoLevel* t = c.two_level;
  if (!t) {
    set_error("no two-level layout on this context");
    return DRAPE_E_INVALID;
  }
  std::vector<float> v(4 * (size_t)c.R);
  for (size_t i = 0; i < v.size(); ++i) v[i] = t->parent[i] >= 0 ? (w ? w[i] : t->bilinear[i]) : 0.0f;
  HIP_TRY(hipMemcpyAsync(t->d_w, v.data(), v.size() * 4, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  t->factored = ~0ull;
  return DRAPE_OK;
}

// A_c = P^T A P of the assembled system, its Cholesky factor and the factor's inverse, unless the current
// factorisation may serve this system (same system, or coarse_reuse with unchanged mu and step size and no
// slowdown). *ok is false when A_c is not positive definite.
int two_level_prepare(Context& c, bool* ok) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  const bool fresh = t->factored == c.system_version;
  const bool reusable = c.rules.coarse_reuse && t->factored != ~0ull && t->ok && !t->refresh &&
                        (c.rules.coarse_reuse == 2 || t->factored_shift == c.shift_version);
  if (fresh || reusable) {
    *ok = t->ok;
    return DRAPE_OK;
  }
  t->factored = c.system_version;
  t->factored_shift = c.shift_version;
  t->refresh = false;
  t->solves = 0;
  t->ok = false;
  *ok = false;
  const auto started = std::chrono::steady_clock::now();
  struct Clock {
    TwoLevel* t;
    std::chrono::steady_clock::time_point t0;
    ~Clock() { t->factor_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
  } clock{t, started};
  const size_t n2 = (size_t)t->N3 * t->N3;
  HIP_TRY(hipMemsetAsync(t->d_Ac, 0, n2 * sizeof(double), s));
  k_galerkin<<<blocks_for(t->npairs), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, t->d_w, c.sell_val,
                                                   c.block_dst, t->npairs, t->Nc, t->d_Ac);
  if (c.rules.coarse_shift != 0.0)
    k_scale_diagonal<<<blocks_for(t->N3), 256, 0, s>>>(t->d_Ac, t->N3, 1.0 + c.rules.coarse_shift);
  HIP_TRY(hipGetLastError());
  if (rocsolver_dpotrf(t->handle, rocblas_fill_lower, t->N3, t->d_Ac, t->N3, t->d_info) != rocblas_status_success) {
    set_error("rocsolver_dpotrf failed");
    return DRAPE_E_DEVICE;
  }
  HIP_TRY(hipMemcpyAsync(t->h_info, t->d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  t->factorisations++;
  if (*t->h_info != 0) {
    t->failures++;
    return DRAPE_OK;
  }
  if (rocsolver_dtrtri(t->handle, rocblas_fill_lower, rocblas_diagonal_non_unit, t->N3, t->d_Ac, t->N3, t->d_info) !=
      rocblas_status_success) {
    set_error("rocsolver_dtrtri failed");
    return DRAPE_E_DEVICE;
  }
  const size_t nm = (size_t)t->N3 * t->N3p;
  k_inverse_to_f32<<<(unsigned)((nm + 255) / 256), 256, 0, s>>>(t->d_Ac, t->N3, t->N3p, t->d_B, t->d_BT);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(t->h_info, t->d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  t->ok = *t->h_info == 0;
  if (!t->ok) t->failures++;
  *ok = t->ok;
  return DRAPE_OK;
}

// The Galerkin operator of the assembled system, row-major [3Nc x 3Nc], into host memory. The current
// factorisation lives in B and B^T and is unaffected.
int two_level_galerkin(Context& c, double* out) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  const size_t n2 = (size_t)t->N3 * t->N3;
  HIP_TRY(hipMemsetAsync(t->d_Ac, 0, n2 * sizeof(double), s));
  k_galerkin<<<blocks_for(t->npairs), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, t->d_w, c.sell_val,
                                                   c.block_dst, t->npairs, t->Nc, t->d_Ac);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, t->d_Ac, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

// After a two-level solve: the first solve after a factorisation sets the iterations-per-decade baseline,
// a later one at more than twice that rate, or a breakdown (iterations < 0), asks for a new factorisation.
void two_level_observe(Context& c, int iterations, double relative_residual) {
  TwoLevel* t = c.two_level;
  if (iterations < 0) {
    t->refresh = true;
    return;
  }
  const double decades = relative_residual > 0 && relative_residual < 1 ? std::max(0.5, -std::log10(relative_residual)) : 0.5;
  const double rate = iterations / decades;
  if (t->solves++ == 0) t->base_rate = rate;
  else if (rate > 2.0 * t->base_rate + 4.0) t->refresh = true;
}

// z = D^-1 r + P B^T B P^T r for the current r (guard: 0 running, 1 replacement gate, 2 setup with p = z).
int two_level_apply(Context& c, float* partials, int guard) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  k_restrict<<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, c.r, t->Nc, t->d_rc, c.state, guard);
  k_tri_gemv<true><<<(t->N3 + 3) / 4, 256, 0, s>>>(t->d_B, t->d_rc, t->N3, t->N3p, t->d_t, c.state, guard);
  k_tri_gemv<false><<<(t->N3 + 3) / 4, 256, 0, s>>>(t->d_BT, t->d_t, t->N3, t->N3p, t->d_yc, c.state, guard);
  k_prolong<<<c.G, 256, 0, s>>>(t->d_parent, t->d_w, t->d_yc, c.dinv, c.r, c.z, c.p, partials, c.R, c.state, guard);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int two_level_info(const Context& c, int32_t* coarse_rows, int32_t* parent, float* weights, int64_t* factorisations,
                   int64_t* failures, double* factor_ms) {
  const TwoLevel* t = c.two_level;
  if (!t) {
    set_error("no two-level layout on this context");
    return DRAPE_E_INVALID;
  }
  if (coarse_rows) *coarse_rows = t->Nc;
  if (parent) std::copy(t->parent.begin(), t->parent.end(), parent);
  if (weights) std::copy(t->bilinear.begin(), t->bilinear.end(), weights);
  if (factorisations) *factorisations = t->factorisations;
  if (failures) *failures = t->failures;
  if (factor_ms) *factor_ms = t->factor_ms;
  return DRAPE_OK;
}

}  // namespace drape
