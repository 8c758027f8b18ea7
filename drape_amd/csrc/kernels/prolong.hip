// Prolongation of a coarse equilibrium to the next grid as the start of a PTC solve (e7_prolong.rs:9-47):
// fine vertex (i, j) of the (2n - 1)^2 grid is the mean of the coarse corners (i/2, j/2), (ceil(i/2), j/2),
// (i/2, ceil(j/2)), (ceil(i/2), ceil(j/2)); then a vertex closer than dhat to a collider moves along that
// collider's normal to d = dhat (one pass over the colliders in order), and a vertex at or inside a collider
// is flagged infeasible. One thread per fine vertex, f64 arithmetic without contraction (the reference's f64
// operation order), f32 out.
#include <algorithm>
#include <mutex>

#include "../internal.h"

namespace drape {

namespace {

// Fine vertex v of the (2 nc - 1)^2 grid: its prolonged position into fine[3v..3v+2] and its flags (1 moved to
// d = dhat, 2 at or inside a collider).
__device__ inline int32_t prolong_vertex(const float* __restrict__ coarse, int nc, int nf, const double* __restrict__ cols,
                                         int K, double dhat, int v, float* __restrict__ fine) {
#pragma clang fp contract(off)
  const int i = v % nf, j = v / nf;
  const int ia[4] = {i / 2, (i + 1) / 2, i / 2, (i + 1) / 2};
  const int ja[4] = {j / 2, j / 2, (j + 1) / 2, (j + 1) / 2};
  double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float* q = coarse + 3 * ((size_t)ja[c] * nc + ia[c]);
#pragma unroll
    for (int k = 0; k < 3; ++k) p[k] = p[k] + (double)q[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = p[k] / 4.0;
  const double b0 = p[0], b1 = p[1], b2 = p[2];
  for (int c = 0; c < K; ++c) {
    const double* col = cols + 4 + 8 * c;
    const bool plane = col[0] == 1.0;
    double d, n[3];
    if (plane) {
      d = col[1] * p[0] + col[2] * p[1] + col[3] * p[2] - col[4];
      n[0] = col[1]; n[1] = col[2]; n[2] = col[3];
    } else {
      const double q0 = p[0] - col[1], q1 = p[1] - col[2], q2 = p[2] - col[3];
      const double len = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
      d = len - col[4];
      n[0] = q0 / len; n[1] = q1 / len; n[2] = q2 / len;
    }
    if (d < dhat) {
#pragma unroll
      for (int k = 0; k < 3; ++k) p[k] = p[k] + (dhat - d) * n[k];
    }
  }
  int32_t flag = (p[0] != b0 || p[1] != b1 || p[2] != b2) ? 1 : 0;
  double dmin = INFINITY;
  for (int c = 0; c < K; ++c) {
    const double* col = cols + 4 + 8 * c;
    double d;
    if (col[0] == 1.0) {
      d = col[1] * p[0] + col[2] * p[1] + col[3] * p[2] - col[4];
    } else {
      const double q0 = p[0] - col[1], q1 = p[1] - col[2], q2 = p[2] - col[3];
      d = sqrt(q0 * q0 + q1 * q1 + q2 * q2) - col[4];
    }
    dmin = fmin(dmin, d);
  }
  if (K > 0 && dmin <= 0.0) flag |= 2;
#pragma unroll
  for (int k = 0; k < 3; ++k) fine[3 * (size_t)v + k] = (float)p[k];
  return flag;
}

__global__ void __launch_bounds__(256) k_prolong_state(const float* __restrict__ coarse, int nc, int nf,
                                                       const double* __restrict__ cols, int K, double dhat,
                                                       float* __restrict__ fine, int32_t* __restrict__ flags) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nf * nf) return;
  flags[v] = prolong_vertex(coarse, nc, nf, cols, K, dhat, v, fine);
}

// The same positions with the flags counted instead of stored: a grid-stride pass whose workgroups each write
// their (moved, infeasible) counts to counts[2 * blockIdx.x ..] (every slot of the grid is written).
__global__ void __launch_bounds__(256) k_prolong_state_counted(const float* __restrict__ coarse, int nc, int nf,
                                                               const double* __restrict__ cols, int K, double dhat,
                                                               float* __restrict__ fine, int32_t* __restrict__ counts) {
  __shared__ int32_t lds[2][256];
  int32_t moved = 0, bad = 0;
  const int n = nf * nf;
  for (int v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
    const int32_t f = prolong_vertex(coarse, nc, nf, cols, K, dhat, v, fine);
    moved += f & 1;
    bad += (f >> 1) & 1;
  }
  lds[0][threadIdx.x] = moved;
  lds[1][threadIdx.x] = bad;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + st];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[2 * blockIdx.x] = lds[0][0];
    counts[2 * blockIdx.x + 1] = lds[1][0];
  }
}

// Sums the `groups` workgroup counts in order into total[0..1].
__global__ void __launch_bounds__(256) k_prolong_counts_total(const int32_t* __restrict__ counts, int groups,
                                                              int32_t* __restrict__ total) {
  __shared__ int32_t lds[2][256];
  int32_t a = 0, b = 0;
  for (int g = threadIdx.x; g < groups; g += 256) {
    a += counts[2 * g];
    b += counts[2 * g + 1];
  }
  lds[0][threadIdx.x] = a;
  lds[1][threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + st];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    total[0] = lds[0][0];
    total[1] = lds[1][0];
  }
}

}  // namespace

// ---- the allocation-free variant ------------------------------------------------------------------------
// Its only device memory beyond the caller's arrays is this file-scope array: the colliders staged from the
// host ([4 + 8 * 16] doubles), the per-workgroup (moved, infeasible) counts of the fixed 1024-workgroup grid
// and their totals. Every GPU holds its own copy until the process ends; every call on a GPU shares it, so
// calls that overlap on one GPU (other streams, other host threads) race. Each call writes every slot it
// reads (the colliders, all 1024 count pairs, the totals) before reading it, so nothing needs clearing.
constexpr int NOALLOC_COLLIDERS = 16;
constexpr int NOALLOC_GROUPS = 1024;
constexpr size_t SCRATCH_COLLIDERS = 1280;  // bytes: 132 doubles, padded to 256
constexpr size_t SCRATCH_COUNTS = 2 * NOALLOC_GROUPS * sizeof(int32_t);
constexpr size_t SCRATCH_BYTES = SCRATCH_COLLIDERS + SCRATCH_COUNTS + 256;
static_assert((4 + 8 * NOALLOC_COLLIDERS) * sizeof(double) <= SCRATCH_COLLIDERS, "collider staging");
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// The device address of `scratch` on the current device, looked up once per device with that device current.
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch[MAX_DEVICES] = {};

static int scratch_address(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) {
    set_error("prolong (no allocation): device index beyond " + std::to_string(MAX_DEVICES));
    return DRAPE_E_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("prolong (no allocation): hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return DRAPE_E_DEVICE;
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return DRAPE_OK;
}

// A caller's array must be device memory of the current device (a host pointer would fault the kernel).
static int check_device_array(const void* p, const char* name) {
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (e != hipSuccess || a.type != hipMemoryTypeDevice || a.device != dev) {
    (void)hipGetLastError();
    set_error(std::string("prolong (no allocation): ") + name + " must be device memory of the current device");
    return DRAPE_E_INVALID;
  }
  return DRAPE_OK;
}

// Argument checks shared by both entry points: K and dhat of a valid collider array.
static int parse_colliders(int32_t coarse_columns, const double* colliders, int32_t collider_doubles, int* K,
                           double* dhat) {
  if (coarse_columns < 2 || coarse_columns > 8193) {
    set_error("prolong: coarse_columns must be in [2, 8193]");
    return DRAPE_E_INVALID;
  }
  *K = 0;
  *dhat = 0.0;
  if (colliders && collider_doubles > 0) {
    *K = (int)colliders[0];
    if (collider_doubles < 4 || *K < 0 || collider_doubles != 4 + 8 * *K) {
      set_error("prolong: colliders must be [count, dhat, kappa, flag] + 8 doubles per collider");
      return DRAPE_E_INVALID;
    }
    *dhat = colliders[1];
    for (int c = 0; c < *K; ++c) {
      const double kind = colliders[4 + 8 * c];
      if (kind != 1.0 && kind != 2.0) {
        set_error("prolong: collider kind must be 1 (plane) or 2 (sphere)");
        return DRAPE_E_INVALID;
      }
    }
  }
  return DRAPE_OK;
}

}  // namespace drape

using namespace drape;

extern "C" int drape_prolong_state(int32_t coarse_columns, const float* coarse, const double* colliders,
                                   int32_t collider_doubles, float* fine, int32_t* moved, int32_t* infeasible) {
  if (!coarse || !fine) {
    set_error("prolong: the state arrays must be non-null");
    return DRAPE_E_INVALID;
  }
  int K = 0;
  double dhat = 0.0;
  TRY_RC(parse_colliders(coarse_columns, colliders, collider_doubles, &K, &dhat));
  const int nc = coarse_columns, nf = 2 * (nc - 1) + 1;
  const size_t vc = (size_t)nc * nc, vf = (size_t)nf * nf;
  float *d_coarse = nullptr, *d_fine = nullptr;
  double* d_cols = nullptr;
  int32_t* d_flags = nullptr;
  std::vector<int32_t> flags(vf);
  int rc = DRAPE_OK;
  auto run = [&]() -> int {
    HIP_TRY(hipMalloc(&d_coarse, 3 * vc * sizeof(float)));
    HIP_TRY(hipMalloc(&d_fine, 3 * vf * sizeof(float)));
    HIP_TRY(hipMalloc(&d_flags, vf * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&d_cols, (4 + 8 * (size_t)K) * sizeof(double)));
    HIP_TRY(hipMemcpy(d_coarse, coarse, 3 * vc * sizeof(float), hipMemcpyHostToDevice));
    if (K > 0) HIP_TRY(hipMemcpy(d_cols, colliders, (4 + 8 * (size_t)K) * sizeof(double), hipMemcpyHostToDevice));
    k_prolong_state<<<blocks_for((int)vf), 256>>>(d_coarse, nc, nf, d_cols, K, dhat, d_fine, d_flags);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(fine, d_fine, 3 * vf * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(flags.data(), d_flags, vf * sizeof(int32_t), hipMemcpyDeviceToHost));
    return DRAPE_OK;
  };
  rc = run();
  if (d_coarse) (void)hipFree(d_coarse);
  if (d_fine) (void)hipFree(d_fine);
  if (d_flags) (void)hipFree(d_flags);
  if (d_cols) (void)hipFree(d_cols);
  if (rc != DRAPE_OK) return rc;
  int32_t m = 0, bad = 0;
  for (int32_t f : flags) {
    m += f & 1;
    bad += (f >> 1) & 1;
  }
  if (moved) *moved = m;
  if (infeasible) *infeasible = bad;
  return DRAPE_OK;
}

extern "C" int drape_prolong_state_noalloc(int32_t coarse_columns, const float* coarse, const double* colliders,
                                           int32_t collider_doubles, float* fine, int32_t* moved,
                                           int32_t* infeasible, void* stream) {
  if (!coarse || !fine) {
    set_error("prolong (no allocation): the state arrays must be non-null");
    return DRAPE_E_INVALID;
  }
  int K = 0;
  double dhat = 0.0;
  TRY_RC(parse_colliders(coarse_columns, colliders, collider_doubles, &K, &dhat));
  if (K > NOALLOC_COLLIDERS) {
    set_error("prolong (no allocation): at most " + std::to_string(NOALLOC_COLLIDERS) +
              " colliders fit its scratch array, got " + std::to_string(K));
    return DRAPE_E_INVALID;
  }
  TRY_RC(check_device_array(coarse, "coarse"));
  TRY_RC(check_device_array(fine, "fine"));
  void* base = nullptr;
  TRY_RC(scratch_address(&base));
  double* d_cols = reinterpret_cast<double*>(base);
  int32_t* d_counts = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(base) + SCRATCH_COLLIDERS);
  int32_t* d_total = d_counts + 2 * NOALLOC_GROUPS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nc = coarse_columns, nf = 2 * (nc - 1) + 1;
  const int groups = std::min(NOALLOC_GROUPS, blocks_for(nf * nf));
  if (K > 0) HIP_TRY(hipMemcpyAsync(d_cols, colliders, (4 + 8 * (size_t)K) * sizeof(double), hipMemcpyHostToDevice, s));
  k_prolong_state_counted<<<groups, 256, 0, s>>>(coarse, nc, nf, d_cols, K, dhat, fine, d_counts);
  k_prolong_counts_total<<<1, 256, 0, s>>>(d_counts, groups, d_total);
  HIP_TRY(hipGetLastError());
  int32_t total[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(total, d_total, sizeof(total), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (moved) *moved = total[0];
  if (infeasible) *infeasible = total[1];
  return DRAPE_OK;
}
