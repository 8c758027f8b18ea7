// Prolongation of a coarse equilibrium to the next grid as the start of a PTC solve (e7_prolong.rs:9-47):
// fine vertex (i, j) of the (2n - 1)^2 grid is the mean of the coarse corners (i/2, j/2), (ceil(i/2), j/2),
// (i/2, ceil(j/2)), (ceil(i/2), ceil(j/2)); then a vertex closer than dhat to a collider moves along that
// collider's normal to d = dhat (one pass over the colliders in order), and a vertex at or inside a collider
// is flagged infeasible. One thread per fine vertex, f64 arithmetic without contraction (the reference's f64
// operation order), f32 out.
#include "../internal.h"

namespace drape {

namespace {

__global__ void __launch_bounds__(256) k_prolong_state(const float* __restrict__ coarse, int nc, int nf,
                                                       const double* __restrict__ cols, int K, double dhat,
                                                       float* __restrict__ fine, int32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nf * nf) return;
  const int i = v % nf, j = v / nf;
  const int ia[4] = {i / 2, (i + 1) / 2, i / 2, (i + 1) / 2};
  const int ja[4] = {j / 2, j / 2, (j + 1) / 2, (j + 1) / 2};
  double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float* q = coarse + 3 * (ja[c] * nc + ia[c]);
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
  for (int k = 0; k < 3; ++k) fine[3 * v + k] = (float)p[k];
  flags[v] = flag;
}

}  // namespace

}  // namespace drape

using namespace drape;

extern "C" int drape_prolong_state(int32_t coarse_columns, const float* coarse, const double* colliders,
                                   int32_t collider_doubles, float* fine, int32_t* moved, int32_t* infeasible) {
  if (coarse_columns < 2 || coarse_columns > 8193 || !coarse || !fine) {
    set_error("prolong: coarse_columns must be in [2, 8193] and the state arrays non-null");
    return DRAPE_E_INVALID;
  }
  int K = 0;
  double dhat = 0.0;
  if (colliders && collider_doubles > 0) {
    K = (int)colliders[0];
    if (collider_doubles < 4 || K < 0 || collider_doubles != 4 + 8 * K) {
      set_error("prolong: colliders must be [count, dhat, kappa, flag] + 8 doubles per collider");
      return DRAPE_E_INVALID;
    }
    dhat = colliders[1];
    for (int c = 0; c < K; ++c) {
      const double kind = colliders[4 + 8 * c];
      if (kind != 1.0 && kind != 2.0) {
        set_error("prolong: collider kind must be 1 (plane) or 2 (sphere)");
        return DRAPE_E_INVALID;
      }
    }
  }
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
