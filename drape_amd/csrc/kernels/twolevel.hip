// Two-level preconditioner for the Newton systems: block-Jacobi plus a Galerkin coarse correction,
//     M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T A P,
// additive and symmetric, so PCG keeps its SPD preconditioner whenever A_c is SPD. There is no reference
// counterpart: the reference's PCG is block-Jacobi only (solver.ts:109-191) and names a hierarchical
// preconditioner as its next step. The coarse nodes are every fourth vertex of the sheet grid in each
// direction (1/16 of the vertices); P = W (x) I3 with four parent slots per fine row whose weights are
// bilinear by default or set by the learned prolongation network.
//
// Device work per factorisation, all hand-written and deterministic:
//  - the Galerkin product into banded tile storage (one thread per coarse 3x3 block pair, FP64, its terms
//    in a fixed order); coarse vertices are numbered row by row, so A_c is banded (half-bandwidth about
//    6 coarse columns of unknowns) and its Cholesky factor keeps that band;
//  - a blocked band Cholesky A_c = L L^T in FP64 on 64 x 64 tiles: per tile column, one workgroup factors
//    the diagonal tile in LDS by 16-column panels and inverts it blockwise, workgroups of 16 rows form the
//    sub-diagonal factor tiles as A_ik L_kk^-T, and workgroups of 16 rows apply the rank-64 trailing update;
//  - B = L^-1 (dense lower triangular) by blocked forward substitution, one workgroup per 16 columns, with
//    FP32 copies of B and B^T written directly.
// The coarse solve is applied as B^T (B r_c): a Gram product, so the FP32 operator stays positive
// semidefinite whatever the rounding of B (an FP32 copy of A_c^-1 itself turns indefinite once cond(A_c)
// nears 1e7). Per PCG iteration: restriction (one wavefront per coarse node), two triangular GEMVs (one
// wavefront per row, float4 loads, fixed shuffle reduction) and the prolongation fused with the block-Jacobi
// term and the (r.z, r.r) partial sums. The factorisation is kept across Newton systems while mu and the
// step size are unchanged and PCG stays within twice the iterations per decade of the first solve after it:
// any fixed SPD operator is a valid preconditioner, and a stale Galerkin operator of a nearby system is one.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <string>
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
  double factored_h = 0.0;      // step size of M/h^2 at the current factorisation
  bool ok = false;            // the last factorisation succeeded
  bool refresh = false;       // PCG slowed down: refactor at the next new system
  double base_rate = 0.0;     // iterations per decade of the first solve after the factorisation
  int solves = 0;             // solves since the factorisation
  std::vector<int32_t> parent;   // [4R] coarse row of each parent slot, -1 when empty
  std::vector<float> bilinear;   // [4R] default weights
  int32_t* d_parent = nullptr;
  float* d_w = nullptr;          // [4R] current weights
  float* d_wbil = nullptr;       // [4R] bilinear weights
  float2* d_fxy = nullptr;       // [R] position of each row inside its coarse cell
  float* d_theta = nullptr;      // learned-prolongation parameters (null: fixed weights)
  float* d_gnn = nullptr;        // its scratch: scale [R], features [13R], edge [B], activations [32R] x 3
  uint32_t* d_child_ptr = nullptr;  // [Nc+1]
  uint32_t* d_child = nullptr;      // 4 i + k of each child slot, in fine-row order
  uint32_t* d_pair_ptr = nullptr;   // [npairs+1]
  uint32_t* d_pair_key = nullptr;   // [npairs] I * Nc + J
  uint4* d_pair_terms = nullptr;    // (block b, slot ik, slot jl, 0)
  int nt = 0;                       // 64-row tiles of the coarse matrix (padded with identity rows)
  int kb = 0;                       // sub-diagonal tiles in the band
  double* d_band = nullptr;         // [nt][kb + 1][64 x 64]: tile (k + d, k) of A_c, then of L, row-major
  double* d_linv = nullptr;         // [nt][64 x 64]: inverse of each diagonal tile of L
  double* d_lpanel = nullptr;       // [nt][kb][64 x 64]: tile (k + d, k) of L, d = 1..kb
  double* d_xring = nullptr;        // [kb + 1][64][N3p]: the last tile rows of L^-1 in FP64 (inverse sweep)
  hipStream_t inv_stream = nullptr; // the inverse sweep, behind each panel of the factorisation chain
  hipEvent_t ev_panel = nullptr, ev_inverse = nullptr;
  double* d_Ac = nullptr;           // [N3 * N3] dense A_c for the check entry point (allocated on demand)
  float* d_B = nullptr;             // [N3 * N3p] row-major B = L^-1 (zero above the diagonal)
  float* d_BT = nullptr;            // [N3 * N3p] row-major B^T (zero below the diagonal)
  _Float16* d_Bh = nullptr;         // [N3 * N3p] FP16 B~ = S^-1 B, S = diag(B) (zero above the diagonal)
  _Float16* d_BTh = nullptr;        // [N3 * N3p] its transpose (the same values)
  bool half_valid = false;          // the FP16 copies belong to the current factorisation
  // coarse_reuse 4: the factorisation of the step size used before the current one, kept for when the
  // step size returns to it (PTC alternates between neighbouring step sizes)
  struct Saved {
    bool valid = false, half_valid = false;
    uint64_t factored = 0, factored_shift = 0;
    double factored_h = 0.0, base_rate = 0.0;
    int solves = 0;
    double* linv = nullptr;
    float *B = nullptr, *BT = nullptr;
    _Float16 *Bh = nullptr, *BTh = nullptr;
  } saved;
  int64_t restores = 0;             // factorisations avoided by returning to the saved one
  // coarse_async: a factorisation running on `side` into these buffers while the solves keep the current one
  struct Pending {
    bool active = false, half_valid = false;
    bool deferred = false;          // decided, waiting for the solve's first launches to be queued
    bool galerkin = false;          // the next assembly waits for `read`: the product still reads the matrix
    uint64_t factored = 0, factored_shift = 0;
    double factored_h = 0.0;
    float* w = nullptr;
    double* linv = nullptr;
    float *B = nullptr, *BT = nullptr;
    _Float16 *Bh = nullptr, *BTh = nullptr;
    int* h_info = nullptr;          // pinned
    hipStream_t side = nullptr;
    hipEvent_t ready = nullptr, read = nullptr, done = nullptr;
    hipEvent_t began = nullptr, ended = nullptr;  // timing events around the chain on the side stream
  } pending;
  float* d_t = nullptr;             // [N3p] B r_c
  float* d_rc = nullptr;            // [N3p]
  float* d_yc = nullptr;            // [N3p]
  int* d_info = nullptr;
  int* h_info = nullptr;
  int64_t factorisations = 0, failures = 0;
  double factor_ms = 0.0;     // time of all factorisations (Galerkin product, Cholesky, inverse): wall time of the
                              // synchronous ones, device time (events on the side stream) of the side-stream ones
};

void two_level_destroy(TwoLevel* t) {
  if (!t) return;
  TwoLevel::Pending& q = t->pending;
  if (q.side) (void)hipStreamSynchronize(q.side);
  if (t->inv_stream) (void)hipStreamSynchronize(t->inv_stream);
  for (hipEvent_t e : {t->ev_panel, t->ev_inverse})
    if (e) (void)hipEventDestroy(e);
  if (t->inv_stream) (void)hipStreamDestroy(t->inv_stream);
  if (t->d_xring) (void)hipFree(t->d_xring);
  for (void* p : {(void*)q.w, (void*)q.linv, (void*)q.B, (void*)q.BT, (void*)q.Bh, (void*)q.BTh})
    if (p) (void)hipFree(p);
  if (q.h_info) (void)hipHostFree(q.h_info);
  for (hipEvent_t e : {q.ready, q.read, q.done, q.began, q.ended})
    if (e) (void)hipEventDestroy(e);
  if (q.side) (void)hipStreamDestroy(q.side);
  for (void* p : {(void*)t->d_parent, (void*)t->d_w, (void*)t->d_wbil, (void*)t->d_fxy, (void*)t->d_theta,
                  (void*)t->d_gnn, (void*)t->d_child_ptr, (void*)t->d_child, (void*)t->d_pair_ptr,
                  (void*)t->d_pair_key, (void*)t->d_pair_terms, (void*)t->d_band, (void*)t->d_linv, (void*)t->d_lpanel, (void*)t->d_Ac, (void*)t->d_B, (void*)t->d_BT, (void*)t->d_Bh, (void*)t->d_BTh,
                  (void*)t->saved.linv, (void*)t->saved.B, (void*)t->saved.BT, (void*)t->saved.Bh, (void*)t->saved.BTh,
                  (void*)t->d_t, (void*)t->d_rc, (void*)t->d_yc, (void*)t->d_info})
    if (p) (void)hipFree(p);
  if (t->h_info) (void)hipHostFree(t->h_info);
  delete t;
}

// ---- kernels ----------------------------------------------------------------------------------
// ---- banded FP64 Cholesky and inverse on 64 x 64 tiles ----------------------------------------------
constexpr int TILE = 64;
constexpr int MAX_KB = 8;   // sub-diagonal tiles: coarse grids up to 83 vertices a side
constexpr int LSTR = 65;    // LDS row stride of column-accessed tiles (bank spread)

__device__ inline double* band_tile(double* band, int kb, int k, int d) {
  return band + (((size_t)k * (kb + 1) + d) << 12);
}

// Copies a ROWS x 64 row-major FP64 tile into LDS rows of stride LSTR: every load of the thread is issued
// before its first store, so the tile costs one memory latency instead of one per element.
template <int ROWS>
__device__ inline void tile_to_lds(const double* __restrict__ src, double* dst) {
  constexpr int PER = ROWS * TILE / 256;
  double v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) v[i] = src[threadIdx.x + 256 * i];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + 256 * i;
    dst[(e >> 6) * LSTR + (e & 63)] = v[i];
  }
}

// The coarse block pair (I, J) into the band, entries on or below the diagonal: one wavefront per pair,
// lane l summing the pair's terms l, l + 64, ... for all nine entries, then a fixed-order wavefront tree, so
// the sum is reproducible. (One thread per entry walking a pair's few hundred terms in sequence was
// latency-bound: 248 us per product at n129.)
__global__ void __launch_bounds__(256) k_galerkin_band(const uint32_t* __restrict__ pair_ptr,
                                                       const uint32_t* __restrict__ pair_key,
                                                       const uint4* __restrict__ terms, const float* __restrict__ w,
                                                       const float* __restrict__ sell_val,
                                                       const uint32_t* __restrict__ block_dst, int npairs, int Nc,
                                                       int kb, double* __restrict__ band) {
  const int p = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= npairs) return;  // uniform per wavefront
  const uint32_t key = pair_key[p];
  const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  const uint32_t t1 = pair_ptr[p + 1];
  for (uint32_t t = pair_ptr[p] + lane; t < t1; t += 64) {
    const uint4 e = terms[t];
    const double ww = (double)w[e.y] * (double)w[e.z];
    const float* a = sell_val + block_dst[e.x];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] += ww * (double)a[64 * q];
  }
#pragma unroll
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_down(acc[q], o, 64);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int R = 3 * I + q / 3, C = 3 * J + q % 3;
      if (R < C) continue;
      const int k = C >> 6, d = (R >> 6) - k;
      band_tile(band, kb, k, d)[((R & 63) << 6) + (C & 63)] = acc[q];
    }
  }
}

// Diagonal of A_c scaled by `factor`; the padding rows past N3 get a unit diagonal.
__global__ void k_band_diagonal(double* __restrict__ band, int kb, int nt, int N3, double factor) {
  const int R = blockIdx.x * blockDim.x + threadIdx.x;
  if (R >= nt * TILE) return;
  double* a = band_tile(band, kb, R >> 6, 0) + ((R & 63) << 6) + (R & 63);
  *a = R < N3 ? *a * factor : 1.0;
}

// LDS traffic between the lanes of one wavefront: LDS instructions of a wavefront execute in order, so a
// wavefront-scope fence (no instruction, an ordering constraint on the compiler) keeps a store before the
// loads that read it back.
__device__ inline void wave_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// 1/sqrt(x) to double precision: the hardware estimate refined by two Newton steps.
__device__ inline double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  y = y * (1.5 - 0.5 * x * y * y);
  return y;
}

// Diagonal tile of tile column k, blocked by 16-column panels. Wavefront 0 factors each panel four columns
// at a time: the 4 x 4 diagonal block is broadcast through LDS and factored redundantly by every lane, each
// lane (= row) solves its four factor entries and broadcasts them, and the lanes apply the rank-4 update to
// the rest of their panel row. All four wavefronts then apply the panel's rank-16 update to the trailing
// lower triangle. The tile's inverse follows blockwise: the four 16 x 16 diagonal blocks in parallel, then
// the off-diagonal blocks by distance from the diagonal, X_ab = -X_aa sum_c L_ac X_cb. A pivot that is not
// positive (or not finite) records 64 k + column + 1 in *info and stops every later kernel.
__global__ void __launch_bounds__(256) k_chol_panel(double* __restrict__ band, double* __restrict__ linv, int k,
                                                    int kb, int* info) {
  __shared__ double Ls[TILE * LSTR];
  __shared__ double Xs[TILE * LSTR];
  __shared__ double Ys[3 * 16 * 17];
  __shared__ double blk[16];
  __shared__ double colbuf[4 * TILE];
  __shared__ double rdiag[TILE];
  __shared__ int fail;
  if (*info != 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* tile = band_tile(band, kb, k, 0);
  tile_to_lds<TILE>(tile, Ls);
  for (int e = tid; e < TILE * TILE; e += 256) Xs[(e >> 6) * LSTR + (e & 63)] = 0.0;
  if (tid == 0) fail = 0;
  __syncthreads();
  for (int b = 0; b < 4; ++b) {
    const int c0 = 16 * b;
    if (wave == 0) {
      const int row = c0 + lane;
      const bool act = row < TILE;
      double p[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) p[j] = act ? Ls[row * LSTR + c0 + j] : 0.0;
      int bad = 0;
#pragma unroll
      for (int q = 0; q < 16; q += 4) {
        if (bad == 0) {
          // the 4 x 4 diagonal block (rows and columns q .. q + 3 of the panel), lower triangle
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (lane == q + u) {
#pragma unroll
              for (int v = 0; v < 4; ++v) blk[u * 4 + v] = p[q + v];
            }
          wave_order();
          double d[4][4];
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) d[u][v] = blk[u * 4 + v];
          // factor it redundantly: l[u][v], u >= v, and the reciprocal diagonal
          double l[4][4] = {}, rd[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double piv = d[v][v];
#pragma unroll
            for (int m = 0; m < v; ++m) piv -= l[v][m] * l[v][m];
            if (!(piv > 0.0) || !(piv < INFINITY)) {
              if (bad == 0) bad = c0 + q + v + 1;
              piv = 1.0;
            }
            rd[v] = rsqrt_f64(piv);
            l[v][v] = piv * rd[v];
#pragma unroll
            for (int u = v + 1; u < 4; ++u) {
              double a = d[u][v];
#pragma unroll
              for (int m = 0; m < v; ++m) a -= l[u][m] * l[v][m];
              l[u][v] = a * rd[v];
            }
          }
          // this lane's factor entries in columns q .. q + 3 (rows below the block), the block's own rows
          double x[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double a = p[q + v];
#pragma unroll
            for (int m = 0; m < v; ++m) a -= x[m] * l[v][m];
            x[v] = a * rd[v];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (lane == q + u) {
#pragma unroll
              for (int v = 0; v < 4; ++v) x[v] = u >= v ? l[u][v] : 0.0;
            }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            p[q + v] = x[v];
            colbuf[v * TILE + lane] = x[v];
          }
          if (lane == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) rdiag[c0 + q + v] = rd[v];
          }
          wave_order();
          // rank-4 update of the rest of the panel row: p[j] -= sum_v x[v] L[c0 + j][c0 + q + v]
#pragma unroll
          for (int j = q + 4; j < 16; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) p[j] -= x[v] * colbuf[v * TILE + j];
          }
          wave_order();
        }
      }
      if (act) {
#pragma unroll
        for (int j = 0; j < 16; ++j) Ls[row * LSTR + c0 + j] = (lane >= 16 || j <= lane) ? p[j] : 0.0;
      }
      if (lane == 0 && bad != 0) fail = bad;
    }
    __syncthreads();
    if (fail != 0) {
      if (tid == 0) *info = TILE * k + fail;
      return;
    }
    // trailing update of rows and columns c0 + 16 .. 63: one 4 x 4 block of the lower triangle per thread
    const int t4 = (TILE - c0 - 16) / 4;
    if (tid < t4 * t4) {
      const int bi = tid / t4, bj = tid % t4;
      if (bj <= bi) {
        const int i0 = c0 + 16 + 4 * bi, j0 = c0 + 16 + 4 * bj;
        double o[4][4] = {};
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          double li[4], lj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            li[u] = Ls[(i0 + u) * LSTR + c0 + m];
            lj[u] = Ls[(j0 + u) * LSTR + c0 + m];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) o[u][v] += li[u] * lj[v];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) Ls[(i0 + u) * LSTR + j0 + v] -= o[u][v];
      }
    }
    __syncthreads();
  }
  // the factor back, zero above the diagonal
  for (int e = tid; e < TILE * TILE; e += 256) {
    const int i = e >> 6, j = e & 63;
    tile[e] = j <= i ? Ls[i * LSTR + j] : 0.0;
  }
  // inverse: diagonal 16 x 16 blocks, wavefront w inverts block w (lane j < 16 holds column j)
  {
    const int o = 16 * wave;
    if (lane < 16) {
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        double acc = lane == i ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < i; ++m) acc -= Ls[(o + i) * LSTR + o + m] * x[m];
        x[i] = acc * rdiag[o + i];
        Xs[(o + i) * LSTR + o + lane] = x[i];
      }
    }
  }
  __syncthreads();
  // off-diagonal blocks (a, b = a - dist): Y = sum_{c=b}^{a-1} L_ac X_cb, then X_ab = -X_aa Y; the sums run
  // over compile-time lengths in four independent partial sums so that the LDS loads pipeline
#pragma unroll
  for (int dist = 1; dist < 4; ++dist) {
    const int nblk = 4 - dist;
    for (int q = tid; q < nblk * 256; q += 256) {
      const int blk = q >> 8, i = (q >> 4) & 15, j = q & 15;
      const int a = dist + blk, b = blk;
      const double* lrow = Ls + (16 * a + i) * LSTR + 16 * b;
      const double* xcol = Xs + (16 * b) * LSTR + 16 * b + j;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 0; m < 16 * dist; ++m) acc[m & 3] += lrow[m] * xcol[m * LSTR];
      Ys[blk * 16 * 17 + i * 17 + j] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    __syncthreads();
    for (int q = tid; q < nblk * 256; q += 256) {
      const int blk = q >> 8, i = (q >> 4) & 15, j = q & 15;
      const int a = dist + blk, b = blk;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 0; m < 16; ++m) acc[m & 3] += Xs[(16 * a + i) * LSTR + 16 * a + m] * Ys[blk * 16 * 17 + m * 17 + j];
      Xs[(16 * a + i) * LSTR + 16 * b + j] = -((acc[0] + acc[1]) + (acc[2] + acc[3]));
    }
    __syncthreads();
  }
  double* inv = linv + ((size_t)k << 12);
  for (int e = tid; e < TILE * TILE; e += 256) inv[e] = Xs[(e >> 6) * LSTR + (e & 63)];
}

// Sub-diagonal factor tiles of tile column k: L(k + d, k) = A(k + d, k) L_kk^-T. Workgroup (d, quarter)
// forms 16 rows of tile d.
__global__ void __launch_bounds__(256) k_chol_trsm(const double* __restrict__ band, double* __restrict__ lpanel,
                                                   const double* __restrict__ linv, int k, int nt, int kb,
                                                   const int* info) {
  __shared__ double Xs[TILE * LSTR];
  __shared__ double As[16 * LSTR];
  if (*info != 0) return;
  const int d = blockIdx.x / 4 + 1, quarter = blockIdx.x % 4;
  if (d > kb || k + d >= nt) return;
  const int tid = threadIdx.x;
  const double* a = band_tile(const_cast<double*>(band), kb, k, d) + (quarter * 16 << 6);
  const double* inv = linv + ((size_t)k << 12);
  tile_to_lds<TILE>(inv, Xs);
  tile_to_lds<16>(a, As);
  __syncthreads();
  // thread: row r (0..15), columns c0 + 16 v (16 consecutive rows of Xs per load: no bank conflicts)
  const int r = tid >> 4, c0 = tid & 15;
  double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int m = 0; m < TILE; ++m) {
    const double av = As[r * LSTR + m];
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] += av * Xs[(c0 + 16 * v) * LSTR + m];
  }
  double* l = lpanel + (((size_t)k * kb + (d - 1)) << 12) + ((quarter * 16 + r) << 6) + c0;
#pragma unroll
  for (int v = 0; v < 4; ++v) l[16 * v] = o[v];
}

// Trailing update after tile column k: tile (k + d1, k + d2) -= L(k + d1, k) L(k + d2, k)^T for
// 1 <= d2 <= d1 <= kb; workgroup (pair, quarter) updates 16 rows of the tile.
__global__ void __launch_bounds__(256) k_chol_update(double* __restrict__ band, const double* __restrict__ lpanel,
                                                     int k, int nt, int kb, const int* info) {
  __shared__ double Ps[16 * LSTR];
  __shared__ double Qs[TILE * LSTR];
  if (*info != 0) return;
  int d1 = 1, idx = blockIdx.x / 4;
  const int quarter = blockIdx.x % 4;
  while (idx >= d1) {
    idx -= d1;
    ++d1;
  }
  const int d2 = idx + 1;
  if (d1 > kb || k + d1 >= nt) return;
  const int tid = threadIdx.x;
  const double* p = lpanel + (((size_t)k * kb + (d1 - 1)) << 12) + (quarter * 16 << 6);
  const double* q = lpanel + (((size_t)k * kb + (d2 - 1)) << 12);
  tile_to_lds<TILE>(q, Qs);
  tile_to_lds<16>(p, Ps);
  __syncthreads();
  const int r = tid >> 4, c0 = tid & 15;
  double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int m = 0; m < TILE; ++m) {
    const double pv = Ps[r * LSTR + m];
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] += pv * Qs[(c0 + 16 * v) * LSTR + m];
  }
  double* out = band_tile(band, kb, k + d2, d1 - d2) + ((quarter * 16 + r) << 6) + c0;
#pragma unroll
  for (int v = 0; v < 4; ++v) out[16 * v] -= o[v];
}

// B = L^-1 by blocked forward substitution: workgroup (J, g) owns the 16 columns 64 J + 16 g .. + 15 and
// walks the row tiles I = J, J + 1, ...: X_I = Linv_II (E_I - sum_d L(I, I - d) X_{I - d}), the last kb
// tiles of X kept in LDS. The factor tiles stream through one LDS tile buffer, the next two prefetched into
// registers while the current one is used. Each product splits its 64-long sums over the four wavefronts
// (16 terms each, a 4 x 4 block of outputs per lane), and the partial sums meet in LDS in a fixed order.
// FP32 B (row-major) and B^T (row-major) receive the rows and columns below N3, and so do the FP16 copies
// B~ = S^-1 B and its transpose when Bh is not null, S = diag(B) = the reciprocal factor diagonal (the
// diagonal of linv), with |B~| clamped to the FP16 range.
constexpr int INV_COLS = 16;
__global__ void __launch_bounds__(256) k_band_inverse(const double* __restrict__ lpanel,
                                                      const double* __restrict__ linv, int nt, int kb, int N3,
                                                      int N3p, float* __restrict__ B, float* __restrict__ BT,
                                                      _Float16* __restrict__ Bh, _Float16* __restrict__ BTh,
                                                      const int* info) {
  __shared__ double Xh[(MAX_KB + 1) * TILE * INV_COLS];
  __shared__ double T[TILE * INV_COLS];
  __shared__ double Part[4 * TILE * INV_COLS];
  __shared__ double Lb[TILE * LSTR];
  if (*info != 0) return;
  const int J = blockIdx.x / (TILE / INV_COLS), g = blockIdx.x % (TILE / INV_COLS);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = (lane >> 2) * 4, c0 = (lane & 3) * 4, m0 = 16 * wave;  // 4 x 4 outputs, 16 terms per wave
  const int j0 = TILE * J + INV_COLS * g;                                // first global column of the group
  // stage (I, d): d = 1 .. min(I - J, kb) are the factor tiles L(I, I - d), d = 0 the diagonal inverse
  auto tile_of = [&](int I, int d) -> const double* {
    return d == 0 ? linv + ((size_t)I << 12) : lpanel + (((size_t)(I - d) * kb + (d - 1)) << 12);
  };
  auto first_stage = [&](int I) { return (I - J < kb ? I - J : kb) > 0 ? 1 : 0; };
  auto next_stage = [&](int I, int d, int* nI, int* nd) {
    const int dmax = I - J < kb ? I - J : kb;
    if (d != 0 && d < dmax) {
      *nI = I;
      *nd = d + 1;
    } else if (d != 0) {
      *nI = I;
      *nd = 0;
    } else {
      *nI = I + 1;
      *nd = first_stage(I + 1);
    }
  };
  // two stages of factor tiles in flight in registers
  double pre[16], pre2[16];
  int I = J, d = first_stage(J);
  int I2, d2;
  next_stage(I, d, &I2, &d2);
  {
    const double* src = tile_of(I, d);
#pragma unroll
    for (int i = 0; i < 16; ++i) pre[i] = src[tid + 256 * i];
    if (I2 < nt) {
      const double* src2 = tile_of(I2, d2);
#pragma unroll
      for (int i = 0; i < 16; ++i) pre2[i] = src2[tid + 256 * i];
    }
  }
  double acc[4][4] = {};   // this wave's part of -sum_d L(I, I - d) X_{I - d}
  while (I < nt) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int e = tid + 256 * i;
      Lb[(e >> 6) * LSTR + (e & 63)] = pre[i];
    }
    if (d == 0) {
      // partial sums of the factor tiles, then T = E_I + their sum
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) Part[wave * TILE * INV_COLS + (r0 + u) * INV_COLS + c0 + v] = acc[u][v];
    }
    __syncthreads();
    int nI, nd, nI2 = nt, nd2 = 0;
    next_stage(I, d, &nI, &nd);
    if (nI < nt) next_stage(nI, nd, &nI2, &nd2);
#pragma unroll
    for (int i = 0; i < 16; ++i) pre[i] = pre2[i];
    if (nI2 < nt) {
      const double* src = tile_of(nI2, nd2);
#pragma unroll
      for (int i = 0; i < 16; ++i) pre2[i] = src[tid + 256 * i];
    }
    if (d != 0) {
      const double* x = Xh + ((I - d) % (kb + 1)) * TILE * INV_COLS;
#pragma unroll
      for (int m = m0; m < m0 + 16; ++m) {
        double lv[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) lv[u] = Lb[(r0 + u) * LSTR + m];
#pragma unroll
        for (int v = 0; v < 4; ++v) xv[v] = x[m * INV_COLS + c0 + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] -= lv[u] * xv[v];
      }
    } else {
      // T = E_I + sum over the four partials
      for (int e = tid; e < TILE * INV_COLS; e += 256) {
        const int rr = e / INV_COLS, cc = e % INV_COLS;
        const double eye = (I == J && rr == INV_COLS * g + cc) ? 1.0 : 0.0;
        T[e] = eye + ((Part[e] + Part[TILE * INV_COLS + e]) +
                      (Part[2 * TILE * INV_COLS + e] + Part[3 * TILE * INV_COLS + e]));
      }
      __syncthreads();
      double xo[4][4] = {};
#pragma unroll
      for (int m = m0; m < m0 + 16; ++m) {
        double lv[4], tv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) lv[u] = Lb[(r0 + u) * LSTR + m];
#pragma unroll
        for (int v = 0; v < 4; ++v) tv[v] = T[m * INV_COLS + c0 + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) xo[u][v] += lv[u] * tv[v];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) Part[wave * TILE * INV_COLS + (r0 + u) * INV_COLS + c0 + v] = xo[u][v];
      __syncthreads();
      double* slot = Xh + (I % (kb + 1)) * TILE * INV_COLS;
      for (int e = tid; e < TILE * INV_COLS; e += 256) {
        const double xv = (Part[e] + Part[TILE * INV_COLS + e]) +
                          (Part[2 * TILE * INV_COLS + e] + Part[3 * TILE * INV_COLS + e]);
        slot[e] = xv;
        const int R = TILE * I + e / INV_COLS, C = j0 + e % INV_COLS;
        if (R < N3 && C < N3) {
          B[(size_t)R * N3p + C] = (float)xv;
          BT[(size_t)C * N3p + R] = (float)xv;
          if (Bh) {
            const double sr = linv[((size_t)(R >> 6) << 12) + (R & 63) * 65];
            const _Float16 h = (_Float16)fmin(fmax(xv / sr, -65504.0), 65504.0);
            Bh[(size_t)R * N3p + C] = h;
            BTh[(size_t)C * N3p + R] = h;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    }
    I = nI;
    d = nd;
  }
}

// Tile row I of B = L^-1 as a sweep over tile rows, so that row I can run as soon as the panel of tile column I
// has produced L_II^-1 (the factor tiles L(I, I - d) came from earlier TRSMs): X(I, J) = L_II^-1 (E(I, J) -
// sum_{d=1..min(kb, I-J)} L(I, I - d) X(I - d, J)) for the 64 (I + 1) columns of B's lower triangle, 16 per
// workgroup. X keeps the last kb + 1 tile rows in FP64 ([kb + 1][64][N3p], slot I % (kb + 1)); B (row-major),
// B^T and, when Bh is not null, the FP16 copies B~ = S^-1 B (S = diag(B), from L_II^-1) and its transpose get
// the rows and columns below N3. Thread (row r, four columns): FP64 sums in a fixed order (d, then m).
__global__ void __launch_bounds__(256) k_inverse_row(const double* __restrict__ lpanel, const double* __restrict__ linv,
                                                     int I, int kb, int N3, int N3p, double* __restrict__ X,
                                                     float* __restrict__ B, float* __restrict__ BT,
                                                     _Float16* __restrict__ Bh, _Float16* __restrict__ BTh,
                                                     const int* info) {
  __shared__ double Ls[TILE * LSTR];
  __shared__ double Xs[TILE * 17];
  if (*info != 0) return;
  const int tid = threadIdx.x, r = tid >> 2, cq = (tid & 3) * 4;
  const int j0 = INV_COLS * blockIdx.x, J = j0 / TILE;
  const int dmax = I - J < kb ? I - J : kb;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int d = 1; d <= dmax; ++d) {
    __syncthreads();
    tile_to_lds<TILE>(lpanel + (((size_t)(I - d) * kb + (d - 1)) << 12), Ls);
    const double* xs = X + (size_t)((I - d) % (kb + 1)) * TILE * N3p + j0;
    for (int e = tid; e < TILE * INV_COLS; e += 256) Xs[(e / INV_COLS) * 17 + e % INV_COLS] = xs[(size_t)(e / INV_COLS) * N3p + e % INV_COLS];
    __syncthreads();
#pragma unroll 8
    for (int m = 0; m < TILE; ++m) {
      const double l = Ls[r * LSTR + m];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] -= l * Xs[m * 17 + cq + u];
    }
  }
  __syncthreads();
  tile_to_lds<TILE>(linv + ((size_t)I << 12), Ls);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int col = j0 + cq + u;
    Xs[r * 17 + cq + u] = (col == TILE * I + r ? 1.0 : 0.0) + acc[u];
  }
  __syncthreads();
  double out[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int m = 0; m < TILE; ++m) {
    const double l = Ls[r * LSTR + m];
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u] += l * Xs[m * 17 + cq + u];
  }
  const int R = TILE * I + r;
  double* xo = X + (size_t)(I % (kb + 1)) * TILE * N3p + (size_t)r * N3p + j0 + cq;
  const double sr = Ls[r * LSTR + r];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    xo[u] = out[u];
    const int C = j0 + cq + u;
    if (R < N3 && C < N3 && C <= R) {
      B[(size_t)R * N3p + C] = (float)out[u];
      BT[(size_t)C * N3p + R] = (float)out[u];
      if (Bh) {
        const _Float16 h = (_Float16)fmin(fmax(out[u] / sr, -65504.0), 65504.0);
        Bh[(size_t)R * N3p + C] = h;
        BTh[(size_t)C * N3p + R] = h;
      }
    }
  }
}

// dense A_c (both triangles, row-major) for the check entry point
__global__ void __launch_bounds__(256) k_galerkin_dense(const uint32_t* __restrict__ pair_ptr,
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

// The FP16 form of k_tri_gemv: T holds B~ (LOWER) or its transpose; eight halves per lane and step, FP32
// products and sums in the same order as k_tri_gemv. LOWER also applies S^2 (the square of the row's
// diagonal of B, read from linv), so that the pair computes (S B~)^T (S B~) x.
template <bool LOWER>
__global__ void __launch_bounds__(256) k_tri_gemv_h(const _Float16* __restrict__ T, const float* __restrict__ x,
                                                    const double* __restrict__ linv, int N3, int N3p,
                                                    float* __restrict__ y, const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= N3) return;
  const uint4* m = reinterpret_cast<const uint4*>(T + (size_t)row * N3p);
  const float4* v = reinterpret_cast<const float4*>(x);
  const int lo = LOWER ? 0 : row / 8, hi = LOWER ? row / 8 + 1 : N3p / 8;
  float acc = 0.0f;
  for (int j0 = lo + lane; j0 < hi; j0 += 4 * 64) {
    uint4 a[4];
    float4 b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + 64 * u;
      if (j < hi) {
        a[u] = m[j];
        b0[u] = v[2 * j];
        b1[u] = v[2 * j + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (j0 + 64 * u < hi) {
        union {
          uint4 q;
          _Float16 h[8];
        } w;
        w.q = a[u];
        acc = acc + ((float)w.h[0] * b0[u].x + (float)w.h[1] * b0[u].y + (float)w.h[2] * b0[u].z +
                     (float)w.h[3] * b0[u].w);
        acc = acc + ((float)w.h[4] * b1[u].x + (float)w.h[5] * b1[u].y + (float)w.h[6] * b1[u].z +
                     (float)w.h[7] * b1[u].w);
      }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) acc = acc + __shfl_down(acc, s, 64);
  if (lane == 0) {
    if (LOWER) {
      const float sr = (float)linv[((size_t)(row >> 6) << 12) + (row & 63) * 65];
      acc = acc * (sr * sr);
    }
    y[row] = acc;
  }
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

// One iteration's x / r update (reduce.wgsl MODE 1 with update_dot.wgsl) folded into the preconditioner
// launches: every wavefront forms alpha = rz / pq from the p.q partials in the fixed order, the restriction
// gathers the children's updated residual r - alpha q on the fly, and the prolongation writes
// x += alpha p and r -= alpha q before z = D^-1 r + P y. The arithmetic is that of the separate update,
// so the iterates are bit-identical to it.
__global__ void __launch_bounds__(256) k_restrict_update(const uint32_t* __restrict__ child_ptr,
                                                         const uint32_t* __restrict__ child,
                                                         const float* __restrict__ w, const float4* __restrict__ r,
                                                         const float4* __restrict__ q, int Nc, float* __restrict__ rc,
                                                         const float* __restrict__ partial_pq, int G, PcgState* st,
                                                         int it) {
  if (st->status != 0) return;
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
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (I >= Nc) return;
  const uint32_t t0 = child_ptr[I], n = child_ptr[I + 1] - t0;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (uint32_t j = lane; j < n; j += 64) {
    const uint32_t s = child[t0 + j];
    const float ww = w[s];
    const float4 v = r[s >> 2], qq = q[s >> 2];
    const float rx = v.x - alpha * qq.x, ry = v.y - alpha * qq.y, rz = v.z - alpha * qq.z;
    a0 = a0 + ww * rx;
    a1 = a1 + ww * ry;
    a2 = a2 + ww * rz;
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

__global__ void __launch_bounds__(256) k_prolong_update(const int32_t* __restrict__ parent, const float* __restrict__ w,
                                                        const float* __restrict__ yc, const float* __restrict__ dinv,
                                                        const float4* __restrict__ p, const float4* __restrict__ q,
                                                        float4* __restrict__ x, float4* __restrict__ r,
                                                        float4* __restrict__ z, float* __restrict__ partials, int R,
                                                        const float* __restrict__ partial_pq, int G,
                                                        const PcgState* st, int it) {
  __shared__ float2 lds[256];
  if (st->status != 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  float4 pr = make_float4(0, 0, 0, 0), qr = pr, xr = pr, rv = pr;
  if (row < R) {
    pr = p[row];
    qr = q[row];
    xr = x[row];
    rv = r[row];
  }
  const float alpha = st->rz[it & 1] / wave_reduce_partials(partial_pq, G);
  float rzs = 0.0f, rrs = 0.0f;
  if (row < R) {
    xr.x = xr.x + alpha * pr.x;
    xr.y = xr.y + alpha * pr.y;
    xr.z = xr.z + alpha * pr.z;
    rv.x = rv.x - alpha * qr.x;
    rv.y = rv.y - alpha * qr.y;
    rv.z = rv.z - alpha * qr.z;
    x[row] = xr;
    r[row] = rv;
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
    z[row] = make_float4(z0, z1, z2, 0.0f);
    rzs = rv.x * z0 + rv.y * z1 + rv.z * z2;
    rrs = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  }
  const float2 sum = tree256x2(rzs, rrs, lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sum.x;
    partials[2 * blockIdx.x + 1] = sum.y;
  }
}

// ---- host ------------------------------------------------------------------------------------
// Coarse layout of a grid sheet with `columns` vertices per side ((columns - 1) divisible by the factor): parents,
// bilinear weights, children and the Galerkin term lists.
// The coarsening factor of rules.coarse_factor; 0 picks the smallest power of two from 4 up that keeps the coarse
// grid at most 33 vertices a side (n129: 4, n257: 8), so that the dense inverse factor stays near 43 MB.
int coarse_factor_of(const drape_rules& r, int columns) {
  if (r.coarse_factor > 0) return r.coarse_factor;
  int f = 4;
  while ((columns - 1) / f + 1 > 33 && (columns - 1) % (2 * f) == 0) f *= 2;
  return f;
}

int two_level_create(Context& c, int columns) {
  const int f = coarse_factor_of(c.rules, columns);
  if (columns < 2 * f + 1 || f < 2 || (columns - 1) % f != 0 || (size_t)columns * columns != (size_t)c.V) {
    set_error("two-level preconditioner: the sheet must be a grid with (columns - 1) divisible by the coarsening "
              "factor (" + std::to_string(f) + ") and at least three coarse vertices a side");
    return DRAPE_E_INVALID;
  }
  TwoLevel* t = new TwoLevel();
  const int n = columns, nc = (n - 1) / f + 1;
  t->nc = nc;
  std::vector<int> row_of(c.V, -1);
  for (int row = 0; row < c.R; ++row) row_of[c.row_vertex[row]] = row;
  // coarse rows: free coarse vertices in vertex order
  std::vector<int> coarse_of(nc * nc, -1);
  int Nc = 0;
  for (int b = 0; b < nc; ++b)
    for (int a = 0; a < nc; ++a)
      if (row_of[(f * b) * n + f * a] >= 0) coarse_of[b * nc + a] = Nc++;
  t->Nc = Nc;
  t->N3 = 3 * Nc;
  t->N3p = (t->N3 + 63) / 64 * 64;
  t->parent.assign(4 * (size_t)c.R, -1);
  t->bilinear.assign(4 * (size_t)c.R, 0.0f);
  for (int row = 0; row < c.R; ++row) {
    const int v = (int)c.row_vertex[row];
    const int x = v % n, y = v / n;
    const int a0 = std::min(x / f, nc - 2), b0 = std::min(y / f, nc - 2);
    const double fx = (x - (double)f * a0) / f, fy = (y - (double)f * b0) / f;
    const int ca[4] = {a0, a0 + 1, a0, a0 + 1}, cb[4] = {b0, b0, b0 + 1, b0 + 1};
    const double wt[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
    for (int k = 0; k < 4; ++k) {
      const int I = coarse_of[cb[k] * nc + ca[k]];
      if (I < 0 || wt[k] == 0.0) continue;
      t->parent[4 * row + k] = I;
      t->bilinear[4 * row + k] = (float)wt[k];
    }
  }
  std::vector<float2> fxy(c.R);
  for (int row = 0; row < c.R; ++row) {
    const int v = (int)c.row_vertex[row];
    const int x = v % n, y = v / n;
    const int a0 = std::min(x / f, nc - 2), b0 = std::min(y / f, nc - 2);
    fxy[row] = make_float2((float)((x - (double)f * a0) / f), (float)((y - (double)f * b0) / f));
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
  // band of the coarse matrix in 64-row tiles
  t->nt = (t->N3 + TILE - 1) / TILE;
  for (uint32_t key : pair_key) {
    const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
    if (I >= J) t->kb = std::max(t->kb, (3 * I + 2) / TILE - (3 * J) / TILE);
  }
  if (t->kb > MAX_KB) {
    set_error("two-level preconditioner: the coarse grid is too large for the banded coarse solve");
    delete t;
    return DRAPE_E_INVALID;
  }
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
  TL_ALLOC(t->d_wbil, t->bilinear.size());
  TL_ALLOC(t->d_fxy, fxy.size());
  TL_ALLOC(t->d_child_ptr, child_ptr.size());
  TL_ALLOC(t->d_child, child.size());
  TL_ALLOC(t->d_pair_ptr, pair_ptr.size());
  TL_ALLOC(t->d_pair_key, pair_key.size());
  TL_ALLOC(t->d_pair_terms, terms.size());
  TL_ALLOC(t->d_band, (size_t)t->nt * (t->kb + 1) * TILE * TILE);
  TL_ALLOC(t->d_linv, (size_t)t->nt * TILE * TILE);
  TL_ALLOC(t->d_lpanel, (size_t)t->nt * std::max(t->kb, 1) * TILE * TILE);
  TL_ALLOC(t->d_xring, (size_t)(t->kb + 1) * TILE * t->N3p);
  if ((e = hipStreamCreateWithFlags(&t->inv_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  if ((e = hipEventCreateWithFlags(&t->ev_panel, hipEventDisableTiming)) != hipSuccess) return fail(e);
  if ((e = hipEventCreateWithFlags(&t->ev_inverse, hipEventDisableTiming)) != hipSuccess) return fail(e);
  TL_ALLOC(t->d_B, (size_t)t->N3 * t->N3p);
  TL_ALLOC(t->d_BT, (size_t)t->N3 * t->N3p);
  TL_ALLOC(t->d_Bh, (size_t)t->N3 * t->N3p);
  TL_ALLOC(t->d_BTh, (size_t)t->N3 * t->N3p);
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
      (e = up(t->d_wbil, t->bilinear.data(), t->bilinear.size() * 4)) != hipSuccess ||
      (e = up(t->d_fxy, fxy.data(), fxy.size() * sizeof(float2))) != hipSuccess ||
      (e = up(t->d_child_ptr, child_ptr.data(), child_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_child, child.data(), child.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_ptr, pair_ptr.data(), pair_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_key, pair_key.data(), pair_key.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_terms, terms.data(), terms.size() * sizeof(uint4))) != hipSuccess ||
      (e = hipMemsetAsync(t->d_rc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_yc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_t, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_B, 0, (size_t)t->N3 * t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_BT, 0, (size_t)t->N3 * t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_Bh, 0, (size_t)t->N3 * t->N3p * 2, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_BTh, 0, (size_t)t->N3 * t->N3p * 2, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return fail(e);
  if (c.two_level) two_level_destroy(c.two_level);
  c.two_level = t;
  return DRAPE_OK;
}

static int adopt_pending(Context& c);

int two_level_set_weights(Context& c, const float* w) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  if (!t) {
    set_error("no two-level layout on this context");
    return DRAPE_E_INVALID;
  }
  std::vector<float> v(4 * (size_t)c.R);
  for (size_t i = 0; i < v.size(); ++i) v[i] = t->parent[i] >= 0 ? (w ? w[i] : t->bilinear[i]) : 0.0f;
  HIP_TRY(hipMemcpyAsync(t->d_w, v.data(), v.size() * 4, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  if (t->d_theta) {  // fixed weights replace the learned prolongation
    HIP_TRY(hipFree(t->d_theta));
    t->d_theta = nullptr;
  }
  t->factored = ~0ull;
  return DRAPE_OK;
}

// The learned prolongation: from the next factorisation on, the weights come from the network evaluated on
// each assembled system (null theta: back to the bilinear weights).
int two_level_set_model(Context& c, const float* theta, int count) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  if (!t) {
    set_error("no two-level layout on this context");
    return DRAPE_E_INVALID;
  }
  if (!theta) return two_level_set_weights(c, nullptr);
  if (count != gnn_param_count()) {
    set_error("prolongation model: expected " + std::to_string(gnn_param_count()) + " parameters");
    return DRAPE_E_INVALID;
  }
  if (!t->d_theta) HIP_TRY(hipMalloc((void**)&t->d_theta, (size_t)count * sizeof(float)));
  if (!t->d_gnn) {
    const size_t n = (size_t)c.R * (1 + 13 + 3 * 32) + (size_t)c.B;
    HIP_TRY(hipMalloc((void**)&t->d_gnn, n * sizeof(float)));
  }
  HIP_TRY(hipMemcpyAsync(t->d_theta, theta, (size_t)count * sizeof(float), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  t->factored = ~0ull;
  return DRAPE_OK;
}

int two_level_weights(Context& c, float* out) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  HIP_TRY(hipMemcpyAsync(out, t->d_w, 4 * (size_t)c.R * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

GnnArgs two_level_gnn_args(Context& c) {
  TwoLevel* t = c.two_level;
  GnnArgs a;
  a.stream = c.stream;
  a.R = c.R;
  a.sell_val = c.sell_val;
  a.block_dst = c.block_dst;
  a.diag_slot = c.diag_slot;
  a.row_ptr = c.csr_row_ptr;
  a.col = c.csr_col;
  a.fxy = t->d_fxy;
  a.wbil = t->d_wbil;
  a.theta = t->d_theta;
  a.scale = t->d_gnn;
  a.features = a.scale + c.R;
  a.edge = a.features + 13 * (size_t)c.R;
  a.h0 = a.edge + c.B;
  a.h1 = a.h0 + 32 * (size_t)c.R;
  a.msg = a.h1 + 32 * (size_t)c.R;
  a.w = t->d_w;
  return a;
}

static int run_model(Context& c) { return gnn_infer(two_level_gnn_args(c)); }

// The FP16 copies are formed and applied under rules.coarse_half = 1, and under 2 (auto) once the FP32 pair
// B, B^T outgrows half of the 256 MB Infinity Cache.
static bool use_half(const Context& c, const TwoLevel* t) {
  const bool big = (size_t)8 * t->N3 * t->N3p > ((size_t)128 << 20);
  return c.rules.coarse_half == 1 || (c.rules.coarse_half == 2 && big);
}

static int factorise(Context& c, bool model, bool* ok);

int two_level_factor_weights(Context& c, bool* ok) {
  TRY_RC(adopt_pending(c));
  return factorise(c, false, ok);
}

// Exchanges the current factorisation (its inverse factor copies, diagonal-tile inverses and bookkeeping)
// with the saved one; `keep` marks the outgoing current one as valid. Allocates the saved buffers on first
// use.
static int swap_saved(Context& c, bool keep) {
  TwoLevel* t = c.two_level;
  TwoLevel::Saved& v = t->saved;
  if (!v.B) {
    const size_t nb = (size_t)t->N3 * t->N3p;
    HIP_TRY(hipMalloc((void**)&v.linv, (size_t)t->nt * TILE * TILE * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&v.B, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&v.BT, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&v.Bh, nb * sizeof(_Float16)));
    HIP_TRY(hipMalloc((void**)&v.BTh, nb * sizeof(_Float16)));
    // the copies are written below/above the diagonal only: the rest stays zero
    HIP_TRY(hipMemsetAsync(v.B, 0, nb * sizeof(float), c.stream));
    HIP_TRY(hipMemsetAsync(v.BT, 0, nb * sizeof(float), c.stream));
    HIP_TRY(hipMemsetAsync(v.Bh, 0, nb * sizeof(_Float16), c.stream));
    HIP_TRY(hipMemsetAsync(v.BTh, 0, nb * sizeof(_Float16), c.stream));
  }
  std::swap(v.linv, t->d_linv);
  std::swap(v.B, t->d_B);
  std::swap(v.BT, t->d_BT);
  std::swap(v.Bh, t->d_Bh);
  std::swap(v.BTh, t->d_BTh);
  std::swap(v.half_valid, t->half_valid);
  const bool was_valid = v.valid;
  const uint64_t f = v.factored, fs = v.factored_shift;
  const double fh = v.factored_h, br = v.base_rate;
  const int so = v.solves;
  v.valid = keep;
  v.factored = t->factored;
  v.factored_shift = t->factored_shift;
  v.factored_h = t->factored_h;
  v.base_rate = t->base_rate;
  v.solves = t->solves;
  if (was_valid) {
    t->factored = f;
    t->factored_shift = fs;
    t->factored_h = fh;
    t->base_rate = br;
    t->solves = so;
    t->ok = true;
  } else {
    t->factored = ~0ull;
    t->ok = false;
  }
  t->refresh = false;
  return DRAPE_OK;
}

float* two_level_weights_device(Context& c) { return c.two_level->d_w; }

// Waits for a factorisation running beside the solves and makes it the current one (its buffers and the
// replaced one's swap).
static int adopt_pending(Context& c) {
  TwoLevel* t = c.two_level;
  TwoLevel::Pending& q = t->pending;
  if (!q.active) return DRAPE_OK;
  TRY_RC(two_level_launch_deferred(c));
  HIP_TRY(hipEventSynchronize(q.done));
  c.read_bytes += sizeof(int);
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, q.began, q.ended));
  t->factor_ms += ms;
  std::swap(q.w, t->d_w);
  std::swap(q.linv, t->d_linv);
  std::swap(q.B, t->d_B);
  std::swap(q.BT, t->d_BT);
  std::swap(q.Bh, t->d_Bh);
  std::swap(q.BTh, t->d_BTh);
  std::swap(q.half_valid, t->half_valid);
  t->factored = q.factored;
  t->factored_shift = q.factored_shift;
  t->factored_h = q.factored_h;
  t->refresh = false;
  t->solves = 0;
  t->factorisations++;
  t->ok = *q.h_info == 0;
  if (!t->ok) t->failures++;
  q.active = false;
  q.galerkin = false;
  return DRAPE_OK;
}

// Drops a side-stream factorisation whose chain could not be queued completely: the side stream is drained
// and nothing is adopted (the kept factorisation stays current).
static void abandon_pending(TwoLevel* t) {
  TwoLevel::Pending& q = t->pending;
  if (q.side) (void)hipStreamSynchronize(q.side);
  q.active = false;
  q.deferred = false;
  q.galerkin = false;
}

int two_level_before_assembly(Context& c) {
  TwoLevel* t = c.two_level;
  if (!t) return DRAPE_OK;
  if (t->pending.deferred) {
    const int rc = two_level_launch_deferred(c);
    if (rc != DRAPE_OK) {
      abandon_pending(t);
      return rc;
    }
  }
  if (t->pending.galerkin) {
    HIP_TRY(hipStreamWaitEvent(c.stream, t->pending.read, 0));
    t->pending.galerkin = false;
  }
  return DRAPE_OK;
}

static int factorise_chain(Context& c, hipStream_t s, const float* w, double* linv, float* B, float* BT,
                           _Float16* Bh, _Float16* BTh, bool half, hipEvent_t after_product);

// The factorisation of the assembled system on the side stream, into the pending buffers; the current one
// keeps serving until the next system adopts it.
static int factorise_async(Context& c) {
  TwoLevel* t = c.two_level;
  TwoLevel::Pending& q = t->pending;
  if (!q.side) {
    const size_t nb = (size_t)t->N3 * t->N3p;
    HIP_TRY(hipStreamCreateWithFlags(&q.side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&q.ready, &q.read, &q.done}) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    for (hipEvent_t* e : {&q.began, &q.ended}) HIP_TRY(hipEventCreate(e));
    HIP_TRY(hipMalloc((void**)&q.w, 4 * (size_t)c.R * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&q.linv, (size_t)t->nt * TILE * TILE * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&q.B, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&q.BT, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&q.Bh, nb * sizeof(_Float16)));
    HIP_TRY(hipMalloc((void**)&q.BTh, nb * sizeof(_Float16)));
    HIP_TRY(hipHostMalloc((void**)&q.h_info, sizeof(int)));
    // the copies are written below/above the diagonal only: the rest stays zero
    HIP_TRY(hipMemsetAsync(q.B, 0, nb * sizeof(float), c.stream));
    HIP_TRY(hipMemsetAsync(q.BT, 0, nb * sizeof(float), c.stream));
    HIP_TRY(hipMemsetAsync(q.Bh, 0, nb * sizeof(_Float16), c.stream));
    HIP_TRY(hipMemsetAsync(q.BTh, 0, nb * sizeof(_Float16), c.stream));
  }
  // the side stream starts once the assembled system is complete; its launches are queued after the
  // solve's first ones (two_level_launch_deferred), so that the device is not idle while the host queues them
  HIP_TRY(hipEventRecord(q.ready, c.stream));
  q.active = true;
  q.deferred = true;
  q.factored = c.system_version;
  q.factored_shift = c.shift_version;
  q.factored_h = c.shift_h;
  t->refresh = false;
  return DRAPE_OK;
}

static int launch_chain(Context& c);

// Queues a decided side-stream factorisation. A failure part-way leaves nothing to adopt: the side stream is
// drained and the pending state cleared, so the kept factorisation stays current.
int two_level_launch_deferred(Context& c) {
  TwoLevel* t = c.two_level;
  if (!t || !t->pending.deferred) return DRAPE_OK;
  t->pending.deferred = false;
  const int rc = launch_chain(c);
  if (rc != DRAPE_OK) abandon_pending(t);
  return rc;
}

static int launch_chain(Context& c) {
  TwoLevel* t = c.two_level;
  TwoLevel::Pending& q = t->pending;
  HIP_TRY(hipStreamWaitEvent(q.side, q.ready, 0));
  HIP_TRY(hipEventRecord(q.began, q.side));
  if (t->d_theta) {
    GnnArgs a = two_level_gnn_args(c);
    a.stream = q.side;
    a.w = q.w;
    TRY_RC(gnn_infer(a));
  } else {
    HIP_TRY(hipMemcpyAsync(q.w, t->d_w, 4 * (size_t)c.R * sizeof(float), hipMemcpyDeviceToDevice, q.side));
  }
  q.half_valid = use_half(c, t);
  TRY_RC(factorise_chain(c, q.side, q.w, q.linv, q.B, q.BT, q.Bh, q.BTh, q.half_valid, q.read));
  HIP_TRY(hipEventRecord(q.ended, q.side));
  HIP_TRY(hipMemcpyAsync(q.h_info, t->d_info, sizeof(int), hipMemcpyDeviceToHost, q.side));
  HIP_TRY(hipEventRecord(q.done, q.side));
  q.galerkin = true;
  return DRAPE_OK;
}

// A_c = P^T A P of the assembled system, its Cholesky factor and the factor's inverse, unless the current
// factorisation may serve this system (same system, or coarse_reuse with unchanged mu and step size and no
// slowdown). *ok is false when A_c is not positive definite.
int two_level_prepare(Context& c, bool* ok) {
  TwoLevel* t = c.two_level;
  TRY_RC(adopt_pending(c));
  const bool fresh = t->factored == c.system_version;
  const int rule = c.rules.coarse_reuse;
  const bool by_h = rule == 3 || rule == 4;
  const bool reusable = rule && t->factored != ~0ull && t->ok && !t->refresh &&
                        (rule == 2 || t->factored_shift == c.shift_version || (by_h && t->factored_h == c.shift_h));
  if (fresh || reusable) {
    *ok = t->ok;
    return DRAPE_OK;
  }
  if (rule == 4) {
    // the step size moved: keep the current factorisation (unless PCG slowed down on it) and return to the
    // saved one if it belongs to the new step size
    const bool keep = t->factored != ~0ull && t->ok && !t->refresh && t->factored_h != c.shift_h;
    const bool back = t->saved.valid && t->saved.factored_h == c.shift_h;
    if (keep || back) TRY_RC(swap_saved(c, keep));
    if (back) {
      t->restores++;
      *ok = t->ok;
      return DRAPE_OK;
    }
  }
  // 2: only a refresh after a slowdown or breakdown at an unchanged step size goes beside the solve
  const bool same_h = t->factored_h == c.shift_h;
  const bool async = c.rules.coarse_async == 1 || (c.rules.coarse_async == 2 && t->refresh && same_h);
  if (async && rule >= 1 && rule <= 3 && t->factored != ~0ull && t->ok) {
    // refactor beside this solve; the kept factorisation serves it
    TRY_RC(factorise_async(c));
    *ok = t->ok;
    return DRAPE_OK;
  }
  return factorise(c, true, ok);
}

// The factorisation itself for the assembled system (with the network's weights first when `model` and a
// network is set, else with the current weights); it becomes the current one.
static int factorise(Context& c, bool model, bool* ok) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  t->factored = c.system_version;
  t->factored_shift = c.shift_version;
  t->factored_h = c.shift_h;
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
  if (model && t->d_theta) TRY_RC(run_model(c));
  t->half_valid = use_half(c, t);
  TRY_RC(factorise_chain(c, s, t->d_w, t->d_linv, t->d_B, t->d_BT, t->d_Bh, t->d_BTh, t->half_valid, nullptr));
  TRY_RC(read_back(c, t->h_info, t->d_info, sizeof(int)));
  t->factorisations++;
  t->ok = *t->h_info == 0;
  if (!t->ok) t->failures++;
  *ok = t->ok;
  return DRAPE_OK;
}

// Galerkin product with weights w, banded Cholesky and the inverse factor into (linv, B, B^T and, when
// `half`, the FP16 copies), on stream s; `after_product` (if not null) is recorded once the product has read
// the assembled matrix.
static int factorise_chain(Context& c, hipStream_t s, const float* w, double* linv, float* B, float* BT,
                           _Float16* Bh, _Float16* BTh, bool half, hipEvent_t after_product) {
  TwoLevel* t = c.two_level;
  HIP_TRY(hipMemsetAsync(t->d_band, 0, (size_t)t->nt * (t->kb + 1) * TILE * TILE * sizeof(double), s));
  HIP_TRY(hipMemsetAsync(t->d_info, 0, sizeof(int), s));
  k_galerkin_band<<<blocks_for(t->npairs, 4), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, w,
                                                        c.sell_val, c.block_dst, t->npairs, t->Nc, t->kb, t->d_band);
  if (after_product) HIP_TRY(hipEventRecord(after_product, s));
  k_band_diagonal<<<blocks_for(t->nt * TILE), 256, 0, s>>>(t->d_band, t->kb, t->nt, t->N3, 1.0 + c.rules.coarse_shift);
  const int updates = t->kb * (t->kb + 1) / 2;
  // The panel chain is latency-bound on one workgroup at a time; tile row k of B = L^-1 runs on the inverse
  // stream as soon as panel k is done, beside the chain's later tile columns, and s waits for the last row.
  for (int k = 0; k < t->nt; ++k) {
    k_chol_panel<<<1, 256, 0, s>>>(t->d_band, linv, k, t->kb, t->d_info);
    HIP_TRY(hipEventRecord(t->ev_panel, s));
    HIP_TRY(hipStreamWaitEvent(t->inv_stream, t->ev_panel, 0));
    k_inverse_row<<<(k + 1) * (TILE / INV_COLS), 256, 0, t->inv_stream>>>(
        t->d_lpanel, linv, k, t->kb, t->N3, t->N3p, t->d_xring, B, BT, half ? Bh : nullptr, half ? BTh : nullptr,
        t->d_info);
    if (k + 1 < t->nt && updates > 0) {
      k_chol_trsm<<<4 * t->kb, 256, 0, s>>>(t->d_band, t->d_lpanel, linv, k, t->nt, t->kb, t->d_info);
      k_chol_update<<<4 * updates, 256, 0, s>>>(t->d_band, t->d_lpanel, k, t->nt, t->kb, t->d_info);
    }
  }
  HIP_TRY(hipEventRecord(t->ev_inverse, t->inv_stream));
  HIP_TRY(hipStreamWaitEvent(s, t->ev_inverse, 0));
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

// The Galerkin operator of the assembled system, row-major [3Nc x 3Nc], into host memory (a dense scratch
// copy allocated on first use). The current factorisation is unaffected.
int two_level_galerkin(Context& c, double* out) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  const size_t n2 = (size_t)t->N3 * t->N3;
  if (!t->d_Ac) HIP_TRY(hipMalloc((void**)&t->d_Ac, n2 * sizeof(double)));
  HIP_TRY(hipMemsetAsync(t->d_Ac, 0, n2 * sizeof(double), s));
  k_galerkin_dense<<<blocks_for(t->npairs), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, t->d_w,
                                                         c.sell_val, c.block_dst, t->npairs, t->Nc, t->d_Ac);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, t->d_Ac, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

// B = L^-1 of the current factorisation (FP32, row-major [3Nc x 3Nc]) into host memory.
int two_level_inverse_factor(Context& c, float* out) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  HIP_TRY(hipMemcpy2DAsync(out, (size_t)t->N3 * sizeof(float), t->d_B, (size_t)t->N3p * sizeof(float),
                           (size_t)t->N3 * sizeof(float), t->N3, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

// After a two-level solve: the first solve after a factorisation sets the iterations-per-decade baseline,
// a later one at more than twice that rate, or a breakdown (iterations < 0), asks for a new factorisation.
void two_level_observe(Context& c, int iterations, double relative_residual) {
  TwoLevel* t = c.two_level;
  if (iterations < 0) {
    if (c.rules.refresh_on_breakdown) t->refresh = true;
    return;
  }
  const double decades = relative_residual > 0 && relative_residual < 1 ? std::max(0.5, -std::log10(relative_residual)) : 0.5;
  const double rate = iterations / decades;
  if (t->solves++ == 0) t->base_rate = rate;
  else if (c.rules.refresh_rate > 0.0 && rate > c.rules.refresh_rate * t->base_rate + 4.0) t->refresh = true;
}

// y_c = B^T B r_c: from the FP16 scaled copies when they are selected and were formed with the current
// factorisation (and not `fp32`), else from FP32 B.
static void coarse_solve(Context& c, const TwoLevel* t, int guard, bool fp32) {
  hipStream_t s = c.stream;
  const int blocks = (t->N3 + 3) / 4;
  if (t->half_valid && use_half(c, t) && !fp32) {
    k_tri_gemv_h<true><<<blocks, 256, 0, s>>>(t->d_Bh, t->d_rc, t->d_linv, t->N3, t->N3p, t->d_t, c.state, guard);
    k_tri_gemv_h<false><<<blocks, 256, 0, s>>>(t->d_BTh, t->d_t, t->d_linv, t->N3, t->N3p, t->d_yc, c.state, guard);
  } else {
    k_tri_gemv<true><<<blocks, 256, 0, s>>>(t->d_B, t->d_rc, t->N3, t->N3p, t->d_t, c.state, guard);
    k_tri_gemv<false><<<blocks, 256, 0, s>>>(t->d_BT, t->d_t, t->N3, t->N3p, t->d_yc, c.state, guard);
  }
}

// One of the coarse solve's triangular GEMVs (lower: t = B r_c, else y_c = B^T t) of the current factorisation,
// `launches` times back to back between two events on the context's stream: average ms per launch and the
// algorithmic bytes of one launch (the triangle in its storage precision, x read and y written once).
int two_level_gemv_timing(Context& c, bool lower, int launches, double* avg_ms, double* bytes) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  const int blocks = (t->N3 + 3) / 4;
  const bool half = t->half_valid && use_half(c, t);
  const double tri = (double)t->N3 * (t->N3 + 1) / 2.0;
  *bytes = tri * (half ? 2.0 : 4.0) + 8.0 * t->N3;
  auto launch = [&]() {
    if (half) {
      if (lower) k_tri_gemv_h<true><<<blocks, 256, 0, s>>>(t->d_Bh, t->d_rc, t->d_linv, t->N3, t->N3p, t->d_t, c.state, 2);
      else k_tri_gemv_h<false><<<blocks, 256, 0, s>>>(t->d_BTh, t->d_t, t->d_linv, t->N3, t->N3p, t->d_yc, c.state, 2);
    } else {
      if (lower) k_tri_gemv<true><<<blocks, 256, 0, s>>>(t->d_B, t->d_rc, t->N3, t->N3p, t->d_t, c.state, 2);
      else k_tri_gemv<false><<<blocks, 256, 0, s>>>(t->d_BT, t->d_t, t->N3, t->N3p, t->d_yc, c.state, 2);
    }
  };
  launch();
  HIP_TRY(hipEventRecord(c.ev_spmv0, s));
  for (int i = 0; i < launches; ++i) launch();
  HIP_TRY(hipEventRecord(c.ev_spmv1, s));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(c.ev_spmv1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c.ev_spmv0, c.ev_spmv1));
  *avg_ms = ms / launches;
  return DRAPE_OK;
}

// z = D^-1 r + P B^T B P^T r for the current r (guard: 0 running, 1 replacement gate, 2 setup with p = z).
int two_level_apply(Context& c, float* partials, int guard) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  k_restrict<<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, c.r, t->Nc, t->d_rc, c.state, guard);
  coarse_solve(c, t, guard, false);
  k_prolong<<<c.G, 256, 0, s>>>(t->d_parent, t->d_w, t->d_yc, c.dinv, c.r, c.z, c.p, partials, c.R, c.state, guard);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int two_level_iterate(Context& c, const float* partial_pq, const float4* p, float* partial_rz, int it) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  k_restrict_update<<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, c.r, c.q, t->Nc, t->d_rc,
                                                    partial_pq, c.G, c.state, it);
  coarse_solve(c, t, 0, false);
  k_prolong_update<<<c.G, 256, 0, s>>>(t->d_parent, t->d_w, t->d_yc, c.dinv, p, c.q, c.x, c.r, c.z, partial_rz, c.R,
                                       partial_pq, c.G, c.state, it);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

CoarseView two_level_view(const Context& c) {
  const TwoLevel* t = c.two_level;
  return {t->d_parent, t->d_w, t->d_yc};
}

int two_level_restrict_solve(Context& c, const float4* r, int guard, bool fp32) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  k_restrict<<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, r, t->Nc, t->d_rc, c.state, guard);
  coarse_solve(c, t, guard, fp32);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int two_level_info(Context& c, int32_t* coarse_rows, int32_t* parent, float* weights, int64_t* factorisations,
                   int64_t* failures, double* factor_ms) {
  const TwoLevel* t = c.two_level;
  if (!t) {
    set_error("no two-level layout on this context");
    return DRAPE_E_INVALID;
  }
  TRY_RC(adopt_pending(c));  // a factorisation still running beside the solves is waited for and counted
  if (coarse_rows) *coarse_rows = t->Nc;
  if (parent) std::copy(t->parent.begin(), t->parent.end(), parent);
  if (weights) std::copy(t->bilinear.begin(), t->bilinear.end(), weights);
  if (factorisations) *factorisations = t->factorisations;
  if (failures) *failures = t->failures;
  if (factor_ms) *factor_ms = t->factor_ms;
  return DRAPE_OK;
}

}  // namespace drape
