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
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "../internal.h"
#include "coarse.hpp"
#include "reduce.hpp"

namespace drape {

struct GramBlock;

struct TwoLevel {
  int nc = 0;        // coarse grid columns
  int Nc = 0;        // coarse rows (free coarse nodes)
  int dof = 3;       // unknowns per coarse node: 3 translations, or 6 with rotations
  int N3 = 0;        // dof Nc coarse unknowns
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
  float4* d_lever = nullptr;     // [4R] rotation lever of each slot at the current factorisation (rotations only)
  float* d_sigma = nullptr;      // [4R] the network's rotation scales of the slots (rotation head only)
  int32_t* d_cvert = nullptr;    // [Nc] vertex of each coarse node
  float inv_spacing = 1.0f;      // 1 / H, H the coarse grid spacing of the rest shape
  double* d_adiag = nullptr;     // [N3p] diagonal of A_c before the factorisation (the pivot floor's reference)
  float* d_theta = nullptr;      // learned-prolongation parameters (null: fixed weights)
  int theta_count = 0;           // their number
  bool rotation_head = false;    // the network sets the rotation scales (rotation unknowns only)
  float* d_gnn = nullptr;        // its scratch: scale [R], features [13R], edge [B], activations [32R] x 3
  uint32_t* d_child_ptr = nullptr;  // [Nc+1]
  uint32_t* d_child = nullptr;      // 4 i + k of each child slot, in fine-row order
  uint32_t* d_pair_ptr = nullptr;   // [npairs+1]
  uint32_t* d_pair_key = nullptr;   // [npairs] I * Nc + J
  uint4* d_pair_terms = nullptr;    // (block b, slot ik, slot jl, 0)
  int nt = 0;                       // 64-row tiles of the coarse matrix (segments padded with identity rows)
  int kb = 0;                       // sub-diagonal tiles in a segment's band
  int ts = 0;                       // separator tiles (0 without the split)
  int seg[4] = {0, 0, 0, 0};        // first tile of the first domain, the second domain, the separator; nt
  int bstart[2] = {0, 0};           // first border tile of each domain
  int real[3] = {0, 0, 0};          // unknowns of each segment
  double tri_entries = 0.0;         // stored entries of B on or below the diagonal (between real unknowns)
  std::vector<int32_t> order;       // [N3] natural unknown 3 I + q at each real position of the factor's ordering
  int32_t* d_pos = nullptr;         // [Nc] position of each coarse node's first unknown in the factor's ordering
  int32_t* d_ppos = nullptr;        // [4R] position of each parent slot's node, -1 when empty
  double* d_band = nullptr;         // [nt][kb + 1][64 x 64]: tile (k + d, k) of A_c, then of L, row-major
  double* d_linv = nullptr;         // [nt][64 x 64]: inverse of each diagonal tile of L
  double* d_lpanel = nullptr;       // [nt][kb][64 x 64]: tile (k + d, k) of L, d = 1..kb
  double* d_border = nullptr;       // [2][kb + 1][ts][64 x 64]: separator rows of A_c against each domain's border
  double* d_lborder = nullptr;      // the same tiles of L
  double* d_sacc = nullptr;         // [ts][ts][64 x 64]: the second domain's Schur update of the separator block
  double* d_xring = nullptr;        // [3][kb + 1][64][N3p]: the last tile rows of L^-1 per segment in FP64
  double* d_ysep = nullptr;         // [ts][64][N3p]: the separator rows' border sums over the domain columns
  float* d_blockval = nullptr;      // [9B]: the assembled blocks' values, contiguous, for the Galerkin product
  hipStream_t inv_stream = nullptr; // the inverse sweep, behind each panel of the factorisation chain
  hipEvent_t ev_panel = nullptr, ev_inverse = nullptr;
  double* d_Ac = nullptr;           // [N3 * N3] dense A_c for the check entry point (allocated on demand)
  float4* d_Ac_lever = nullptr;     // [4R] the check entry point's levers (rotations only, allocated on demand)
  float* d_B = nullptr;             // [N3p * N3p] row-major B = L^-1 in the factor's ordering (zero above the
                                    // diagonal, between the domains and at padding)
  float* d_BT = nullptr;            // [N3p * N3p] row-major B^T (only for the two-GEMV coarse solve)
  _Float16* d_Bh = nullptr;         // [N3p * N3p] FP16 B~ = S^-1 B, S = diag(B)
  _Float16* d_BTh = nullptr;        // [N3p * N3p] its transpose (the same values)
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
    float* w = nullptr;             // the prolongation the saved factorisation was formed with
    float4* lever = nullptr;
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
    float4* lever = nullptr;
    double* linv = nullptr;
    float *B = nullptr, *BT = nullptr;
    _Float16 *Bh = nullptr, *BTh = nullptr;
    int* h_info = nullptr;          // pinned [2]: failure, replaced pivots
    hipStream_t side = nullptr;
    hipEvent_t ready = nullptr, read = nullptr, done = nullptr;
    hipEvent_t began = nullptr, ended = nullptr;  // timing events around the chain on the side stream
  } pending;
  float* d_t = nullptr;             // [N3p] B r_c
  float* d_rc = nullptr;            // [N3p]
  float* d_yc = nullptr;            // [N3p]
  GramBlock* d_gram = nullptr;      // [gram_blocks] row blocks of the fused coarse solve (null: two GEMVs)
  int gram_blocks = 0;
  int4* d_gram_cover = nullptr;     // [N3p / 4] the blocks that cover each column group (k_gram_sum)
  float4* d_gram_part = nullptr;    // the blocks' partials of y_c
  int* d_info = nullptr;             // [2]: 64 tile + column + 1 of a failed pivot (0: none), replaced pivots
  int* h_info = nullptr;             // pinned [2]
  int64_t factorisations = 0, failures = 0;
  int64_t replaced = 0;              // pivots replaced by the pivot floor, over all factorisations
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
  if (t->d_ysep) (void)hipFree(t->d_ysep);
  if (t->d_blockval) (void)hipFree(t->d_blockval);
  for (void* p : {(void*)q.w, (void*)q.lever, (void*)q.linv, (void*)q.B, (void*)q.BT, (void*)q.Bh, (void*)q.BTh})
    if (p) (void)hipFree(p);
  if (q.h_info) (void)hipHostFree(q.h_info);
  for (hipEvent_t e : {q.ready, q.read, q.done, q.began, q.ended})
    if (e) (void)hipEventDestroy(e);
  if (q.side) (void)hipStreamDestroy(q.side);
  for (void* p : {(void*)t->d_lever, (void*)t->d_sigma, (void*)t->d_cvert, (void*)t->d_adiag, (void*)t->saved.w,
                  (void*)t->saved.lever, (void*)t->d_Ac_lever})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)t->d_parent, (void*)t->d_w, (void*)t->d_wbil, (void*)t->d_fxy, (void*)t->d_theta,
                  (void*)t->d_gnn, (void*)t->d_child_ptr, (void*)t->d_child, (void*)t->d_pair_ptr,
                  (void*)t->d_pair_key, (void*)t->d_pair_terms, (void*)t->d_band, (void*)t->d_linv, (void*)t->d_lpanel,
                  (void*)t->d_border, (void*)t->d_lborder, (void*)t->d_sacc, (void*)t->d_pos, (void*)t->d_ppos, (void*)t->d_Ac, (void*)t->d_B, (void*)t->d_BT, (void*)t->d_Bh, (void*)t->d_BTh,
                  (void*)t->saved.linv, (void*)t->saved.B, (void*)t->saved.BT, (void*)t->saved.Bh, (void*)t->saved.BTh,
                  (void*)t->d_t, (void*)t->d_rc, (void*)t->d_yc, (void*)t->d_info, (void*)t->d_gram,
                  (void*)t->d_gram_part, (void*)t->d_gram_cover})
    if (p) (void)hipFree(p);
  if (t->h_info) (void)hipHostFree(t->h_info);
  delete t;
}

// ---- kernels ----------------------------------------------------------------------------------
// ---- banded FP64 Cholesky and inverse on 64 x 64 tiles ----------------------------------------------
constexpr int TILE = 64;
constexpr int MAX_KB = 8;   // sub-diagonal tiles: coarse grids up to 83 vertices a side
constexpr int MAX_TS = 8;   // separator tiles
constexpr int LSTR = 65;    // LDS row stride of column-accessed tiles (bank spread)

// The factor's ordering of the coarse unknowns: up to three segments, each padded to whole 64-row tiles.
// With the split (rules.coarse_split), the first domain holds the coarse grid rows above a separator of as
// many rows as the coarse operator's coupling reaches, in row order; the second the rows below it, in reverse
// order, so that its rows next to the separator come last; the separator follows both (one level of nested
// dissection). Without it the first segment holds every unknown in row order. A_c and L then have nonzero
// tiles only inside each segment's band and between the separator and the last tiles of each domain (its
// border): the two domains factor side by side, and B = L^-1 is zero between them.
struct Tiles {
  double* band;     // [nt][kb + 1]: tile (k + d, k) inside a segment, of A_c, then the factor's diagonal tiles
  double* lpanel;   // [nt][kb]: factor tile (k + d, k), d = 1..kb, inside a segment
  double* border;   // [2][kb + 1][ts]: tile (separator tile a, border tile bstart[e] + j) of A_c
  double* lborder;  // the same tiles of the factor
  double* sacc;     // [ts][ts]: the second domain's Schur update of the separator block (below the diagonal)
  double* ysep;     // [ts][64][N3p]: -sum over the border tiles K of L(I, K) X(K, J) for separator rows I and the
                    // domain columns J, formed once the domains are factored (null: the rows sum them themselves)
  int kb, nt, ts;
  int seg[4];       // first tile of each segment (first domain, second domain, separator), seg[3] = nt
  int bstart[2];    // first border tile of each domain (= its segment's end without a separator)
  int real[3];      // unknowns of each segment; the rest of its tiles is padding with a unit diagonal
  double* adiag;    // [64 nt] the diagonal of A_c as factored (padding: 1)
  double floor;     // pivot floor: a pivot <= floor * adiag is replaced by adiag (0: such a pivot fails)
};

__device__ inline int seg_of(const Tiles& t, int T) { return T < t.seg[1] ? 0 : (T < t.seg[2] ? 1 : 2); }
__device__ inline bool real_row(const Tiles& t, int R) {
  const int s = seg_of(t, R >> 6);
  return R - TILE * t.seg[s] < t.real[s];
}
__device__ inline double* band_tile(const Tiles& t, int k, int d) {
  return t.band + (((size_t)k * (t.kb + 1) + d) << 12);
}
__device__ inline size_t border_offset(const Tiles& t, int e, int J, int a) {
  return ((size_t)(e * (t.kb + 1) + J - t.bstart[e]) * t.ts + a) << 12;
}
// Tile (I, J), I >= J, of A_c (inside a segment or on a border).
__device__ inline double* a_tile(const Tiles& t, int I, int J) {
  const int sJ = seg_of(t, J);
  if (seg_of(t, I) == sJ) return band_tile(t, J, I - J);
  return t.border + border_offset(t, sJ, J, I - t.seg[2]);
}
// Factor tile (I, J), I > J.
__device__ inline double* l_tile(const Tiles& t, int I, int J) {
  const int sJ = seg_of(t, J);
  if (seg_of(t, I) == sJ) return t.lpanel + (((size_t)J * t.kb + (I - J - 1)) << 12);
  return t.lborder + border_offset(t, sJ, J, I - t.seg[2]);
}
// The tile row of slot j below the diagonal tile of column k, or -1: slots 0 .. kb-1 are the band tiles
// k + 1 + j of k's segment, slots kb .. kb+ts-1 the separator's tiles when k is a border tile of a domain.
__device__ inline int below_tile(const Tiles& t, int k, int j) {
  const int s = seg_of(t, k);
  if (j < t.kb) {
    const int I = k + 1 + j;
    return I < t.seg[s + 1] ? I : -1;
  }
  return (s < 2 && k >= t.bstart[s]) ? t.seg[2] + (j - t.kb) : -1;
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

// One Galerkin term P_ik^T A_b P_jl of the coarse block pair: the fine block A_b (SELL planes, stride 64) between
// the slots ik and jl, added to the DOF x DOF accumulator in FP64. With rotations the slot blocks are
// [w I | S], S = -[v]x (coarse.hpp), so the pair block is [[wi wj A, wi A Sj], [wj Si^T A, Si^T A Sj]].
// a: the fine 3 x 3 block's first value, its values STRIDE floats apart (64 in the SELL-64 slots, 1 in the
// contiguous copy k_block_values makes)
template <int DOF, int STRIDE = 64>
__device__ inline void galerkin_term(double* acc, const float* __restrict__ a, const float* __restrict__ w,
                                     const float4* __restrict__ lever, uint32_t si, uint32_t sj) {
  if (DOF == 3) {
    const double ww = (double)w[si] * (double)w[sj];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] += ww * (double)a[STRIDE * q];
  } else {
    double A[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) A[q] = (double)a[STRIDE * q];
    const double wi = w[si], wj = w[sj];
    const float4 fi = lever[si], fj = lever[sj];
    const double vi[3] = {fi.x, fi.y, fi.z}, vj[3] = {fj.x, fj.y, fj.z};
    // Sj = -[vj]x and Si^T = [vi]x, row-major
    const double Sj[9] = {0.0, vj[2], -vj[1], -vj[2], 0.0, vj[0], vj[1], -vj[0], 0.0};
    const double Ti[9] = {0.0, -vi[2], vi[1], vi[2], 0.0, -vi[0], -vi[1], vi[0], 0.0};
    double X[9], Y[9];  // A Sj and Si^T A
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        X[3 * r + c] = A[3 * r] * Sj[c] + A[3 * r + 1] * Sj[3 + c] + A[3 * r + 2] * Sj[6 + c];
        Y[3 * r + c] = Ti[3 * r] * A[c] + Ti[3 * r + 1] * A[3 + c] + Ti[3 * r + 2] * A[6 + c];
      }
    const double ww = wi * wj;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[6 * r + c] += ww * A[3 * r + c];
        acc[6 * r + 3 + c] += wi * X[3 * r + c];
        acc[6 * (3 + r) + c] += wj * Y[3 * r + c];
        acc[6 * (3 + r) + 3 + c] += Ti[3 * r] * X[c] + Ti[3 * r + 1] * X[3 + c] + Ti[3 * r + 2] * X[6 + c];
      }
  }
}

// The assembled blocks' values in block order, 9 contiguous floats each, copied from their SELL-64 slots (where a
// block's values lie 64 floats apart) once per factorisation: the Galerkin product reads each fine block once per
// pair of its rows' parent slots (16 times), and reads it whole from here.
__global__ void __launch_bounds__(256) k_block_values(const float* __restrict__ sell_val,
                                                      const uint32_t* __restrict__ block_dst, int nblocks,
                                                      float* __restrict__ out) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nblocks) return;
  const float* a = sell_val + block_dst[b];
  float v[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) v[q] = a[64 * q];
#pragma unroll
  for (int q = 0; q < 9; ++q) out[9 * (size_t)b + q] = v[q];
}

// The coarse block pair (I, J) into the tiles, entries on or below the diagonal of the factor's ordering
// (pos: the position of each coarse node's first unknown): one wavefront per pair, lane l summing the pair's
// terms l, l + 64, ... for all DOF x DOF entries, then a fixed-order wavefront tree, so the sum is
// reproducible. (One thread per entry walking a pair's few hundred terms in sequence was latency-bound: 248 us
// per product at n129.) The fine blocks come from blockval, 9 contiguous floats each (k_block_values).
template <int DOF>
__global__ void __launch_bounds__(256) k_galerkin_band(const uint32_t* __restrict__ pair_ptr,
                                                       const uint32_t* __restrict__ pair_key,
                                                       const uint4* __restrict__ terms, const float* __restrict__ w,
                                                       const float4* __restrict__ lever,
                                                       const float* __restrict__ blockval, int npairs, int Nc,
                                                       const int32_t* __restrict__ pos, Tiles tl) {
  constexpr int NE = DOF * DOF;
  const int p = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= npairs) return;  // uniform per wavefront
  const uint32_t key = pair_key[p];
  const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
  double acc[NE];
#pragma unroll
  for (int q = 0; q < NE; ++q) acc[q] = 0.0;
  const uint32_t t1 = pair_ptr[p + 1];
  for (uint32_t t = pair_ptr[p] + lane; t < t1; t += 64) {
    const uint4 e = terms[t];
    galerkin_term<DOF, 1>(acc, blockval + 9 * (size_t)e.x, w, lever, e.y, e.z);
  }
#pragma unroll
  for (int q = 0; q < NE; ++q)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_down(acc[q], o, 64);
  if (lane == 0) {
    const int pI = pos[I], pJ = pos[J];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int R = pI + q / DOF, C = pJ + q % DOF;
      if (R < C) continue;
      a_tile(tl, R >> 6, C >> 6)[((R & 63) << 6) + (C & 63)] = acc[q];
    }
  }
}

// The rotation lever of every parent slot from the current positions: v = w sigma (x_i - X_I) / H
// (sigma: the network's rotation scale of the slot, 1 when null); empty slots get zero.
__global__ void __launch_bounds__(256) k_lever(const int32_t* __restrict__ parent, const float* __restrict__ w,
                                               const float* __restrict__ sigma, const uint32_t* __restrict__ rvert,
                                               const int32_t* __restrict__ cvert, const float4* __restrict__ pos,
                                               float inv_h, int slots, float4* __restrict__ lever) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= slots) return;
  const int I = parent[s];
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (I >= 0) {
    const float4 xf = pos[rvert[s >> 2]], xc = pos[cvert[I]];
    const float g = w[s] * (sigma ? sigma[s] : 1.0f) * inv_h;
    v = make_float4(g * (xf.x - xc.x), g * (xf.y - xc.y), g * (xf.z - xc.z), 0.0f);
  }
  lever[s] = v;
}

// Diagonal of A_c scaled by `factor`; the padding rows of each segment get a unit diagonal. The result is kept
// in adiag for the pivot floor.
__global__ void k_band_diagonal(Tiles tl, double factor) {
  const int R = blockIdx.x * blockDim.x + threadIdx.x;
  if (R >= tl.nt * TILE) return;
  double* a = band_tile(tl, R >> 6, 0) + ((R & 63) << 6) + (R & 63);
  const double d = real_row(tl, R) ? *a * factor : 1.0;
  *a = d;
  tl.adiag[R] = d;
}

// The second domain's Schur update into the separator block (tiles on or below its diagonal).
__global__ void k_add_separator(Tiles tl) {
  const int tile = blockIdx.x >> 4, e = ((blockIdx.x & 15) << 8) + threadIdx.x;
  const int a = tile / tl.ts, b = tile % tl.ts;
  if (b > a) return;
  band_tile(tl, tl.seg[2] + b, a - b)[e] += tl.sacc[((size_t)tile << 12) + e];
}

// LDS traffic between the lanes of one wavefront: LDS instructions of a wavefront execute in order, so a
// wavefront-scope fence (no instruction, an ordering constraint on the compiler) keeps a store before the
// loads that read it back.
__device__ inline void wave_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// 1/sqrt(x) to double precision: the hardware estimate refined by two Newton steps.
__device__ inline double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  // y += y (1 - x y^2) / 2, twice, with fused products
  y = fma(0.5 * y, fma(-x, y * y, 1.0), y);
  y = fma(0.5 * y, fma(-x, y * y, 1.0), y);
  return y;
}

typedef double v4d __attribute__((ext_vector_type(4)));

// A 16 x 16 block of X Y^T over columns k0 .. k0 + K - 1 of X and Y (16 rows each, in LDS with stride LSTR),
// on one wavefront: v_mfma_f64_16x16x4f64 takes A[i][k] from lane 16 k + i and B[k][j] from lane 16 k + j, and
// returns D[(lane >> 4) + 4 v][lane & 15] in element v.
template <int K>
__device__ inline v4d mfma_xyt(const double* X, const double* Y, int k0) {
  const int l = threadIdx.x & 63;
  const double* xr = X + (l & 15) * LSTR + k0 + (l >> 4);
  const double* yr = Y + (l & 15) * LSTR + k0 + (l >> 4);
  v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < K; k += 4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[k], yr[k], acc, 0, 0, 0);
  return acc;
}

// One step of the factorisation chain for up to two tile columns (one per domain), in one launch:
//  - a panel workgroup per column c = st.c[e]: with p = c - 1 in c's segment, it forms the factor tile
//    L(c, p) = A(c, p) L_pp^-T (A(c, p) already holds every update from the columns before p; L_pp^-1 came
//    from p's panel) and subtracts L(c, p) L(c, p)^T from A(c, c), both on FP64 MFMA; then it factors A(c, c)
//    by 16-column panels: wavefront 0 factors each panel four columns at a time (the 4 x 4 diagonal block
//    broadcast through LDS and factored redundantly by every lane, each lane (= row) solving its four factor
//    entries and broadcasting them for the rank-4 update of the rest of its panel row), and the panel's
//    rank-16 update of the trailing lower triangle runs on MFMA; finally it inverts L_cc blockwise (the four
//    16 x 16 diagonal blocks by forward substitution in parallel, then the off-diagonal blocks by distance
//    from the diagonal, X_ab = -X_aa sum_c L_ac X_cb, on MFMA). A pivot that is not positive (or not finite)
//    records 64 c + column + 1 in *info and stops every later kernel;
//  - worker workgroups for column p = st.p[e] (the previous column of a segment, or its last one after its
//    panel): the factor tiles L(I, p) = A(I, p) L_pp^-T of the below tiles I other than the panel's, and one
//    workgroup per pair of below tiles (I >= J) other than the panel's (c, c), forming L(I, p) and L(J, p)
//    itself and subtracting L(I, p) L(J, p)^T from tile (I, J) (the second domain's separator block goes to
//    the separate Schur block).
// Every tile a workgroup writes is read by no other workgroup of the launch.
struct Step {
  int panels, workers;
  int c[2], p[2];
};
constexpr int STEP_THREADS = 512;
constexpr int STEP_WAVES = STEP_THREADS / 64;

template <int ROWS>
__device__ inline void tile_to_lds_step(const double* __restrict__ src, double* dst) {
  constexpr int PER = ROWS * TILE / STEP_THREADS;
  double v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) v[i] = src[threadIdx.x + STEP_THREADS * i];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + STEP_THREADS * i;
    dst[(e >> 6) * LSTR + (e & 63)] = v[i];
  }
}

// Up to three 64 x 64 tiles into LDS with every load issued before the first store (src2 may be null).
__device__ inline void tiles_to_lds_step(const double* __restrict__ src0, double* dst0, const double* __restrict__ src1,
                                         double* dst1, const double* __restrict__ src2, double* dst2) {
  constexpr int PER = TILE * TILE / STEP_THREADS;
  double v0[PER], v1[PER], v2[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    v0[i] = src0[threadIdx.x + STEP_THREADS * i];
    v1[i] = src1[threadIdx.x + STEP_THREADS * i];
    if (src2) v2[i] = src2[threadIdx.x + STEP_THREADS * i];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + STEP_THREADS * i, o = (e >> 6) * LSTR + (e & 63);
    dst0[o] = v0[i];
    dst1[o] = v1[i];
    if (src2) dst2[o] = v2[i];
  }
}

__device__ inline void lds_to_tile_step(const double* src, double* __restrict__ dst) {
  for (int e = threadIdx.x; e < TILE * TILE; e += STEP_THREADS) dst[e] = src[(e >> 6) * LSTR + (e & 63)];
}

// This wavefront's two 16 x 16 blocks (q = wave, wave + 8; block (q >> 2, q & 3)) of A Linv^T, A and Linv
// 64 x 64 in LDS.
__device__ inline void times_inverse_t(const double* A, const double* Linv, v4d out[2]) {
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = wave + STEP_WAVES * h;
    out[h] = mfma_xyt<TILE>(A + 16 * (q >> 2) * LSTR, Linv + 16 * (q & 3) * LSTR, 0);
  }
}

__device__ inline void store_blocks(const v4d in[2], double* X) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = wave + STEP_WAVES * h;
#pragma unroll
    for (int v = 0; v < 4; ++v) X[(16 * (q >> 2) + (lane >> 4) + 4 * v) * LSTR + 16 * (q & 3) + (lane & 15)] = in[h][v];
  }
}

// X = L^-1 for the lower factor tile L (LDS; rd[i] = 1 / L_ii); X (LDS) must be zero on entry; Y: 16 x 17 x 3
// scratch. The diagonal blocks by forward substitution (wavefront w < 4, lane j < 16 holding column j), then
// the off-diagonal blocks by distance from the diagonal on MFMA, one wavefront per block.
__device__ inline void invert_tile(const double* L, double* X, double* Y, const double* rd) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave < 4 && lane < 16) {
    const int o = 16 * wave;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      double acc[4] = {lane == i ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 0; m < i; ++m) acc[m & 3] = fma(-L[(o + i) * LSTR + o + m], x[m], acc[m & 3]);
      x[i] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) * rd[o + i];
      X[(o + i) * LSTR + o + lane] = x[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int dist = 1; dist < 4; ++dist) {
    if (wave < 4 - dist) {
      const int b = wave, a = b + dist;
      // Y = sum_{c=b}^{a-1} L_ac X_cb: A[i][k] = L[16 a + i][16 b + k'], B[k][j] = X[16 b + k'][16 b + j]
      const double* lr = L + (16 * a + (lane & 15)) * LSTR + 16 * b + (lane >> 4);
      const double* xc = X + (16 * b + (lane >> 4)) * LSTR + 16 * b + (lane & 15);
      v4d y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 16 * dist; k += 4) y = __builtin_amdgcn_mfma_f64_16x16x4f64(lr[k], xc[k * LSTR], y, 0, 0, 0);
      double* yb = Y + wave * 16 * 17;
#pragma unroll
      for (int v = 0; v < 4; ++v) yb[((lane >> 4) + 4 * v) * 17 + (lane & 15)] = y[v];
      wave_order();
      // X_ab = -X_aa Y
      const double* xa = X + (16 * a + (lane & 15)) * LSTR + 16 * a + (lane >> 4);
      const double* yc = yb + (lane >> 4) * 17 + (lane & 15);
      v4d z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 16; k += 4) z = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[k], yc[k * 17], z, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) X[(16 * a + (lane >> 4) + 4 * v) * LSTR + 16 * b + (lane & 15)] = -z[v];
    }
    __syncthreads();
  }
}

// Lane l's copy of lane `src`'s value of v (src wave-uniform): two v_readlane_b32.
__device__ inline double lane_value(double v, int src) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The 16 columns c0 .. c0 + 15 of the panel on wavefront 0 (lane = row c0 + lane, p = its 16 entries) without the
// pivot floor and the failure check, one column at a time: the pivot comes from the lane that owns it, 1/sqrt is
// taken once for the wavefront, the next pivot is formed from two values read ahead of it, and the column's other
// factor entries reach the rows through LDS (colw). The operations, and their order per entry, are those of the
// checked chain when it replaces no pivot, so the results are the same bits. Each column's pivot goes to piv[] and
// its reciprocal square root to rd[]; the caller checks the pivots afterwards and factors the panel again with the
// checked chain when one of them would have been floored or is not positive and finite.
__device__ inline void panel_columns_unchecked(double (&p)[16], double* colw, double* piv, double* rd) {
  const int lane = threadIdx.x & 63;
  double w = lane_value(p[0], 0);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double pj = w;
    const double r = rsqrt_f64(pj);
    piv[j] = pj;
    rd[j] = r;
    const double lj = p[j] * r;
    if (j + 1 < 16) {
      const double l1 = lane_value(p[j], j + 1) * r;  // the factor entry of row j + 1
      w = fma(-l1, l1, lane_value(p[j + 1], j + 1));  // the next pivot, as row j + 1 forms it below
      colw[j * TILE + lane] = lj;
      p[j + 1] = fma(-lj, l1, p[j + 1]);
#pragma unroll
      for (int k = j + 2; k < 16; ++k) p[k] = fma(-lj, colw[j * TILE + k], p[k]);
    }
    p[j] = lj;
  }
}

__global__ void __launch_bounds__(STEP_THREADS) k_chol_step(Tiles t, double* __restrict__ linv, Step st, int* info) {
  __shared__ double sa[TILE * LSTR];
  __shared__ double sb[TILE * LSTR];
  __shared__ double sc[TILE * LSTR];
  __shared__ double sy[3 * 16 * 17];
  __shared__ double blk[16];
  __shared__ double colbuf[4 * TILE];
  __shared__ double colw[15 * TILE];  // the unchecked chain's factor columns
  __shared__ double pivs[16];         // the unchecked chain's pivots
  __shared__ double rdiag[TILE];
  __shared__ double adg[TILE];  // the panel's diagonal of A_c for the pivot floor (read in the pivot chain)
  __shared__ int fail;
  if (*info != 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int slots = t.kb + t.ts, per = slots + slots * (slots + 1) / 2;
  const int bid = blockIdx.x;
  if (bid < st.panels) {
    const int k = st.c[bid];
    const int prev = k - 1 >= t.seg[seg_of(t, k)] ? k - 1 : -1;
    double* tile = band_tile(t, k, 0);
    double* Ls = sc;
    if (t.floor > 0.0 && tid < TILE) adg[tid] = t.adiag[TILE * k + tid];  // visible after the next barrier
    if (prev >= 0) {
      tiles_to_lds_step(linv + ((size_t)prev << 12), sa, a_tile(t, k, prev), sb, tile, sc);
      __syncthreads();
      v4d lk[2];
      times_inverse_t(sb, sa, lk);  // L(k, prev)
      __syncthreads();
      store_blocks(lk, sb);
      __syncthreads();
      // A(k, k) -= L(k, prev) L(k, prev)^T on the ten blocks on or below the diagonal
      for (int q = wave; q < 10; q += STEP_WAVES) {
        const int bi = q < 1 ? 0 : q < 3 ? 1 : q < 6 ? 2 : 3;
        const int bj = q - bi * (bi + 1) / 2;
        const v4d acc = mfma_xyt<TILE>(sb + 16 * bi * LSTR, sb + 16 * bj * LSTR, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ls[(16 * bi + (lane >> 4) + 4 * v) * LSTR + 16 * bj + (lane & 15)] -= acc[v];
      }
    } else {
      tile_to_lds_step<TILE>(tile, sc);
    }
    // sa becomes L_kk^-1
    for (int e = tid; e < TILE * TILE; e += STEP_THREADS) sa[(e >> 6) * LSTR + (e & 63)] = 0.0;
    if (tid == 0) fail = 0;
    __syncthreads();
    int nrep = 0;  // pivots replaced by the floor (wavefront 0, the same on every lane)
    for (int b = 0; b < 4; ++b) {
      const int c0 = 16 * b;
      if (wave == 0) {
      const int row = c0 + lane;
      const bool act = row < TILE;
      double p[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) p[j] = act ? Ls[row * LSTR + c0 + j] : 0.0;
      int bad = 0;
      panel_columns_unchecked(p, colw, pivs, rdiag + c0);
      // lane l < 16 checks column l's pivot as the checked chain would
      bool trip = false;
      if (lane < 16) {
        const double pv = pivs[lane];
        if (t.floor > 0.0) {
          const double a = adg[c0 + lane];
          trip = !(pv > t.floor * a) && pv > -INFINITY && a > 0.0 && a < INFINITY;
        }
        trip = trip || !(pv > 0.0) || !(pv < INFINITY);
      }
      // the checked chain, from the panel as it was, when a pivot is floored or fails
      if (__any(trip)) {
#pragma unroll
      for (int j = 0; j < 16; ++j) p[j] = act ? Ls[row * LSTR + c0 + j] : 0.0;
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
            for (int m = 0; m < v; ++m) piv = fma(-l[v][m], l[v][m], piv);
            if (t.floor > 0.0) {
              // modified Cholesky: a collapsed or negative pivot becomes the column's diagonal entry of A_c
              const double a = adg[c0 + q + v];
              if (!(piv > t.floor * a) && piv > -INFINITY && a > 0.0 && a < INFINITY) {
                piv = a;
                ++nrep;
              }
            }
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
              for (int m = 0; m < v; ++m) a = fma(-l[u][m], l[v][m], a);
              l[u][v] = a * rd[v];
            }
          }
          // this lane's factor entries in columns q .. q + 3 (rows below the block), the block's own rows
          double x[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double a = p[q + v];
#pragma unroll
            for (int m = 0; m < v; ++m) a = fma(-x[m], l[v][m], a);
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
            for (int v = 0; v < 4; ++v) p[j] = fma(-x[v], colbuf[v * TILE + j], p[j]);
          }
          wave_order();
        }
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
      // rank-16 update of rows and columns c0 + 16 .. 63 (16 x 16 blocks on or below the diagonal)
      const int nb = 3 - b;
      for (int q = wave; q < nb * (nb + 1) / 2; q += STEP_WAVES) {
        const int bi = q < 1 ? 0 : q < 3 ? 1 : 2;
        const int bj = q - bi * (bi + 1) / 2;
        const int i0 = c0 + 16 + 16 * bi, j0 = c0 + 16 + 16 * bj;
        const v4d acc = mfma_xyt<16>(Ls + i0 * LSTR, Ls + j0 * LSTR, c0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ls[(i0 + (lane >> 4) + 4 * v) * LSTR + j0 + (lane & 15)] -= acc[v];
      }
      __syncthreads();
    }
    if (wave == 0 && lane == 0 && nrep > 0) atomicAdd(info + 1, nrep);
    // waves 4 .. 7, idle in the blockwise inverse (waves 0 .. 3), store the factor tile (zero above the diagonal)
    // and L(k, prev) meanwhile; then every wave stores the inverse
    if (wave >= 4) {
      constexpr int HALF = STEP_THREADS / 2;
      for (int e = tid - HALF; e < TILE * TILE; e += HALF) {
        const int i = e >> 6, j = e & 63;
        tile[e] = j <= i ? Ls[i * LSTR + j] : 0.0;
      }
      if (prev >= 0) {
        double* lt = l_tile(t, k, prev);
        for (int e = tid - HALF; e < TILE * TILE; e += HALF) lt[e] = sb[(e >> 6) * LSTR + (e & 63)];
      }
    }
    invert_tile(Ls, sa, sy, rdiag);
    lds_to_tile_step(sa, linv + ((size_t)k << 12));
    return;
  }
  const int w = bid - st.panels;
  const int p = st.p[w / per], role = w % per;
  // the panel of column p + 1 in this launch (same segment) owns tile (p + 1, p) and the update of (p + 1, p + 1)
  int excl = -1;
  for (int e = 0; e < st.panels; ++e)
    if (st.c[e] == p + 1 && seg_of(t, p + 1) == seg_of(t, p)) excl = p + 1;
  if (role < slots) {
    const int I = below_tile(t, p, role);
    if (I < 0 || I == excl) return;
    tiles_to_lds_step(linv + ((size_t)p << 12), sa, a_tile(t, I, p), sb, nullptr, nullptr);
    __syncthreads();
    v4d li[2];
    times_inverse_t(sb, sa, li);
    double* out = l_tile(t, I, p);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = wave + STEP_WAVES * h;
#pragma unroll
      for (int v = 0; v < 4; ++v) out[((16 * (q >> 2) + (lane >> 4) + 4 * v) << 6) + 16 * (q & 3) + (lane & 15)] = li[h][v];
    }
    return;
  }
  int i = 0, idx = role - slots;
  while (idx > i) {
    idx -= i + 1;
    ++i;
  }
  const int I = below_tile(t, p, i), J = below_tile(t, p, idx);
  if (I < 0 || J < 0 || (I == excl && J == excl)) return;
  tiles_to_lds_step(linv + ((size_t)p << 12), sa, a_tile(t, J, p), sb, I != J ? a_tile(t, I, p) : nullptr, sc);
  __syncthreads();
  v4d lj[2], li[2];
  times_inverse_t(sb, sa, lj);
  if (I != J) times_inverse_t(sc, sa, li);
  __syncthreads();
  store_blocks(lj, sb);
  if (I != J) store_blocks(li, sc);
  __syncthreads();
  const double* LI = I != J ? sc : sb;
  double* out;
  if (seg_of(t, I) == 2 && seg_of(t, J) == 2 && seg_of(t, p) == 1)
    out = t.sacc + ((size_t)((I - t.seg[2]) * t.ts + (J - t.seg[2])) << 12);
  else
    out = a_tile(t, I, J);
  for (int q = wave; q < 16; q += STEP_WAVES) {
    const int bi = q >> 2, bj = q & 3;
    const v4d acc = mfma_xyt<TILE>(LI + 16 * bi * LSTR, sb + 16 * bj * LSTR, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) out[((16 * bi + (lane >> 4) + 4 * v) << 6) + 16 * bj + (lane & 15)] -= acc[v];
  }
}

constexpr int INV_COLS = 16;
__device__ inline int inverse_first_column(const Tiles& t, int I) { return seg_of(t, I) == 1 ? TILE * t.seg[1] : 0; }
// tile row K of X in its segment's ring
template <typename T>
__device__ inline T* inverse_ring(const Tiles& t, T* X, int N3p, int K) {
  const int sk = seg_of(t, K);
  return X + (size_t)(sk * (t.kb + 1) + (K - t.seg[sk]) % (t.kb + 1)) * TILE * N3p;
}
// acc[c] -= L(I, K) X(K, jb + 16 c ..), c < NB, for this wavefront's 16 rows: FP64 MFMA with the operands read
// from L2 (A[i][k] = L(I, K)[16 bi + i][k0 + k] from lane 16 k + i, B[k][j] = X(K)[k0 + k][jb + 16 c + j] from
// lane 16 k + j), the factor rows read once for the NB column blocks
template <int NB>
__device__ inline void inverse_terms(const Tiles& t, const double* X, int N3p, int I, int K, int jb, v4d (&acc)[NB]) {
  const int lane = threadIdx.x & 63, bi = threadIdx.x >> 6;
  const double* a = l_tile(t, I, K) + (16 * bi + (lane & 15)) * TILE + (lane >> 4);
  const double* x = inverse_ring(t, X, N3p, K) + (size_t)(lane >> 4) * N3p + jb + (lane & 15);
  double av[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) av[q] = -a[4 * q];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    double xv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) xv[q] = x[(size_t)4 * q * N3p + INV_COLS * c];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], xv[q], acc[c], 0, 0, 0);
  }
}

// Tile rows of B = L^-1 as a sweep over tile rows, so that row I can run as soon as the panel of tile column I
// has produced L_II^-1 (the factor tiles L(I, K) came from earlier TRSMs): X(I, J) = L_II^-1 (E(I, J) -
// sum_K L(I, K) X(K, J)) over the structurally nonzero X(K, J): K >= J in I's segment band, and for a separator
// row the border tiles of J's domain (summed beforehand by k_separator_border when t.ysep is set). One launch
// serves tile row rows.x (the first w1 workgroups) and rows.y; 16 columns per workgroup, over the row's nonzero
// columns (the second domain's rows start at its first column): short-lived workgroups, which leave the chain's
// steps their CUs (64-column workgroups, a quarter of the factor-tile reads, held CUs long enough to delay the
// chain by more than they saved). X keeps the last kb + 1 tile rows of each segment in FP64
// ([3][kb + 1][64][N3p]); B (row-major), B^T and, when Bh is not null, the FP16 copies B~ = S^-1 B (S = diag(B),
// from L_II^-1) and its transpose get the entries between real unknowns. Wavefront w forms rows 16 w .. 16 w + 15
// with FP64 MFMA (a fixed order).
__global__ void __launch_bounds__(256) k_inverse_row(Tiles t, const double* __restrict__ linv, int2 rows, int w1,
                                                     int N3p, double* __restrict__ X, float* __restrict__ B,
                                                     float* __restrict__ BT, _Float16* __restrict__ Bh,
                                                     _Float16* __restrict__ BTh, const int* info) {
  __shared__ double Ts[TILE * 17];
  if (*info != 0) return;
  const bool second = (int)blockIdx.x >= w1;
  const int I = second ? rows.y : rows.x;
  const int s = seg_of(t, I);
  const int lane = threadIdx.x & 63, bi = threadIdx.x >> 6;  // wavefront bi: rows 16 bi .. 16 bi + 15
  const int j0 = inverse_first_column(t, I) + INV_COLS * (second ? blockIdx.x - w1 : blockIdx.x), J = j0 / TILE;
  // acc = -sum_K L(I, K) X(K, j0 ..) for this wavefront's 16 rows
  v4d acc[1] = {v4d{0.0, 0.0, 0.0, 0.0}};
  if (s == 2 && seg_of(t, J) < 2) {
    if (t.ysep) {
      // the border sum k_separator_border formed, in the order the loop below would have formed it
      const double* y = t.ysep + (size_t)(I - t.seg[2]) * TILE * N3p;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[0][v] = y[(size_t)(16 * bi + (lane >> 4) + 4 * v) * N3p + j0 + (lane & 15)];
    } else {
      const int sJ = seg_of(t, J);
      for (int K = max(t.bstart[sJ], J); K < t.seg[sJ + 1]; ++K) inverse_terms(t, X, N3p, I, K, j0, acc);
    }
  }
  for (int d = 1; d <= t.kb && I - d >= t.seg[s] && I - d >= J; ++d) inverse_terms(t, X, N3p, I, I - d, j0, acc);
  // T = E(I, j0 ..) + acc in LDS, then X(I, j0 ..) = L_II^-1 T
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 16 * bi + (lane >> 4) + 4 * v, col = j0 + (lane & 15);
    Ts[r * 17 + (lane & 15)] = (col == TILE * I + r ? 1.0 : 0.0) + acc[0][v];
  }
  __syncthreads();
  const double* li = linv + ((size_t)I << 12) + (16 * bi + (lane & 15)) * TILE + (lane >> 4);
  double lv[16], tv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    lv[q] = li[4 * q];
    tv[q] = Ts[(4 * q + (lane >> 4)) * 17 + (lane & 15)];
  }
  v4d out = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 16; ++q) out = __builtin_amdgcn_mfma_f64_16x16x4f64(lv[q], tv[q], out, 0, 0, 0);
  double* xo = inverse_ring(t, X, N3p, I);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 16 * bi + (lane >> 4) + 4 * v, C = j0 + (lane & 15), R = TILE * I + r;
    xo[(size_t)r * N3p + C] = out[v];
    if (C <= R && real_row(t, R) && real_row(t, C)) {
      B[(size_t)R * N3p + C] = (float)out[v];
      if (BT) BT[(size_t)C * N3p + R] = (float)out[v];
      if (Bh) {
        const double sr = linv[((size_t)I << 12) + r * 65];
        const _Float16 h = (_Float16)fmin(fmax(out[v] / sr, -65504.0), 65504.0);
        Bh[(size_t)R * N3p + C] = h;
        BTh[(size_t)C * N3p + R] = h;
      }
    }
  }
}

// The separator rows' border sums over the domain columns, Y(I, J) = -sum_{K >= J} L(I, K) X(K, J) over the
// border tiles K of J's domain, for every separator tile row I at once (ts * seg[2] workgroups of one 64-column
// tile; the factor tiles are read once per tile of Y, half the time of 16-column workgroups): the terms
// k_inverse_row would sum first for these columns, in its order, so that the rows, which must follow one another,
// only add the separator's own terms. It runs on the inverse stream once every domain tile row
// of L^-1 (and so every border row of X in the rings) is formed, beside the separator's steps of the chain.
__global__ void __launch_bounds__(256) k_separator_border(Tiles t, int N3p, const double* __restrict__ X,
                                                          const int* info) {
  if (*info != 0) return;
  const int I = t.seg[2] + (int)blockIdx.x / t.seg[2], J = (int)blockIdx.x % t.seg[2], jb = TILE * J;
  const int lane = threadIdx.x & 63, bi = threadIdx.x >> 6;
  const int sJ = seg_of(t, J);
  constexpr int NB = TILE / INV_COLS;
  v4d acc[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int K = max(t.bstart[sJ], J); K < t.seg[sJ + 1]; ++K) inverse_terms(t, X, N3p, I, K, jb, acc);
  double* y = t.ysep + (size_t)(I - t.seg[2]) * TILE * N3p + jb + (lane & 15);
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) y[(size_t)(16 * bi + (lane >> 4) + 4 * v) * N3p + INV_COLS * c] = acc[c][v];
}

// dense A_c (both triangles, row-major, DOF unknowns per coarse node) for the check entry point
template <int DOF>
__global__ void __launch_bounds__(256) k_galerkin_dense(const uint32_t* __restrict__ pair_ptr,
                                                        const uint32_t* __restrict__ pair_key,
                                                        const uint4* __restrict__ terms, const float* __restrict__ w,
                                                        const float4* __restrict__ lever,
                                                        const float* __restrict__ sell_val,
                                                        const uint32_t* __restrict__ block_dst, int npairs, int Nc,
                                                        double* __restrict__ Ac) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double acc[DOF * DOF];
#pragma unroll
  for (int q = 0; q < DOF * DOF; ++q) acc[q] = 0.0;
  for (uint32_t t = pair_ptr[p]; t < pair_ptr[p + 1]; ++t) {
    const uint4 e = terms[t];
    galerkin_term<DOF>(acc, sell_val + block_dst[e.x], w, lever, e.y, e.z);
  }
  const uint32_t key = pair_key[p];
  const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
  const size_t N = (size_t)DOF * Nc;
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int c = 0; c < DOF; ++c) Ac[(DOF * (size_t)I + r) * N + DOF * (size_t)J + c] = acc[DOF * r + c];
}

// The restriction's fixed shuffle tree and store: r_c at the node's position P, three translation sums and,
// with rotations, three rotation sums.
template <bool ROT>
__device__ inline void restrict_store(float* a, int lane, int P, float* __restrict__ rc) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < (ROT ? 6 : 3); ++q) a[q] = a[q] + __shfl_down(a[q], o, 64);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < (ROT ? 6 : 3); ++q) rc[P + q] = a[q];
}

// r_c[I] = sum over the children (i, k) of I of P_ik^T r_i at the node's position in the factor's ordering
// (pos): w_ik r_i, and with rotations v_ik x r_i; one wavefront per coarse node, lane j takes child j (a node
// has at most 49), then a fixed shuffle tree.
template <bool ROT>
__global__ void __launch_bounds__(256) k_restrict(const uint32_t* __restrict__ child_ptr,
                                                  const uint32_t* __restrict__ child, const float* __restrict__ w,
                                                  const float4* __restrict__ lever,
                                                  const int32_t* __restrict__ pos, const float4* __restrict__ r,
                                                  int Nc, float* __restrict__ rc, const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (I >= Nc) return;
  const uint32_t t0 = child_ptr[I], n = child_ptr[I + 1] - t0;
  float a[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t j = lane; j < n; j += 64) {
    const uint32_t s = child[t0 + j];
    const float ww = w[s];
    const float4 v = r[s >> 2];
    a[0] = a[0] + ww * v.x;
    a[1] = a[1] + ww * v.y;
    a[2] = a[2] + ww * v.z;
    if (ROT) {
      const float3 c = lever_cross(lever[s], v.x, v.y, v.z);
      a[3] = a[3] + c.x;
      a[4] = a[4] + c.y;
      a[5] = a[5] + c.z;
    }
  }
  restrict_store<ROT>(a, lane, pos[I], rc);
}

// Row ranges of B = L^-1 in the factor's ordering: d2 and sep are the first positions of the second domain
// and of the separator (both N3p without the split), r1 .. r3 the real unknowns of the three segments. A row
// of B (LOWER) spans [0, R], or [d2, R] in the second domain; a row of B^T spans [C, N3p), without the second
// domain [d2, sep) when C lies in the first; padding rows are skipped (their y stays zero).
struct Split {
  int d2, sep, r1, r2, r3;
};
__device__ inline bool split_real(const Split& p, int R) {
  return R < p.d2 ? R < p.r1 : (R < p.sep ? R - p.d2 < p.r2 : R - p.sep < p.r3);
}
// [lo, hi) in units of `width` elements, without [hs, he)
template <bool LOWER>
__device__ inline void split_range(const Split& p, int row, int N3p, int width, int* lo, int* hi, int* hs, int* he) {
  if (LOWER) {
    *lo = (row >= p.d2 && row < p.sep) ? p.d2 / width : 0;
    *hi = row / width + 1;
    *hs = *he = *hi;
  } else {
    *lo = row / width;
    *hi = N3p / width;
    *hs = row < p.d2 ? p.d2 / width : *hi;
    *he = row < p.d2 ? p.sep / width : *hi;
  }
}

// y = T x for a triangular T stored row-major in full (zeros outside the structure): one wavefront per row
// over the float4 groups of its range, lane sums then a fixed shuffle tree. A lane's loads go out GEMV_DEPTH at a
// time; its elements are summed in index order whatever the depth, so the depth changes no result.
constexpr int GEMV_DEPTH = 8;
template <bool LOWER>
__global__ void __launch_bounds__(256) k_tri_gemv(const float* __restrict__ T, const float* __restrict__ x, Split sp,
                                                  int N3p, float* __restrict__ y, const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int lane = threadIdx.x & 63;
  // longest rows first: a row of B (LOWER) grows with its index, so the last rows are dispatched first and the
  // short ones fill in behind them (a row of B^T shrinks with its index: natural order)
  const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int row = LOWER ? N3p - 1 - slot : slot;
  if (row < 0 || row >= N3p || !split_real(sp, row)) return;
  const float4* m = reinterpret_cast<const float4*>(T + (size_t)row * N3p);
  const float4* v = reinterpret_cast<const float4*>(x);
  int lo, hi, hs, he;
  split_range<LOWER>(sp, row, N3p, 4, &lo, &hi, &hs, &he);
  const int count = hi - lo - (he - hs);
  float acc = 0.0f;
  // GEMV_DEPTH independent loads of B (and of x) in flight per lane before the sequential accumulation
  for (int v0 = lane; v0 < count; v0 += GEMV_DEPTH * 64) {
    float4 a[GEMV_DEPTH], b[GEMV_DEPTH];
#pragma unroll
    for (int u = 0; u < GEMV_DEPTH; ++u) {
      const int vv = v0 + 64 * u;
      if (vv < count) {
        const int j = lo + vv < hs ? lo + vv : lo + vv + (he - hs);
        a[u] = m[j];
        b[u] = v[j];
      }
    }
#pragma unroll
    for (int u = 0; u < GEMV_DEPTH; ++u)
      if (v0 + 64 * u < count) acc = acc + (a[u].x * b[u].x + a[u].y * b[u].y + a[u].z * b[u].z + a[u].w * b[u].w);
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
                                                    const double* __restrict__ linv, Split sp, int N3p,
                                                    float* __restrict__ y, const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= N3p || !split_real(sp, row)) return;
  const uint4* m = reinterpret_cast<const uint4*>(T + (size_t)row * N3p);
  const float4* v = reinterpret_cast<const float4*>(x);
  int lo, hi, hs, he;
  split_range<LOWER>(sp, row, N3p, 8, &lo, &hi, &hs, &he);
  const int count = hi - lo - (he - hs);
  float acc = 0.0f;
  for (int v0 = lane; v0 < count; v0 += 4 * 64) {
    uint4 a[4];
    float4 b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int vv = v0 + 64 * u;
      if (vv < count) {
        const int j = lo + vv < hs ? lo + vv : lo + vv + (he - hs);
        a[u] = m[j];
        b0[u] = v[2 * j];
        b1[u] = v[2 * j + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (v0 + 64 * u < count) {
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

// The coarse solve y_c = B^T (B r_c) in one pass over the stored triangle of B (FP32). Each workgroup owns a
// block of consecutive real rows of one segment; their nonzero columns lie in one range [g0, g1) of float4
// groups (a row R of the second domain spans groups d2 / 4 .. R / 4, any other row 0 .. R / 4). The block's row
// slots (one wavefront each, or a pair when the range exceeds one wavefront's registers) take its rows P at a
// time: per row, t = B_R . r_c (each lane summing its groups in increasing order, a pair adding its two
// wavefronts' sums in order, then a fixed shuffle tree), and the slot adds t B_R to its partial of y in
// registers from the row it still holds, so B is read once. The slots' partials are summed in slot order
// through LDS and stored per block; k_gram_sum adds the blocks' partials per group in block order. Every sum
// has a fixed order, so the result is reproducible.
struct GramBlock {
  int r0, r1;  // rows [r0, r1): real rows of one segment
  int g0, g1;  // the float4 groups of their columns
  int kr;      // groups per lane and row: 1, 2, 3, 4, 6 or 13
  int wpr;     // wavefronts per row slot: 1 or 2
  int off;     // the block's first float4 in the partials
  int pad;
};
constexpr int GRAM_THREADS = 512;
constexpr int GRAM_WAVES = GRAM_THREADS / 64;
constexpr int GRAM_KR_MAX = 13;
constexpr int GRAM_ROW_REGS = 13;                  // float4 registers of a lane for the rows of a chunk
constexpr int GRAM_COMB = GRAM_WAVES * 64 * GRAM_KR_MAX;  // float4: every slot's partial of the widest block
constexpr int GRAM_SUM_LANES = 32;                  // lanes per group in k_gram_sum
constexpr int GRAM_SUM_PER = 16;                    // covering blocks per lane in k_gram_sum, at most
constexpr int GRAM_MAX_BLOCKS = 1024;

__device__ inline float dot4(const float4 a, const float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// DPP steps of a wavefront sum (v + the value the DPP control `ctrl` moves in; lanes outside row_mask add 0)
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
// The sum over the 64 lanes in lane 63, in a fixed order: pairs and quads (quad_perm), rows of 16 (row_ror 4,
// 8), then rows 0 + 1 and 2 + 3 (row_bcast 15) and (2 + 3) + (0 + 1) (row_bcast 31).
__device__ inline float wave_sum_lane63(float v) {
  v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_add<0x124, 0xF>(v);  // row_ror 4
  v = dpp_add<0x128, 0xF>(v);  // row_ror 8
  v = dpp_add<0x142, 0xA>(v);  // row_bcast 15 into rows 1 and 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast 31 into rows 2 and 3
  return v;
}

// The sums over lanes 0-31 and 32-63 in lanes 31 and 63, in a fixed order (the steps of wave_sum_lane63 up to
// row_bcast 15)
__device__ inline float half_wave_sum_lane31(float v) {
  v = dpp_add<0xB1, 0xF>(v);
  v = dpp_add<0x4E, 0xF>(v);
  v = dpp_add<0x124, 0xF>(v);
  v = dpp_add<0x128, 0xF>(v);
  v = dpp_add<0x142, 0xA>(v);
  return v;
}

// v where c holds, else zero: componentwise, so that the select stays in registers (a select of whole vectors
// became a select of their addresses in private memory)
__device__ inline float4 keep_if(bool c, const float4 v) {
  return make_float4(c ? v.x : 0.0f, c ? v.y : 0.0f, c ? v.z : 0.0f, c ? v.w : 0.0f);
}

template <int KR, int WPR>
__device__ inline void gram_block(const GramBlock& gb, const float4* __restrict__ B4, const float4* __restrict__ x4,
                                  int n4, float4* __restrict__ part, float4* comb, float* red) {
  constexpr int P = GRAM_ROW_REGS / KR;     // rows of a slot per chunk
  constexpr int S = GRAM_WAVES / WPR;      // row slots
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: the row branches stay scalar
  const int slot = wave / WPR, half = wave % WPR;
  const int ncols = gb.g1 - gb.g0;
  // A lane reads its groups of every row over the block's whole range: B is stored in full, so the groups past
  // a row's end hold zeros (above the diagonal). The groups past the range (the last lanes) are read from the
  // range's last group, and the lane's r_c there is zero, so they add nothing to t; their y is never stored.
  // Every load is unconditional: a load under a lane condition made the compiler wait for each load before
  // issuing the next.
  int col[KR];
  float4 x[KR], y[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    col[k] = 64 * (WPR * k + half) + lane;
    x[k] = x4[gb.g0 + min(col[k], ncols - 1)];
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    x[k] = keep_if(col[k] < ncols, x[k]);
    y[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const int chunks = (gb.r1 - gb.r0 + S * P - 1) / (S * P);
  // the rows of chunk ch, loaded unconditionally (a row past the block's end reads its last row again, and
  // live marks the rows that exist; wave-uniform)
  auto load = [&](int ch, float4 (&a)[P][KR], bool (&live)[P]) {
    const int rbase = gb.r0 + (ch * S + slot) * P;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      live[p] = rbase + p < gb.r1;
      const float4* row = B4 + (size_t)min(rbase + p, gb.r1 - 1) * n4 + gb.g0;
#pragma unroll
      for (int k = 0; k < KR; ++k) a[p][k] = row[min(col[k], ncols - 1)];
    }
  };
  // t = B_R . r_c per row (a lane's groups in order, fused multiply-adds, then the DPP tree), y += t B_R
  auto process = [&](int ch, const float4 (&a)[P][KR], const bool (&live)[P]) {
    float t[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float acc = 0.0f;
      if (live[p]) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
          acc = __builtin_fmaf(a[p][k].x, x[k].x, acc);
          acc = __builtin_fmaf(a[p][k].y, x[k].y, acc);
          acc = __builtin_fmaf(a[p][k].z, x[k].z, acc);
          acc = __builtin_fmaf(a[p][k].w, x[k].w, acc);
        }
      }
      t[p] = wave_sum_lane63(acc);
    }
    if (WPR == 1) {
#pragma unroll
      for (int p = 0; p < P; ++p) t[p] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t[p]), 63));
    } else {
      // the pair's sums through LDS, double-buffered by chunk so that one barrier per chunk orders them
      float* rb = red + (ch & 1) * (2 * S * P);
      if (lane == 63)
#pragma unroll
        for (int p = 0; p < P; ++p) rb[(slot * P + p) * 2 + half] = t[p];
      __syncthreads();
#pragma unroll
      for (int p = 0; p < P; ++p) t[p] = rb[(slot * P + p) * 2] + rb[(slot * P + p) * 2 + 1];
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (live[p]) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
          y[k].x = __builtin_fmaf(t[p], a[p][k].x, y[k].x);
          y[k].y = __builtin_fmaf(t[p], a[p][k].y, y[k].y);
          y[k].z = __builtin_fmaf(t[p], a[p][k].z, y[k].z);
          y[k].w = __builtin_fmaf(t[p], a[p][k].w, y[k].w);
        }
      }
  };
  // two row buffers in turn: the next chunk's rows go out before this one is reduced
  float4 a0[P][KR], a1[P][KR];
  bool l0[P], l1[P];
  load(0, a0, l0);
  for (int ch = 0; ch < chunks; ch += 2) {
    load(ch + 1, a1, l1);
    process(ch, a0, l0);
    load(ch + 2, a0, l0);
    if (ch + 1 < chunks) process(ch + 1, a1, l1);
  }
  // the slots' partials summed in slot order
  const int stride = 64 * WPR * KR;
#pragma unroll
  for (int k = 0; k < KR; ++k)
    if (col[k] < ncols) comb[slot * stride + col[k]] = y[k];
  __syncthreads();
  for (int c = threadIdx.x; c < ncols; c += GRAM_THREADS) {
    float4 s = comb[c];
#pragma unroll
    for (int q = 1; q < S; ++q) {
      const float4 v = comb[q * stride + c];
      s.x = s.x + v.x;
      s.y = s.y + v.y;
      s.z = s.z + v.z;
      s.w = s.w + v.w;
    }
    part[gb.off + c] = s;
  }
}

__global__ void __launch_bounds__(GRAM_THREADS) k_gram(const GramBlock* __restrict__ blocks, const float* __restrict__ B,
                                                      const float* __restrict__ x, int N3p, float4* __restrict__ part,
                                                      const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  __shared__ float4 comb[GRAM_COMB];
  __shared__ float red[2 * GRAM_WAVES * GRAM_ROW_REGS];
  const GramBlock gb = blocks[blockIdx.x];
  const float4* B4 = reinterpret_cast<const float4*>(B);
  const float4* x4 = reinterpret_cast<const float4*>(x);
  const int n4 = N3p / 4;
  if (gb.wpr == 2) {
    gram_block<13, 2>(gb, B4, x4, n4, part, comb, red);
    return;
  }
  switch (gb.kr) {
    case 1: gram_block<1, 1>(gb, B4, x4, n4, part, comb, red); break;
    case 2: gram_block<2, 1>(gb, B4, x4, n4, part, comb, red); break;
    case 3: gram_block<3, 1>(gb, B4, x4, n4, part, comb, red); break;
    case 4: gram_block<4, 1>(gb, B4, x4, n4, part, comb, red); break;
    case 6: gram_block<6, 1>(gb, B4, x4, n4, part, comb, red); break;
    default: gram_block<13, 1>(gb, B4, x4, n4, part, comb, red); break;
  }
}

// y_c = the sum of the blocks' partials, per float4 group in block order. The blocks that cover group g are the
// blocks [a, e) of its segment whose rows reach it and then the blocks [s2, nb) (the separator's, which cover every
// column of the domains): cover[g] = (a, e, s2), C of them. GRAM_SUM_LANES lanes per group, lane j adding the
// covering blocks j, j + 32, ... in order, then a fixed DPP tree over the lanes. The loads are issued together
// (a wave-uniform count, clamped indices, values of the lanes past C replaced by zero).
__global__ void __launch_bounds__(256) k_gram_sum(const GramBlock* __restrict__ blocks, const int4* __restrict__ cover,
                                                  int nb, const float4* __restrict__ part, int n4,
                                                  float* __restrict__ y, const PcgState* st, int guard) {
  if (guard == 0 && st->status != 0) return;
  if (guard == 1 && (st->status != 0 || !(st->rr > st->above2 * st->rr0))) return;
  __shared__ int3 tab[GRAM_MAX_BLOCKS];
  for (int b = threadIdx.x; b < nb; b += 256) tab[b] = make_int3(blocks[b].g0, blocks[b].g1, blocks[b].off);
  __syncthreads();
  const int gl = blockIdx.x * (256 / GRAM_SUM_LANES) + (int)threadIdx.x / GRAM_SUM_LANES;
  const int g = min(gl, n4 - 1);
  const int j = threadIdx.x % GRAM_SUM_LANES;
  const int4 cv = cover[g];
  const int first = cv.y - cv.x, count = first + nb - cv.z;
  // loads per lane: the larger count of the wavefront's two groups, wave-uniform
  const int c0 = __builtin_amdgcn_readlane(count, 0), c1 = __builtin_amdgcn_readlane(count, 32);
  const int per = (max(c0, c1) + GRAM_SUM_LANES - 1) / GRAM_SUM_LANES;
  float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (per > 0) {
    float4 v[GRAM_SUM_PER];
#pragma unroll
    for (int i = 0; i < GRAM_SUM_PER; ++i) {
      if (i < per) {
        const int m = min(j + GRAM_SUM_LANES * i, max(count - 1, 0));
        const int3 e = tab[min(m < first ? cv.x + m : cv.z + m - first, nb - 1)];  // (count 0: any block)
        v[i] = part[e.z + min(max(g - e.x, 0), e.y - e.x - 1)];
      }
    }
#pragma unroll
    for (int i = 0; i < GRAM_SUM_PER; ++i) {
      if (i < per) {
        const float4 u = keep_if(j + GRAM_SUM_LANES * i < count, v[i]);
        s.x = s.x + u.x;
        s.y = s.y + u.y;
        s.z = s.z + u.z;
        s.w = s.w + u.w;
      }
    }
  }
  static_assert(GRAM_SUM_LANES == 32, "the lane tree sums halves of a wavefront");
  s.x = half_wave_sum_lane31(s.x);
  s.y = half_wave_sum_lane31(s.y);
  s.z = half_wave_sum_lane31(s.z);
  s.w = half_wave_sum_lane31(s.w);
  if (j == GRAM_SUM_LANES - 1 && gl < n4) reinterpret_cast<float4*>(y)[g] = s;
}

// z = D^-1 r + sum_k w_ik y_c[parent(i, k)], partial (r.z, r.r); mode 2 (setup) also sets p = z. ppos holds the
// position of each parent slot's node in the factor's ordering (-1: empty slot).
__global__ void __launch_bounds__(256) k_prolong(const int32_t* __restrict__ ppos, const float* __restrict__ w,
                                                 const float4* __restrict__ lever,
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
    coarse_prolong_add(ppos, w, lever, yc, row, z0, z1, z2);
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
template <bool ROT>
__global__ void __launch_bounds__(256) k_restrict_update(const uint32_t* __restrict__ child_ptr,
                                                         const uint32_t* __restrict__ child,
                                                         const float* __restrict__ w,
                                                         const float4* __restrict__ lever,
                                                         const int32_t* __restrict__ pos,
                                                         const float4* __restrict__ r,
                                                         const float4* __restrict__ q, int Nc, float* __restrict__ rc,
                                                         const float* __restrict__ partial_pq, int G, PcgState* st,
                                                         int it) {
  const int entry = st->status;  // first: a launch queued past the solve's stop returns after this one load
  float raw[4];
  partials_start(partial_pq, G, raw);  // before the children's loads: loads complete in issue order
  const float rzv = st->rz[it & 1];
  if (entry != 0) return;
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // each lane's first child (a node has at most 49) is loaded before alpha is known, and the partials' tree runs
  // while those loads are in flight
  uint32_t t0 = 0, n = 0, s0 = 0;
  float w0 = 0.0f;
  float4 r0 = make_float4(0, 0, 0, 0), q0 = r0, l0 = r0;
  if (I < Nc) {
    t0 = child_ptr[I];
    n = child_ptr[I + 1] - t0;
  }
  if ((uint32_t)lane < n) {
    s0 = child[t0 + lane];
    w0 = w[s0];
    r0 = r[s0 >> 2];
    q0 = q[s0 >> 2];
    if (ROT) l0 = lever[s0];
  }
  const float pq = partials_finish(partial_pq, G, raw);  // every lane: the shuffles need the whole wavefront
  const float alpha = rzv / pq;
  if (!(pq > 0.0f && finite32(alpha))) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->status = 3;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq;
    st->alpha = alpha;
  }
  if (I >= Nc) return;
  float a[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if ((uint32_t)lane < n) {
    const float rx = r0.x - alpha * q0.x, ry = r0.y - alpha * q0.y, rz = r0.z - alpha * q0.z;
    a[0] = a[0] + w0 * rx;
    a[1] = a[1] + w0 * ry;
    a[2] = a[2] + w0 * rz;
    if (ROT) {
      const float3 c = lever_cross(l0, rx, ry, rz);
      a[3] = a[3] + c.x;
      a[4] = a[4] + c.y;
      a[5] = a[5] + c.z;
    }
  }
  for (uint32_t j = lane + 64; j < n; j += 64) {
    const uint32_t s = child[t0 + j];
    const float ww = w[s];
    const float4 v = r[s >> 2], qq = q[s >> 2];
    const float rx = v.x - alpha * qq.x, ry = v.y - alpha * qq.y, rz = v.z - alpha * qq.z;
    a[0] = a[0] + ww * rx;
    a[1] = a[1] + ww * ry;
    a[2] = a[2] + ww * rz;
    if (ROT) {
      const float3 c = lever_cross(lever[s], rx, ry, rz);
      a[3] = a[3] + c.x;
      a[4] = a[4] + c.y;
      a[5] = a[5] + c.z;
    }
  }
  restrict_store<ROT>(a, lane, pos[I], rc);
}

__global__ void __launch_bounds__(256) k_prolong_update(const int32_t* __restrict__ ppos, const float* __restrict__ w,
                                                        const float4* __restrict__ lever,
                                                        const float* __restrict__ yc, const float* __restrict__ dinv,
                                                        const float4* __restrict__ p, const float4* __restrict__ q,
                                                        float4* __restrict__ x, float4* __restrict__ r,
                                                        float4* __restrict__ z, float* __restrict__ partials, int R,
                                                        const float* __restrict__ partial_pq, int G,
                                                        const PcgState* st, int it) {
  __shared__ float2 lds[256];
  const int entry = st->status;  // first: a launch queued past the solve's stop returns after this one load
  float raw[4];
  partials_start(partial_pq, G, raw);  // before the row's loads: loads complete in issue order
  const float rzv = st->rz[it & 1];
  if (entry != 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  // every operand that does not depend on alpha is loaded before the p.q partials are reduced, and the tree
  // runs while the coarse solution's gathers are in flight
  float4 pr = make_float4(0, 0, 0, 0), qr = pr, xr = pr, rv = pr;
  float d[9];
  SlotOperands so;
  if (row < R) {
    pr = p[row];
    qr = q[row];
    xr = x[row];
    rv = r[row];
#pragma unroll
    for (int e = 0; e < 9; ++e) d[e] = dinv[9 * row + e];
    coarse_prolong_load(ppos, w, lever, yc, row, so);
  }
  const float pq = partials_finish(partial_pq, G, raw);  // every lane: the shuffles need the whole wavefront
  const float alpha = rzv / pq;
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
    float z0 = d[0] * rv.x + d[1] * rv.y + d[2] * rv.z;
    float z1 = d[3] * rv.x + d[4] * rv.y + d[5] * rv.z;
    float z2 = d[6] * rv.x + d[7] * rv.y + d[8] * rv.z;
    coarse_prolong_apply(so, lever != nullptr, z0, z1, z2);
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
// The rows of the fused coarse solve in blocks of about equal cost (the groups a row reads, plus
// GRAM_ROW_COST for its fixed work), inside each segment's real rows; false when a block's columns exceed what
// k_gram holds in registers (then the coarse solve runs as two GEMVs).
constexpr int GRAM_TARGET_BLOCKS = 256;
constexpr int GRAM_ROW_COST = 32;
static bool gram_partition(const TwoLevel* t, std::vector<GramBlock>& out, std::vector<int4>& cover, int* part4) {
  struct Seg {
    int first, count, lo;
  };
  const Seg segs[3] = {{0, t->real[0], 0},
                       {TILE * t->seg[1], t->real[1], TILE * t->seg[1] / 4},
                       {TILE * t->seg[2], t->real[2], 0}};
  double total = 0.0;
  for (const Seg& g : segs)
    for (int R = g.first; R < g.first + g.count; ++R) total += R / 4 + 1 - g.lo + GRAM_ROW_COST;
  const double target = total / GRAM_TARGET_BLOCKS;
  out.clear();
  int off = 0;
  for (const Seg& g : segs) {
    int start = g.first;
    double acc = 0.0;
    for (int R = g.first; R < g.first + g.count; ++R) {
      acc += R / 4 + 1 - g.lo + GRAM_ROW_COST;
      if (acc < target && R + 1 < g.first + g.count) continue;
      GramBlock b{};
      b.r0 = start;
      b.r1 = R + 1;
      b.g0 = g.lo;
      b.g1 = R / 4 + 1;
      const int ncols = b.g1 - b.g0;
      b.wpr = ncols <= 64 * GRAM_KR_MAX ? 1 : 2;
      if (ncols > 64 * b.wpr * GRAM_KR_MAX) return false;
      const int need = (ncols + 64 * b.wpr - 1) / (64 * b.wpr);
      b.kr = b.wpr == 2 ? GRAM_KR_MAX : (need <= 4 ? need : (need <= 6 ? 6 : GRAM_KR_MAX));
      b.off = off;
      off += ncols;
      out.push_back(b);
      start = R + 1;
      acc = 0.0;
    }
  }
  *part4 = off;
  const int nb = (int)out.size();
  if (nb == 0 || nb > GRAM_MAX_BLOCKS) return false;
  // the covering blocks of each column group: a suffix [a, e) of its segment's blocks, then the separator's
  // blocks [s2, nb) for the columns of the domains (they span every column below the separator)
  int segb[4] = {0, 0, 0, nb};
  for (int q = 1; q < 3; ++q) {
    segb[q] = segb[q - 1];
    while (segb[q] < nb && out[segb[q]].r0 < segs[q].first) ++segb[q];
  }
  const int n4 = t->N3p / 4;
  cover.assign(n4, make_int4(0, 0, nb, 0));
  for (int g = 0; g < n4; ++g) {
    const int q = 4 * g < segs[1].first ? 0 : (4 * g < segs[2].first ? 1 : 2);
    int a = segb[q];
    while (a < segb[q + 1] && out[a].g1 <= g) ++a;
    cover[g] = make_int4(a, segb[q + 1], q < 2 ? segb[2] : nb, 0);
    int count = cover[g].y - cover[g].x + nb - cover[g].z;
    for (int b = cover[g].x; b < cover[g].y; ++b)
      if (g < out[b].g0 || g >= out[b].g1) return false;
    for (int b = cover[g].z; b < nb; ++b)
      if (g < out[b].g0 || g >= out[b].g1) return false;
    for (int b = 0; b < nb; ++b)
      count -= (g >= out[b].g0 && g < out[b].g1) ? 1 : 0;
    if (count != 0 || cover[g].y - cover[g].x + nb - cover[g].z > GRAM_SUM_LANES * GRAM_SUM_PER) return false;
  }
  return true;
}

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
  t->dof = c.rules.coarse_rotations ? 6 : 3;
  const int dof = t->dof;
  t->N3 = dof * Nc;
  std::vector<int32_t> cvert(Nc);
  for (int b = 0; b < nc; ++b)
    for (int a = 0; a < nc; ++a)
      if (coarse_of[b * nc + a] >= 0) cvert[coarse_of[b * nc + a]] = (f * b) * n + f * a;
  {
    // the coarse grid spacing of the rest shape: vertex 0 to vertex f of the first grid row
    const float* r0 = c.rest_host.data();
    const double dx = (double)r0[3 * f] - r0[0], dy = (double)r0[3 * f + 1] - r0[1], dz = (double)r0[3 * f + 2] - r0[2];
    const double h = std::sqrt(dx * dx + dy * dy + dz * dz);
    t->inv_spacing = h > 0.0 ? (float)(1.0 / h) : 1.0f;
  }
  std::vector<int> node_b(Nc);
  for (int b = 0; b < nc; ++b)
    for (int a = 0; a < nc; ++a)
      if (coarse_of[b * nc + a] >= 0) node_b[coarse_of[b * nc + a]] = b;
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
  // The factor's ordering. The separator is as many coarse grid rows as the pairs reach across rows (two for
  // the hinges' two-cell stencil); the first domain takes the rows above it in row order, the second the rows
  // below it in reverse order. A layout whose separator or borders outgrow the tile structure falls back to
  // one segment in row order.
  int reach = 0;
  for (uint32_t key : pair_key)
    reach = std::max(reach, std::abs(node_b[key / (uint32_t)Nc] - node_b[key % (uint32_t)Nc]));
  std::vector<int32_t> pos(Nc);
  auto layout = [&](bool split) -> bool {
    const int m0 = (nc - reach) / 2;
    split = split && reach >= 1 && m0 >= 1 && nc - m0 - reach >= 1;
    auto seg_of_row = [&](int b) { return !split ? 0 : (b < m0 ? 0 : (b < m0 + reach ? 2 : 1)); };
    int count[3] = {0, 0, 0};
    std::vector<int> idx(Nc);
    for (int b = 0; b < nc; ++b)  // rows in order (the second domain is numbered from its far end below)
      for (int a = 0; a < nc; ++a) {
        const int I = coarse_of[b * nc + a];
        if (I >= 0 && seg_of_row(b) != 1) idx[I] = count[seg_of_row(b)]++;
      }
    for (int b = nc - 1; b >= 0; --b)
      for (int a = nc - 1; a >= 0; --a) {
        const int I = coarse_of[b * nc + a];
        if (I >= 0 && seg_of_row(b) == 1) idx[I] = count[1]++;
      }
    int tiles[3];
    for (int g = 0; g < 3; ++g) {
      t->real[g] = dof * count[g];
      tiles[g] = (t->real[g] + TILE - 1) / TILE;
    }
    t->seg[0] = 0;
    t->seg[1] = tiles[0];
    t->seg[2] = tiles[0] + tiles[1];
    t->seg[3] = t->nt = tiles[0] + tiles[1] + tiles[2];
    t->ts = tiles[2];
    for (int I = 0; I < Nc; ++I) pos[I] = TILE * t->seg[seg_of_row(node_b[I])] + dof * idx[I];
    auto seg_of_tile = [&](int T) { return T < t->seg[1] ? 0 : (T < t->seg[2] ? 1 : 2); };
    t->kb = 0;
    t->bstart[0] = t->seg[1];
    t->bstart[1] = t->seg[2];
    for (uint32_t key : pair_key) {
      const int I = (int)(key / (uint32_t)Nc), J = (int)(key % (uint32_t)Nc);
      for (int q = 0; q < dof * dof; ++q) {
        const int R = pos[I] + q / dof, C = pos[J] + q % dof;
        if (R < C) continue;
        const int sR = seg_of_tile(R / TILE), sC = seg_of_tile(C / TILE);
        if (sR == sC) t->kb = std::max(t->kb, R / TILE - C / TILE);
        else if (sR == 2) t->bstart[sC] = std::min(t->bstart[sC], C / TILE);
        else return false;  // the domains touch: no separator
      }
    }
    if (t->ts > MAX_TS || (t->ts > 0 && t->ts - 1 > t->kb)) return false;
    for (int e = 0; e < 2; ++e)
      if (t->seg[e + 1] - t->bstart[e] > t->kb + 1) return false;
    return true;
  };
  if (!layout(c.rules.coarse_split != 0) && !layout(false)) {
    set_error("two-level preconditioner: no tile layout for the coarse operator");
    delete t;
    return DRAPE_E_INVALID;
  }
  if (t->kb > MAX_KB) {
    set_error("two-level preconditioner: the coarse grid is too large for the banded coarse solve");
    delete t;
    return DRAPE_E_INVALID;
  }
  t->N3p = TILE * t->nt;
  // the factor's order of the real unknowns, and the entries of B's structure on or below its diagonal (a row
  // meets the real unknowns before it in its own segment, and a separator row those of both domains too)
  t->order.assign(t->N3, 0);
  {
    std::vector<int32_t> at(t->N3p, -1);
    for (int I = 0; I < Nc; ++I)
      for (int q = 0; q < dof; ++q) at[pos[I] + q] = dof * I + q;
    int n = 0;
    for (int P = 0; P < t->N3p; ++P)
      if (at[P] >= 0) t->order[n++] = at[P];
    double entries = 0.0;
    for (int g = 0; g < 3; ++g)
      for (int i = 0; i < t->real[g]; ++i) entries += i + 1 + (g == 2 ? t->real[0] + t->real[1] : 0);
    t->tri_entries = entries;
  }
  std::vector<int32_t> ppos(4 * (size_t)c.R, -1);
  for (size_t i = 0; i < ppos.size(); ++i)
    if (t->parent[i] >= 0) ppos[i] = pos[t->parent[i]];
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
  TL_ALLOC(t->d_cvert, cvert.size());
  TL_ALLOC(t->d_adiag, (size_t)t->nt * TILE);
  if (dof == 6) TL_ALLOC(t->d_lever, t->parent.size());
  TL_ALLOC(t->d_child_ptr, child_ptr.size());
  TL_ALLOC(t->d_child, child.size());
  TL_ALLOC(t->d_pair_ptr, pair_ptr.size());
  TL_ALLOC(t->d_pair_key, pair_key.size());
  TL_ALLOC(t->d_pair_terms, terms.size());
  TL_ALLOC(t->d_pos, pos.size());
  TL_ALLOC(t->d_ppos, ppos.size());
  TL_ALLOC(t->d_band, (size_t)t->nt * (t->kb + 1) * TILE * TILE);
  TL_ALLOC(t->d_linv, (size_t)t->nt * TILE * TILE);
  TL_ALLOC(t->d_lpanel, (size_t)t->nt * std::max(t->kb, 1) * TILE * TILE);
  TL_ALLOC(t->d_border, (size_t)2 * (t->kb + 1) * std::max(t->ts, 1) * TILE * TILE);
  TL_ALLOC(t->d_lborder, (size_t)2 * (t->kb + 1) * std::max(t->ts, 1) * TILE * TILE);
  TL_ALLOC(t->d_sacc, (size_t)std::max(t->ts * t->ts, 1) * TILE * TILE);
  TL_ALLOC(t->d_xring, (size_t)3 * (t->kb + 1) * TILE * t->N3p);
  if (t->ts > 0) TL_ALLOC(t->d_ysep, (size_t)t->ts * TILE * t->N3p);
  TL_ALLOC(t->d_blockval, (size_t)9 * c.B);
  if ((e = hipStreamCreateWithFlags(&t->inv_stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  // the chain's events order the two device streams only: a device-scope release is enough, and spares each
  // panel the system-scope cache writeback that stalled the next launch of the chain
  if ((e = hipEventCreateWithFlags(&t->ev_panel, hipEventDisableTiming | hipEventReleaseToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipEventCreateWithFlags(&t->ev_inverse, hipEventDisableTiming | hipEventReleaseToDevice)) != hipSuccess)
    return fail(e);
  std::vector<GramBlock> gram;
  std::vector<int4> gram_cover;
  int gram_part = 0;
  if (gram_partition(t, gram, gram_cover, &gram_part)) {
    TL_ALLOC(t->d_gram, gram.size());
    TL_ALLOC(t->d_gram_part, (size_t)gram_part);
    TL_ALLOC(t->d_gram_cover, gram_cover.size());
    t->gram_blocks = (int)gram.size();
  }
  TL_ALLOC(t->d_B, (size_t)t->N3p * t->N3p);
  if (t->gram_blocks == 0) TL_ALLOC(t->d_BT, (size_t)t->N3p * t->N3p);
  TL_ALLOC(t->d_Bh, (size_t)t->N3p * t->N3p);
  TL_ALLOC(t->d_BTh, (size_t)t->N3p * t->N3p);
  TL_ALLOC(t->d_t, (size_t)t->N3p);
  TL_ALLOC(t->d_rc, (size_t)t->N3p);
  TL_ALLOC(t->d_yc, (size_t)t->N3p);
  TL_ALLOC(t->d_info, 2);
#undef TL_ALLOC
  if ((e = hipHostMalloc((void**)&t->h_info, 2 * sizeof(int))) != hipSuccess) return fail(e);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  if ((e = up(t->d_parent, t->parent.data(), t->parent.size() * 4)) != hipSuccess ||
      (e = up(t->d_pos, pos.data(), pos.size() * 4)) != hipSuccess ||
      (e = up(t->d_ppos, ppos.data(), ppos.size() * 4)) != hipSuccess ||
      (e = up(t->d_w, t->bilinear.data(), t->bilinear.size() * 4)) != hipSuccess ||
      (e = up(t->d_wbil, t->bilinear.data(), t->bilinear.size() * 4)) != hipSuccess ||
      (e = up(t->d_fxy, fxy.data(), fxy.size() * sizeof(float2))) != hipSuccess ||
      (e = up(t->d_cvert, cvert.data(), cvert.size() * 4)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_info, 0, 2 * sizeof(int), s)) != hipSuccess ||
      (e = up(t->d_child_ptr, child_ptr.data(), child_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_child, child.data(), child.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_ptr, pair_ptr.data(), pair_ptr.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_key, pair_key.data(), pair_key.size() * 4)) != hipSuccess ||
      (e = up(t->d_pair_terms, terms.data(), terms.size() * sizeof(uint4))) != hipSuccess ||
      (e = up(t->d_gram, gram.data(), t->gram_blocks * sizeof(GramBlock))) != hipSuccess ||
      (e = up(t->d_gram_cover, gram_cover.data(), t->gram_blocks ? gram_cover.size() * sizeof(int4) : 0)) !=
          hipSuccess ||
      (e = hipMemsetAsync(t->d_rc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_yc, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_t, 0, (size_t)t->N3p * 4, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_B, 0, (size_t)t->N3p * t->N3p * 4, s)) != hipSuccess ||
      (e = t->d_BT ? hipMemsetAsync(t->d_BT, 0, (size_t)t->N3p * t->N3p * 4, s) : hipSuccess) != hipSuccess ||
      (e = hipMemsetAsync(t->d_Bh, 0, (size_t)t->N3p * t->N3p * 2, s)) != hipSuccess ||
      (e = hipMemsetAsync(t->d_BTh, 0, (size_t)t->N3p * t->N3p * 2, s)) != hipSuccess ||
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
  if (count != gnn_param_count(false) && count != gnn_param_count(true)) {
    set_error("prolongation model: expected " + std::to_string(gnn_param_count(false)) + " parameters (" +
              std::to_string(gnn_param_count(true)) + " with the rotation head)");
    return DRAPE_E_INVALID;
  }
  if (t->d_theta && t->theta_count != count) {
    HIP_TRY(hipFree(t->d_theta));
    t->d_theta = nullptr;
  }
  if (!t->d_theta) HIP_TRY(hipMalloc((void**)&t->d_theta, (size_t)count * sizeof(float)));
  t->theta_count = count;
  t->rotation_head = count == gnn_param_count(true) && t->dof == 6;
  if (t->rotation_head && !t->d_sigma) HIP_TRY(hipMalloc((void**)&t->d_sigma, 4 * (size_t)c.R * sizeof(float)));
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
  a.sigma = t->rotation_head ? t->d_sigma : nullptr;
  return a;
}

static int run_model(Context& c) { return gnn_infer(two_level_gnn_args(c)); }

static Tiles tiles_of(const TwoLevel* t) {
  Tiles v;
  v.band = t->d_band;
  v.lpanel = t->d_lpanel;
  v.border = t->d_border;
  v.lborder = t->d_lborder;
  v.sacc = t->d_sacc;
  v.ysep = t->d_ysep;
  v.kb = t->kb;
  v.nt = t->nt;
  v.ts = t->ts;
  for (int i = 0; i < 4; ++i) v.seg[i] = t->seg[i];
  for (int i = 0; i < 2; ++i) v.bstart[i] = t->bstart[i];
  for (int i = 0; i < 3; ++i) v.real[i] = t->real[i];
  v.adiag = t->d_adiag;
  v.floor = 0.0;
  return v;
}

static Split split_of(const TwoLevel* t) {
  return {TILE * t->seg[1], TILE * t->seg[2], t->real[0], t->real[1], t->real[2]};
}

// The FP16 copies are formed and applied under rules.coarse_half = 1, and under 2 (auto) once the FP32 pair
// B, B^T outgrows half of the 256 MB Infinity Cache.
static bool use_half(const Context& c, const TwoLevel* t) {
  const bool big = 8.0 * t->tri_entries > (double)((size_t)128 << 20);
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
    const size_t nb = (size_t)t->N3p * t->N3p;
    HIP_TRY(hipMalloc((void**)&v.linv, (size_t)t->nt * TILE * TILE * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&v.B, nb * sizeof(float)));
    if (t->d_BT) HIP_TRY(hipMalloc((void**)&v.BT, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&v.Bh, nb * sizeof(_Float16)));
    HIP_TRY(hipMalloc((void**)&v.BTh, nb * sizeof(_Float16)));
    // the copies are written below/above the diagonal only: the rest stays zero
    HIP_TRY(hipMemsetAsync(v.B, 0, nb * sizeof(float), c.stream));
    if (v.BT) HIP_TRY(hipMemsetAsync(v.BT, 0, nb * sizeof(float), c.stream));
    HIP_TRY(hipMemsetAsync(v.Bh, 0, nb * sizeof(_Float16), c.stream));
    HIP_TRY(hipMemsetAsync(v.BTh, 0, nb * sizeof(_Float16), c.stream));
    // the prolongation travels with its factorisation; a fresh copy starts as the current one
    HIP_TRY(hipMalloc((void**)&v.w, 4 * (size_t)c.R * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(v.w, t->d_w, 4 * (size_t)c.R * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
    if (t->dof == 6) {
      HIP_TRY(hipMalloc((void**)&v.lever, 4 * (size_t)c.R * sizeof(float4)));
      HIP_TRY(hipMemcpyAsync(v.lever, t->d_lever, 4 * (size_t)c.R * sizeof(float4), hipMemcpyDeviceToDevice,
                             c.stream));
    }
  }
  std::swap(v.w, t->d_w);
  std::swap(v.lever, t->d_lever);
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
  std::swap(q.lever, t->d_lever);
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
  t->replaced += q.h_info[1];
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

static int factorise_chain(Context& c, hipStream_t s, const float* w, const float4* lever, double* linv, float* B,
                           float* BT, _Float16* Bh, _Float16* BTh, bool half, hipEvent_t after_product);
static int compute_levers(Context& c, hipStream_t s, const float* w, float4* lever);

// The factorisation of the assembled system on the side stream, into the pending buffers; the current one
// keeps serving until the next system adopts it.
static int factorise_async(Context& c) {
  TwoLevel* t = c.two_level;
  TwoLevel::Pending& q = t->pending;
  if (!q.side) {
    const size_t nb = (size_t)t->N3p * t->N3p;
    HIP_TRY(hipStreamCreateWithFlags(&q.side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&q.ready, &q.read, &q.done}) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    for (hipEvent_t* e : {&q.began, &q.ended}) HIP_TRY(hipEventCreate(e));
    HIP_TRY(hipMalloc((void**)&q.w, 4 * (size_t)c.R * sizeof(float)));
    if (t->dof == 6) HIP_TRY(hipMalloc((void**)&q.lever, 4 * (size_t)c.R * sizeof(float4)));
    HIP_TRY(hipMalloc((void**)&q.linv, (size_t)t->nt * TILE * TILE * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&q.B, nb * sizeof(float)));
    if (t->d_BT) HIP_TRY(hipMalloc((void**)&q.BT, nb * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&q.Bh, nb * sizeof(_Float16)));
    HIP_TRY(hipMalloc((void**)&q.BTh, nb * sizeof(_Float16)));
    HIP_TRY(hipHostMalloc((void**)&q.h_info, 2 * sizeof(int)));
    // the copies are written below/above the diagonal only: the rest stays zero
    HIP_TRY(hipMemsetAsync(q.B, 0, nb * sizeof(float), c.stream));
    if (q.BT) HIP_TRY(hipMemsetAsync(q.BT, 0, nb * sizeof(float), c.stream));
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
  TRY_RC(compute_levers(c, q.side, q.w, q.lever));
  TRY_RC(factorise_chain(c, q.side, q.w, q.lever, q.linv, q.B, q.BT, q.Bh, q.BTh, q.half_valid, q.read));
  HIP_TRY(hipEventRecord(q.ended, q.side));
  HIP_TRY(hipMemcpyAsync(q.h_info, t->d_info, 2 * sizeof(int), hipMemcpyDeviceToHost, q.side));
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
  TRY_RC(compute_levers(c, s, t->d_w, t->d_lever));
  TRY_RC(factorise_chain(c, s, t->d_w, t->d_lever, t->d_linv, t->d_B, t->d_BT, t->d_Bh, t->d_BTh, t->half_valid,
                         nullptr));
  TRY_RC(read_back(c, t->h_info, t->d_info, 2 * sizeof(int)));
  t->factorisations++;
  t->replaced += t->h_info[1];
  t->ok = *t->h_info == 0;
  if (!t->ok) t->failures++;
  *ok = t->ok;
  return DRAPE_OK;
}

// The rotation levers of the slots for the weights w (and the network's rotation scales when it has a
// rotation head) at the current positions, on stream s; nothing without rotation unknowns.
static int compute_levers(Context& c, hipStream_t s, const float* w, float4* lever) {
  TwoLevel* t = c.two_level;
  if (t->dof != 6) return DRAPE_OK;
  const int slots = 4 * c.R;
  k_lever<<<blocks_for(slots), 256, 0, s>>>(t->d_parent, w, t->d_theta && t->rotation_head ? t->d_sigma : nullptr,
                                            c.rvert, t->d_cvert,
                                            c.pos, t->inv_spacing, slots, lever);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

// Galerkin product with weights w (and rotation levers), banded Cholesky and the inverse factor into (linv, B,
// B^T and, when `half`, the FP16 copies), on stream s; `after_product` (if not null) is recorded once the
// product has read the assembled matrix.
static int factorise_chain(Context& c, hipStream_t s, const float* w, const float4* lever, double* linv, float* B,
                           float* BT, _Float16* Bh, _Float16* BTh, bool half, hipEvent_t after_product) {
  TwoLevel* t = c.two_level;
  Tiles tl = tiles_of(t);
  tl.floor = c.rules.coarse_pivot_floor;
  const size_t tile_bytes = (size_t)TILE * TILE * sizeof(double);
  HIP_TRY(hipMemsetAsync(t->d_band, 0, (size_t)t->nt * (t->kb + 1) * tile_bytes, s));
  if (t->ts > 0) {
    HIP_TRY(hipMemsetAsync(t->d_border, 0, (size_t)2 * (t->kb + 1) * t->ts * tile_bytes, s));
    HIP_TRY(hipMemsetAsync(t->d_sacc, 0, (size_t)t->ts * t->ts * tile_bytes, s));
  }
  HIP_TRY(hipMemsetAsync(t->d_info, 0, 2 * sizeof(int), s));
  k_block_values<<<blocks_for(c.B), 256, 0, s>>>(c.sell_val, c.block_dst, c.B, t->d_blockval);
  if (t->dof == 6)
    k_galerkin_band<6><<<blocks_for(t->npairs, 4), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, w,
                                                                 lever, t->d_blockval, t->npairs, t->Nc, t->d_pos, tl);
  else
    k_galerkin_band<3><<<blocks_for(t->npairs, 4), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, w,
                                                                 nullptr, t->d_blockval, t->npairs, t->Nc, t->d_pos, tl);
  if (after_product) HIP_TRY(hipEventRecord(after_product, s));
  k_band_diagonal<<<blocks_for(t->nt * TILE), 256, 0, s>>>(tl, 1.0 + c.rules.coarse_shift);
  const int slots = t->kb + t->ts;
  const int per = slots + slots * (slots + 1) / 2;  // worker workgroups per finished column
  auto inverse_columns = [&](int I) { return (TILE * (I + 1) - (I >= t->seg[1] && I < t->seg[2] ? TILE * t->seg[1] : 0)) / INV_COLS; };
  // Tile rows of B = L^-1 go to the inverse stream once their panel has run (its L^-1 tile formed), a few steps
  // at a time behind one event, beside the chain's later steps.
  std::vector<int2> rows;
  int since = 0;
  auto flush_rows = [&]() -> int {
    if (rows.empty()) return DRAPE_OK;
    HIP_TRY(hipEventRecord(t->ev_panel, s));
    HIP_TRY(hipStreamWaitEvent(t->inv_stream, t->ev_panel, 0));
    for (const int2 r : rows) {
      const int w1 = inverse_columns(r.x), w2 = r.y >= 0 ? inverse_columns(r.y) : 0;
      k_inverse_row<<<w1 + w2, 256, 0, t->inv_stream>>>(tl, linv, r, w1, t->N3p, t->d_xring, B, BT,
                                                       half ? Bh : nullptr, half ? BTh : nullptr, t->d_info);
    }
    rows.clear();
    since = 0;
    return DRAPE_OK;
  };
  // step k of a group of segments: the panel of each segment's column k, the workers of its column k - 1
  auto steps = [&](int first, int last) -> int {
    int longest = 0;
    for (int g = first; g <= last; ++g) longest = std::max(longest, t->seg[g + 1] - t->seg[g]);
    for (int k = 0; k <= longest; ++k) {
      Step st{};
      for (int g = first; g <= last; ++g) {
        const int n = t->seg[g + 1] - t->seg[g];
        if (k < n) st.c[st.panels++] = t->seg[g] + k;
        if (k >= 1 && k <= n) st.p[st.workers++] = t->seg[g] + k - 1;
      }
      if (st.panels + st.workers == 0) continue;
      k_chol_step<<<st.panels + st.workers * per, STEP_THREADS, 0, s>>>(tl, linv, st, t->d_info);
      if (st.panels > 0) rows.push_back(make_int2(st.c[0], st.panels > 1 ? st.c[1] : -1));
      if (++since == 4) TRY_RC(flush_rows());
    }
    return DRAPE_OK;
  };
  // the two domains side by side, then the separator
  TRY_RC(steps(0, 1));
  if (t->ts > 0) {
    // every domain row of L^-1 is queued, and the event after the domains' last step (whose workers form the
    // last border factor tiles) orders the separator rows' border sums behind them on the inverse stream
    TRY_RC(flush_rows());
    HIP_TRY(hipEventRecord(t->ev_panel, s));
    HIP_TRY(hipStreamWaitEvent(t->inv_stream, t->ev_panel, 0));
    k_separator_border<<<t->ts * t->seg[2], 256, 0, t->inv_stream>>>(tl, t->N3p, t->d_xring, t->d_info);
    k_add_separator<<<16 * t->ts * t->ts, 256, 0, s>>>(tl);
    TRY_RC(steps(2, 2));
  }
  TRY_RC(flush_rows());
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
  if (t->dof == 6) {
    // the levers of the current weights at the current positions (the factorisation's are left as they are)
    if (!t->d_Ac_lever) HIP_TRY(hipMalloc((void**)&t->d_Ac_lever, 4 * (size_t)c.R * sizeof(float4)));
    TRY_RC(compute_levers(c, s, t->d_w, t->d_Ac_lever));
    k_galerkin_dense<6><<<blocks_for(t->npairs), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, t->d_w,
                                                              t->d_Ac_lever, c.sell_val, c.block_dst, t->npairs, t->Nc,
                                                              t->d_Ac);
  } else {
    k_galerkin_dense<3><<<blocks_for(t->npairs), 256, 0, s>>>(t->d_pair_ptr, t->d_pair_key, t->d_pair_terms, t->d_w,
                                                              nullptr, c.sell_val, c.block_dst, t->npairs, t->Nc,
                                                              t->d_Ac);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, t->d_Ac, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

// B = L^-1 of the current factorisation (FP32, row-major [3Nc x 3Nc]) between the real unknowns, in the
// factor's ordering (two_level_ordering), into host memory.
int two_level_inverse_factor(Context& c, float* out) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  std::vector<float> full((size_t)t->N3p * t->N3p);
  HIP_TRY(hipMemcpyAsync(full.data(), t->d_B, full.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  std::vector<int> real_pos;
  for (int g = 0; g < 3; ++g)
    for (int i = 0; i < t->real[g]; ++i) real_pos.push_back(TILE * t->seg[g] + i);
  const size_t n = real_pos.size();
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) out[i * n + j] = full[(size_t)real_pos[i] * t->N3p + real_pos[j]];
  return DRAPE_OK;
}

// Unknowns per coarse node, the rotation levers of the current factorisation [4R] (x, y, z, 0 per slot; zero
// without rotations) and the pivots replaced by the pivot floor over all factorisations. Any output may be null.
int two_level_rotations(Context& c, int32_t* dof, float* levers, int64_t* replaced) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  if (dof) *dof = t->dof;
  if (replaced) *replaced = t->replaced;
  if (levers) {
    const size_t n = 4 * (size_t)c.R;
    if (t->dof == 6) {
      HIP_TRY(hipMemcpyAsync(levers, t->d_lever, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
      HIP_TRY(hipStreamSynchronize(c.stream));
    } else {
      std::fill(levers, levers + 4 * n, 0.0f);
    }
  }
  return DRAPE_OK;
}

// The factor's ordering: order[i] = dof I + q, the natural coarse unknown at its i-th real position.
int two_level_ordering(Context& c, int32_t* order) {
  const TwoLevel* t = c.two_level;
  std::copy(t->order.begin(), t->order.end(), order);
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
static void launch_gram(const TwoLevel* t, hipStream_t s, const PcgState* st, int guard) {
  k_gram<<<t->gram_blocks, GRAM_THREADS, 0, s>>>(t->d_gram, t->d_B, t->d_rc, t->N3p, t->d_gram_part, st, guard);
  const int n4 = t->N3p / 4, per = 256 / GRAM_SUM_LANES;
  k_gram_sum<<<(n4 + per - 1) / per, 256, 0, s>>>(t->d_gram, t->d_gram_cover, t->gram_blocks, t->d_gram_part, n4,
                                                  t->d_yc, st, guard);
}

static void coarse_solve(Context& c, const TwoLevel* t, int guard, bool fp32) {
  hipStream_t s = c.stream;
  const int blocks = t->N3p / 4;
  const Split sp = split_of(t);
  if (t->half_valid && use_half(c, t) && !fp32) {
    k_tri_gemv_h<true><<<blocks, 256, 0, s>>>(t->d_Bh, t->d_rc, t->d_linv, sp, t->N3p, t->d_t, c.state, guard);
    k_tri_gemv_h<false><<<blocks, 256, 0, s>>>(t->d_BTh, t->d_t, t->d_linv, sp, t->N3p, t->d_yc, c.state, guard);
  } else if (t->gram_blocks > 0) {
    launch_gram(t, s, c.state, guard);
  } else {
    k_tri_gemv<true><<<blocks, 256, 0, s>>>(t->d_B, t->d_rc, sp, t->N3p, t->d_t, c.state, guard);
    k_tri_gemv<false><<<blocks, 256, 0, s>>>(t->d_BT, t->d_t, sp, t->N3p, t->d_yc, c.state, guard);
  }
}

// One of the coarse solve's triangular GEMVs (which 0: t = B r_c, 1: y_c = B^T t) of the current factorisation,
// or (which 2) the fused y_c = B^T B r_c (k_gram and k_gram_sum), `launches` times back to back between two
// events on the context's stream: average ms per launch and the algorithmic bytes of one launch (the triangle
// in its storage precision, x read and y written once; the fused solve reads the triangle once as well).
int two_level_gemv_timing(Context& c, int which, int launches, double* avg_ms, double* bytes) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  const int blocks = t->N3p / 4;
  const Split sp = split_of(t);
  const bool half = t->half_valid && use_half(c, t);
  const bool lower = which == 0;
  if (which == 2 && (half || t->gram_blocks == 0)) {
    set_error("kernel timing: the coarse solve of this layout runs as two GEMVs");
    return DRAPE_E_INVALID;
  }
  if (which == 1 && !half && !t->d_BT) {
    set_error("kernel timing: B^T is not stored (the coarse solve of this layout is fused)");
    return DRAPE_E_INVALID;
  }
  *bytes = t->tri_entries * (half ? 2.0 : 4.0) + 8.0 * t->N3;
  auto launch = [&]() {
    if (which == 2) {
      launch_gram(t, s, c.state, 2);
    } else if (half) {
      if (lower) k_tri_gemv_h<true><<<blocks, 256, 0, s>>>(t->d_Bh, t->d_rc, t->d_linv, sp, t->N3p, t->d_t, c.state, 2);
      else k_tri_gemv_h<false><<<blocks, 256, 0, s>>>(t->d_BTh, t->d_t, t->d_linv, sp, t->N3p, t->d_yc, c.state, 2);
    } else {
      if (lower) k_tri_gemv<true><<<blocks, 256, 0, s>>>(t->d_B, t->d_rc, sp, t->N3p, t->d_t, c.state, 2);
      else k_tri_gemv<false><<<blocks, 256, 0, s>>>(t->d_BT, t->d_t, sp, t->N3p, t->d_yc, c.state, 2);
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

// FP64 flops of the tile algorithm of one factorisation (the Galerkin product's terms not counted), with T = 64^3:
// per tile column with m tiles below the diagonal, the diagonal factor (2/3) T, its inverse (1/3) T, m TRSMs of T
// each and the trailing updates (m tiles of T on the diagonal, m (m - 1) / 2 of 2 T below it); per tile (I, J) of
// L^-1 below the diagonal, 2 T for each nonzero L(I, K), J <= K < I, and 2 T for the product with L_II^-1.
static double factor_flops(const TwoLevel* t) {
  const double T = 64.0 * 64.0 * 64.0;
  double f = 0.0;
  for (int s = 0; s < 3; ++s)
    for (int k = t->seg[s]; k < t->seg[s + 1]; ++k) {
      int m = std::min(t->kb, t->seg[s + 1] - 1 - k);
      if (s < 2 && t->ts > 0 && k >= t->bstart[s]) m += t->ts;
      f += T * (2.0 / 3.0 + 1.0 / 3.0 + 2.0 * m + (double)m * (m - 1));
    }
  for (int s = 0; s < 3; ++s)
    for (int I = t->seg[s]; I < t->seg[s + 1]; ++I)
      for (int J = s == 1 ? t->seg[1] : 0; J < I; ++J) {
        const int sJ = J < t->seg[1] ? 0 : (J < t->seg[2] ? 1 : 2);
        int terms = std::max(0, std::min(t->kb, I - std::max(t->seg[s], J)));
        if (s == 2 && sJ < 2) terms += std::max(0, t->seg[sJ + 1] - std::max(t->bstart[sJ], J));
        f += 2.0 * T * (terms + 1);
      }
  return f;
}

// Device time of one factorisation of the current system with the current weights and levers (the Galerkin
// product, the chain and the inverse sweep; HIP events on the context's stream, which waits for the inverse stream
// at the end of each), averaged over `launches` back-to-back ones, and its FP64 flops (factor_flops). The factor is
// formed again from the same inputs, so the preconditioner afterwards is that of the current system; the
// factorisation count is not changed.
int two_level_factor_timing(Context& c, int launches, double* avg_ms, double* flops) {
  TRY_RC(adopt_pending(c));
  TwoLevel* t = c.two_level;
  if (t->factored == ~0ull) {
    set_error("kernel timing: no coarse factorisation yet");
    return DRAPE_E_INVALID;
  }
  hipStream_t s = c.stream;
  auto once = [&]() -> int {
    return factorise_chain(c, s, t->d_w, t->d_lever, t->d_linv, t->d_B, t->d_BT, t->d_Bh, t->d_BTh, t->half_valid,
                           nullptr);
  };
  TRY_RC(once());
  HIP_TRY(hipEventRecord(c.ev_spmv0, s));
  for (int i = 0; i < launches; ++i) TRY_RC(once());
  HIP_TRY(hipEventRecord(c.ev_spmv1, s));
  HIP_TRY(hipEventSynchronize(c.ev_spmv1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c.ev_spmv0, c.ev_spmv1));
  *avg_ms = ms / launches;
  *flops = factor_flops(t);
  TRY_RC(read_back(c, t->h_info, t->d_info, 2 * sizeof(int)));
  t->factored = c.system_version;
  t->factored_shift = c.shift_version;
  t->factored_h = c.shift_h;
  t->ok = *t->h_info == 0;
  t->solves = 0;
  return DRAPE_OK;
}

// z = D^-1 r + P B^T B P^T r for the current r (guard: 0 running, 1 replacement gate, 2 setup with p = z).
static void launch_restrict(Context& c, const TwoLevel* t, const float4* r, int guard) {
  hipStream_t s = c.stream;
  if (t->dof == 6)
    k_restrict<true><<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, t->d_lever, t->d_pos, r, t->Nc,
                                                     t->d_rc, c.state, guard);
  else
    k_restrict<false><<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, nullptr, t->d_pos, r, t->Nc,
                                                      t->d_rc, c.state, guard);
}

static const float4* lever_of(const TwoLevel* t) { return t->dof == 6 ? t->d_lever : nullptr; }

int two_level_apply(Context& c, float* partials, int guard) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  launch_restrict(c, t, c.r, guard);
  coarse_solve(c, t, guard, false);
  k_prolong<<<c.G, 256, 0, s>>>(t->d_ppos, t->d_w, lever_of(t), t->d_yc, c.dinv, c.r, c.z, c.p, partials, c.R, c.state,
                                guard);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int two_level_iterate(Context& c, const float* partial_pq, const float4* p, float* partial_rz, int it) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  if (t->dof == 6)
    k_restrict_update<true><<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, t->d_lever, t->d_pos,
                                                             c.r, c.q, t->Nc, t->d_rc, partial_pq, c.G, c.state, it);
  else
    k_restrict_update<false><<<(t->Nc + 3) / 4, 256, 0, s>>>(t->d_child_ptr, t->d_child, t->d_w, nullptr, t->d_pos,
                                                              c.r, c.q, t->Nc, t->d_rc, partial_pq, c.G, c.state, it);
  coarse_solve(c, t, 0, false);
  k_prolong_update<<<c.G, 256, 0, s>>>(t->d_ppos, t->d_w, lever_of(t), t->d_yc, c.dinv, p, c.q, c.x, c.r, c.z,
                                       partial_rz, c.R, partial_pq, c.G, c.state, it);
  HIP_TRY(hipGetLastError());
  return DRAPE_OK;
}

int two_level_dof(const Context& c) { return c.two_level ? c.two_level->dof : 3; }

CoarseView two_level_view(const Context& c) {
  const TwoLevel* t = c.two_level;
  return {t->d_ppos, t->d_w, lever_of(t), t->d_yc};
}

int two_level_restrict_solve(Context& c, const float4* r, int guard, bool fp32) {
  TwoLevel* t = c.two_level;
  hipStream_t s = c.stream;
  launch_restrict(c, t, r, guard);
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
