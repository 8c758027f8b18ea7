// Device builder of a sheet problem's topology and of the assembly scatter maps: interior hinges
// (mesh.rs:83-107), the BSR pattern of the coupling groups (bsr.rs:12-29, drape.rs:126-134), element
// geometry and hinge stiffness (cases.rs:94-105), lumped masses (mesh.rs:27-36) and the per-block and
// per-row contribution lists (e3_generate.rs:111-161), built with stable radix sorts instead of the
// reference's hash map and per-row vectors.
//
// Every order the reference's builder fixes is reproduced: a stable sort of (key, value) pairs generated
// in the reference's loop order keeps that order inside each key, so the hinge opposites, the block
// contribution lists (mass, membranes, hinges) and the force lists come out entry for entry as the
// host builder and problem.bin have them. Geometry is f64 with the host builder's operation order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "../internal.h"

namespace drape {

namespace {

constexpr uint64_t NO_KEY = ~0ull;
constexpr uint32_t NO_ROW = 0xffffffffu;

// Device temporaries released at scope exit.
struct Pool {
  std::vector<void*> ptrs;
  ~Pool() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T** out, size_t n) {
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(p);
      *out = static_cast<T*>(p);
    }
    return e;
  }
};

template <class K, class V>
int sort_pairs(Pool& pool, const K* kin, K* kout, const V* vin, V* vout, int n, hipStream_t s) {
  if (n == 0) return DRAPE_OK;
  size_t bytes = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, (int)(8 * sizeof(K)), s));
  char* tmp = nullptr;
  HIP_TRY(pool.get(&tmp, bytes));
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, (int)(8 * sizeof(K)), s));
  return DRAPE_OK;
}

int exclusive_sum(Pool& pool, const uint32_t* in, uint32_t* out, int n, hipStream_t s) {
  if (n == 0) return DRAPE_OK;
  size_t bytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
  char* tmp = nullptr;
  HIP_TRY(pool.get(&tmp, bytes));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s));
  return DRAPE_OK;
}

// the count of flagged items: exclusive sum of the last item plus its flag
int flagged_total(const uint32_t* flags, const uint32_t* excl, int n, hipStream_t s, uint32_t* total) {
  if (n == 0) {
    *total = 0;
    return DRAPE_OK;
  }
  uint32_t a = 0, b = 0;
  HIP_TRY(hipMemcpyAsync(&a, excl + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&b, flags + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *total = a + b;
  return DRAPE_OK;
}

template <class K>
__device__ inline uint32_t lower_bound(const K* __restrict__ keys, uint32_t n, K key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// n items below the first key that is not smaller than `key` (one thread)
template <class K>
__global__ void k_count_below(const K* __restrict__ keys, uint32_t n, K key, uint32_t* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = lower_bound(keys, n, key);
}

// ptr[i] = first sorted position with key >= i, for i in [0, count]
__global__ void __launch_bounds__(256) k_ptr_from_sorted(const uint32_t* __restrict__ keys, uint32_t n, int count,
                                                         uint32_t* __restrict__ ptr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > count) return;
  ptr[i] = lower_bound(keys, n, (uint32_t)i);
}

__device__ inline double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

// ---- rows -----------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_free_flags(const uint8_t* __restrict__ free_mask, int V,
                                                    uint32_t* __restrict__ flags) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < V) flags[v] = free_mask[v] ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_rows(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ excl,
                                              int V, int32_t* __restrict__ row_of, uint32_t* __restrict__ row_vertex) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  if (flags[v]) {
    row_of[v] = (int32_t)excl[v];
    row_vertex[excl[v]] = (uint32_t)v;
  } else {
    row_of[v] = -1;
  }
}

// ---- hinges ---------------------------------------------------------------------------------------
// edge (min, max) of side k of triangle e, valued 3e + k (mesh.rs:86-91 visiting order)
__global__ void __launch_bounds__(256) k_edge_keys(const uint32_t* __restrict__ tris, int T,
                                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 3 * T) return;
  const int e = q / 3, k = q % 3;
  const uint32_t a = tris[3 * e + k], b = tris[3 * e + (k + 1) % 3];
  keys[q] = ((uint64_t)min(a, b) << 32) | (uint64_t)max(a, b);
  vals[q] = (uint32_t)q;
}

// the first of exactly two sorted entries of an edge: an interior edge with its two opposites
__global__ void __launch_bounds__(256) k_hinge_flags(const uint64_t* __restrict__ keys, int n,
                                                     uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const bool start = i == 0 || keys[i - 1] != k;
  const bool two = i + 1 < n && keys[i + 1] == k && (i + 2 >= n || keys[i + 2] != k);
  flags[i] = start && two ? 1u : 0u;
}

// oriented hinge [left opposite, right opposite, edge start, edge end] (mesh.rs:99-106), with its low sort key
__global__ void __launch_bounds__(256) k_hinge_emit(const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ pos, int n,
                                                    const uint32_t* __restrict__ tris,
                                                    const double* __restrict__ rest, uint32_t* __restrict__ hv,
                                                    uint64_t* __restrict__ key_lo, uint32_t* __restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const uint32_t a = (uint32_t)(keys[i] >> 32), b = (uint32_t)(keys[i] & 0xffffffffu);
  const uint32_t v0 = vals[i], v1 = vals[i + 1];
  const uint32_t first = tris[3 * (v0 / 3) + (v0 % 3 + 2) % 3];
  const uint32_t second = tris[3 * (v1 / 3) + (v1 % 3 + 2) % 3];
  const double ax = rest[2 * a], ay = rest[2 * a + 1];
  const bool left = cross2(rest[2 * b] - ax, rest[2 * b + 1] - ay, rest[2 * first] - ax, rest[2 * first + 1] - ay) > 0.0;
  const uint32_t h = pos[i];
  const uint32_t o0 = left ? first : second, o1 = left ? second : first;
  hv[4 * h] = o0;
  hv[4 * h + 1] = o1;
  hv[4 * h + 2] = a;
  hv[4 * h + 3] = b;
  key_lo[h] = ((uint64_t)a << 32) | b;
  idx[h] = h;
}

__global__ void __launch_bounds__(256) k_hinge_key_hi(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ hv,
                                                      int H, uint64_t* __restrict__ key_hi) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= H) return;
  const uint32_t h = idx[i];
  key_hi[i] = ((uint64_t)hv[4 * h] << 32) | hv[4 * h + 1];
}

// sorted hinges and their stiffness k_b(e) 3 |e|^2 / (A1 + A2) (cases.rs:94-105)
__global__ void __launch_bounds__(256) k_hinge_gather(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ hv,
                                                      int H, const double* __restrict__ rest, double bend_wale,
                                                      double bend_course, uint32_t* __restrict__ out,
                                                      float* __restrict__ stiffness) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= H) return;
  const uint32_t h = idx[i];
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = hv[4 * h + k];
    out[4 * i + k] = v[k];
  }
  const double p1x = rest[2 * v[0]], p1y = rest[2 * v[0] + 1], p2x = rest[2 * v[1]], p2y = rest[2 * v[1] + 1];
  const double p3x = rest[2 * v[2]], p3y = rest[2 * v[2] + 1], p4x = rest[2 * v[3]], p4y = rest[2 * v[3] + 1];
  const double length = hypot(p4x - p3x, p4y - p3y);
  const double dx = (p4x - p3x) / length, dy = (p4y - p3y) / length;
  const double bending = bend_wale * dy * dy + bend_course * dx * dx;
  const double e0 = p4x - p3x, e1 = p4y - p3y;
  const double length2 = e0 * e0 + e1 * e1;
  const double a1 = 0.5 * fabs((p3x - p1x) * (p4y - p1y) - (p3y - p1y) * (p4x - p1x));
  const double a2 = 0.5 * fabs((p3x - p2x) * (p4y - p2y) - (p3y - p2y) * (p4x - p2x));
  stiffness[i] = (float)(bending * 3.0 * length2 / (a1 + a2));
}

// ---- elements and masses --------------------------------------------------------------------------
// shape gradients and area (e3_generate.rs:108-111), the f64 area, and the (vertex, triangle) pairs
__global__ void __launch_bounds__(256) k_elements(const uint32_t* __restrict__ tris, int T,
                                                  const double* __restrict__ rest, float* __restrict__ geom,
                                                  double* __restrict__ area, uint32_t* __restrict__ vkeys,
                                                  uint32_t* __restrict__ vvals) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= T) return;
  const uint32_t t0 = tris[3 * e], t1 = tris[3 * e + 1], t2 = tris[3 * e + 2];
  const double p0x = rest[2 * t0], p0y = rest[2 * t0 + 1];
  const double m00 = rest[2 * t1] - p0x, m01 = rest[2 * t2] - p0x;
  const double m10 = rest[2 * t1 + 1] - p0y, m11 = rest[2 * t2 + 1] - p0y;
  const double det = m00 * m11 - m01 * m10;
  const double g1x = m11 / det, g1y = -m01 / det, g2x = -m10 / det, g2y = m00 / det;
  const double g0x = -g1x - g2x, g0y = -g1y - g2y;
  const double g[7] = {g0x, g0y, g1x, g1y, g2x, g2y, 0.5 * fabs(det)};
#pragma unroll
  for (int k = 0; k < 7; ++k) geom[7 * e + k] = (float)g[k];
  area[e] = 0.5 * fabs(cross2(m00, m10, m01, m11));
  vkeys[3 * e] = t0;
  vkeys[3 * e + 1] = t1;
  vkeys[3 * e + 2] = t2;
  vvals[3 * e] = vvals[3 * e + 1] = vvals[3 * e + 2] = (uint32_t)e;
}

// lumped mass: density times the sum of area / 3 over the vertex's triangles in triangle order
__global__ void __launch_bounds__(256) k_vertex_mass(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                     uint32_t n, int V, const double* __restrict__ area,
                                                     double density, double* __restrict__ mass) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= V) return;
  uint32_t i = lower_bound(keys, n, (uint32_t)v);
  double acc = 0.0;
  for (; i < n && keys[i] == (uint32_t)v; ++i) acc += area[vals[i]] / 3.0;
  mass[v] = density * acc;
}

// ---- pattern and block contributions -----------------------------------------------------------------
// (row, column) pairs valued by local 3x3 block (units of 9 floats), in the reference's order: the row
// terms on the diagonal, membranes (9e + 3k + l), hinges (9T + 16h + 4k + l) (e3_generate.rs:118-130)
__global__ void __launch_bounds__(256) k_block_pairs(const uint32_t* __restrict__ tris, const uint32_t* __restrict__ hv,
                                                     const int32_t* __restrict__ row_of, int T, int H, int R,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t nd = (size_t)R, nt = 9 * (size_t)T, nh = 16 * (size_t)H;
  if (q >= nd + nt + nh) return;
  int r, c;
  uint32_t val;
  if (q < nd) {
    r = c = (int)q;
    val = (uint32_t)(9 * (size_t)T + 16 * (size_t)H + q);
  } else if (q < nd + nt) {
    const size_t p = q - nd;
    const size_t e = p / 9;
    const int k = (int)(p % 9) / 3, l = (int)(p % 3);
    r = row_of[tris[3 * e + k]];
    c = row_of[tris[3 * e + l]];
    val = (uint32_t)p;
  } else {
    const size_t p = q - nd - nt;
    const size_t h = p / 16;
    const int k = (int)(p % 16) / 4, l = (int)(p % 4);
    r = row_of[hv[4 * h + k]];
    c = row_of[hv[4 * h + l]];
    val = (uint32_t)(9 * (size_t)T + p);
  }
  keys[q] = (r >= 0 && c >= 0) ? (((uint64_t)(uint32_t)r << 32) | (uint32_t)c) : NO_KEY;
  vals[q] = val;
}

__global__ void __launch_bounds__(256) k_block_flags(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || keys[i - 1] != keys[i]) ? 1u : 0u;
}

// the pattern from the sorted unique pairs: column, first contribution, row start and diagonal position
__global__ void __launch_bounds__(256) k_block_emit(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                                                   const uint32_t* __restrict__ pos, int n,
                                                   uint32_t* __restrict__ col_idx, uint32_t* __restrict__ bcp,
                                                   uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ diag) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const uint32_t b = pos[i];
  const uint32_t r = (uint32_t)(keys[i] >> 32), c = (uint32_t)(keys[i] & 0xffffffffu);
  col_idx[b] = c;
  bcp[b] = (uint32_t)i;
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != r) row_ptr[r] = b;
  if (c == r) diag[r] = b;
}

// per-row force pairs valued by local gradient (units of 3 floats): the row term (3T + 4H + row),
// membranes (3e + k), hinges (3T + 4h + k) (e3_generate.rs:142-152)
__global__ void __launch_bounds__(256) k_force_pairs(const uint32_t* __restrict__ tris, const uint32_t* __restrict__ hv,
                                                     const int32_t* __restrict__ row_of, int T, int H, int R,
                                                     uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t nd = (size_t)R, nt = 3 * (size_t)T, nh = 4 * (size_t)H;
  if (q >= nd + nt + nh) return;
  int r;
  uint32_t val;
  if (q < nd) {
    r = (int)q;
    val = (uint32_t)(3 * (size_t)T + 4 * (size_t)H + q);
  } else if (q < nd + nt) {
    const size_t p = q - nd;
    r = row_of[tris[p]];
    val = (uint32_t)p;
  } else {
    const size_t p = q - nd - nt;
    r = row_of[hv[p]];
    val = (uint32_t)(3 * (size_t)T + p);
  }
  keys[q] = r >= 0 ? (uint32_t)r : NO_ROW;
  vals[q] = val;
}

// ---- scatter maps of a given pattern (context creation) ----------------------------------------------
// block position of (r, c) in the CSR pattern, NO_ROW when outside it
__device__ inline uint32_t block_of(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col_idx, int r,
                                    int c) {
  const uint32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
  const uint32_t b = b0 + lower_bound(col_idx + b0, b1 - b0, (uint32_t)c);
  return (b < b1 && col_idx[b] == (uint32_t)c) ? b : NO_ROW;
}

// Kind flags of the list entries (membrane record < m_end <= hinge record < h_end <= row record), with a zero
// entry at n so that the exclusive sums end with the totals.
__global__ void __launch_bounds__(256) k_kind_flags(const uint32_t* __restrict__ bc_idx, uint32_t n, uint32_t m_end,
                                                    uint32_t h_end, uint32_t* __restrict__ fm,
                                                    uint32_t* __restrict__ fh, uint32_t* __restrict__ fr) {
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j > n) return;
  const uint32_t rec = j < n ? bc_idx[j] : 0;
  fm[j] = j < n && rec < m_end;
  fh[j] = j < n && rec >= m_end && rec < h_end;
  fr[j] = j < n && rec >= h_end;
}

// rec_pos[bc_idx[j]] = the slot of entry j in its kind's region (membranes from 0, hinges from hinge0, rows
// from row0): its rank among the entries of its kind before it
__global__ void __launch_bounds__(256) k_record_positions(const uint32_t* __restrict__ bc_idx, uint32_t n,
                                                          uint32_t m_end, uint32_t h_end,
                                                          const uint32_t* __restrict__ sm,
                                                          const uint32_t* __restrict__ sh,
                                                          const uint32_t* __restrict__ sr, uint32_t hinge0,
                                                          uint32_t row0, uint32_t* __restrict__ rec_pos) {
  const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const uint32_t rec = bc_idx[j];
  rec_pos[rec] = rec < m_end ? sm[j] : (rec < h_end ? hinge0 + sh[j] : row0 + sr[j]);
}

// kind_ptr[3 b + (0, 1, 2)] = the first row, membrane and hinge slot of block b in their regions (b = 0 .. B)
__global__ void __launch_bounds__(256) k_kind_ptrs(const uint32_t* __restrict__ bc_ptr, int B,
                                                   const uint32_t* __restrict__ sm, const uint32_t* __restrict__ sh,
                                                   const uint32_t* __restrict__ sr, uint32_t* __restrict__ kind_ptr) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b > B) return;
  const uint32_t p = bc_ptr[b];
  kind_ptr[3 * (size_t)b] = sr[p];
  kind_ptr[3 * (size_t)b + 1] = sm[p];
  kind_ptr[3 * (size_t)b + 2] = sh[p];
}

__global__ void __launch_bounds__(256) k_map_pairs(const int4* __restrict__ tri, const int4* __restrict__ hin,
                                                   const int32_t* __restrict__ row_of, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ col_idx, int T, int H, int R,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   int* __restrict__ bad) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t nd = (size_t)R, nt = 9 * (size_t)T, nh = 16 * (size_t)H;
  if (q >= nd + nt + nh) return;
  int r, c;
  uint32_t val;
  if (q < nd) {
    r = c = (int)q;
    val = (uint32_t)(9 * (size_t)T + 16 * (size_t)H + q);
  } else if (q < nd + nt) {
    const size_t p = q - nd;
    const int4 t = tri[p / 9];
    const int vs[3] = {t.x, t.y, t.z};
    r = row_of[vs[(p % 9) / 3]];
    c = row_of[vs[p % 3]];
    val = (uint32_t)p;
  } else {
    const size_t p = q - nd - nt;
    const int4 h = hin[p / 16];
    const int vs[4] = {h.x, h.y, h.z, h.w};
    r = row_of[vs[(p % 16) / 4]];
    c = row_of[vs[p % 4]];
    val = (uint32_t)(9 * (size_t)T + p);
  }
  uint32_t b = NO_ROW;
  if (r >= 0 && c >= 0) {
    b = block_of(row_ptr, col_idx, r, c);
    if (b == NO_ROW) bad[0] = 1;
  }
  keys[q] = b;
  vals[q] = val;
}

__global__ void __launch_bounds__(256) k_map_forces(const int4* __restrict__ tri, const int4* __restrict__ hin,
                                                    const int32_t* __restrict__ row_of, int T, int H, int R,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t nd = (size_t)R, nt = 3 * (size_t)T, nh = 4 * (size_t)H;
  if (q >= nd + nt + nh) return;
  int r;
  uint32_t val;
  if (q < nd) {
    r = (int)q;
    val = (uint32_t)(3 * (size_t)T + 4 * (size_t)H + q);
  } else if (q < nd + nt) {
    const size_t p = q - nd;
    const int4 t = tri[p / 3];
    const int vs[3] = {t.x, t.y, t.z};
    r = row_of[vs[p % 3]];
    val = (uint32_t)p;
  } else {
    const size_t p = q - nd - nt;
    const int4 h = hin[p / 4];
    const int vs[4] = {h.x, h.y, h.z, h.w};
    r = row_of[vs[p % 4]];
    val = (uint32_t)(3 * (size_t)T + p);
  }
  keys[q] = r >= 0 ? (uint32_t)r : NO_ROW;
  vals[q] = val;
}

inline int grid_of(size_t n) { return n == 0 ? 1 : (int)((n + 255) / 256); }

// stable sort of (key, value) pairs, keeping the valued entries below `limit`: the sorted values and,
// for each key in [0, count], the first position with that key or above (ptr, count + 1 entries)
int lists_from_pairs(Pool& pool, uint32_t* keys, uint32_t* vals, size_t n, int count, hipStream_t s,
                     uint32_t* ptr_out, uint32_t** idx_out, uint32_t* total) {
  uint32_t *ks = nullptr, *vs = nullptr, *cnt = nullptr;
  HIP_TRY(pool.get(&ks, n));
  HIP_TRY(pool.get(&vs, n));
  HIP_TRY(pool.get(&cnt, 1));
  TRY_RC(sort_pairs(pool, keys, ks, vals, vs, (int)n, s));
  k_count_below<<<1, 64, 0, s>>>(ks, (uint32_t)n, NO_ROW, cnt);
  k_ptr_from_sorted<<<grid_of((size_t)count + 1), 256, 0, s>>>(ks, (uint32_t)n, count, ptr_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(total, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *idx_out = vs;
  return DRAPE_OK;
}

}  // namespace

int build_topology_device(int V, const double* rest_xy, int T, const uint32_t* tris_host, const uint8_t* free_host,
                          double density, double bend_wale, double bend_course, MeshTopology& out) {
  if (V <= 0 || T < 0) {
    set_error("device builder: invalid sizes");
    return DRAPE_E_INVALID;
  }
  if ((size_t)V + 9 * (size_t)T + 16 * (size_t)(3 * (size_t)T / 2 + 1) >= (size_t)INT32_MAX) {
    set_error("device builder: the mesh is too large for 32-bit sort counts");
    return DRAPE_E_INVALID;
  }
  Pool pool;
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  double* rest = nullptr;
  uint32_t* tris = nullptr;
  uint8_t* free_mask = nullptr;
  HIP_TRY(pool.get(&rest, 2 * (size_t)V));
  HIP_TRY(pool.get(&tris, 3 * (size_t)T));
  HIP_TRY(pool.get(&free_mask, (size_t)V));
  HIP_TRY(hipMemcpyAsync(rest, rest_xy, 2 * (size_t)V * sizeof(double), hipMemcpyHostToDevice, s));
  if (T > 0) HIP_TRY(hipMemcpyAsync(tris, tris_host, 3 * (size_t)T * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(free_mask, free_host, (size_t)V, hipMemcpyHostToDevice, s));

  // rows: free vertices in vertex order
  uint32_t *vflags = nullptr, *vexcl = nullptr, *row_vertex = nullptr;
  int32_t* row_of = nullptr;
  HIP_TRY(pool.get(&vflags, V));
  HIP_TRY(pool.get(&vexcl, V));
  HIP_TRY(pool.get(&row_of, V));
  HIP_TRY(pool.get(&row_vertex, V));
  k_free_flags<<<grid_of(V), 256, 0, s>>>(free_mask, V, vflags);
  TRY_RC(exclusive_sum(pool, vflags, vexcl, V, s));
  k_rows<<<grid_of(V), 256, 0, s>>>(vflags, vexcl, V, row_of, row_vertex);
  HIP_TRY(hipGetLastError());
  uint32_t R = 0;
  TRY_RC(flagged_total(vflags, vexcl, V, s, &R));
  if (R == 0) {
    set_error("device builder: no free vertex");
    return DRAPE_E_INVALID;
  }

  // hinges
  const int ne = 3 * T;
  uint64_t *ekeys = nullptr, *ekeys_s = nullptr;
  uint32_t *evals = nullptr, *evals_s = nullptr, *hflags = nullptr, *hpos = nullptr;
  HIP_TRY(pool.get(&ekeys, ne));
  HIP_TRY(pool.get(&ekeys_s, ne));
  HIP_TRY(pool.get(&evals, ne));
  HIP_TRY(pool.get(&evals_s, ne));
  HIP_TRY(pool.get(&hflags, ne));
  HIP_TRY(pool.get(&hpos, ne));
  k_edge_keys<<<grid_of(ne), 256, 0, s>>>(tris, T, ekeys, evals);
  TRY_RC(sort_pairs(pool, ekeys, ekeys_s, evals, evals_s, ne, s));
  k_hinge_flags<<<grid_of(ne), 256, 0, s>>>(ekeys_s, ne, hflags);
  TRY_RC(exclusive_sum(pool, hflags, hpos, ne, s));
  HIP_TRY(hipGetLastError());
  uint32_t H = 0;
  TRY_RC(flagged_total(hflags, hpos, ne, s, &H));
  uint32_t *hv = nullptr, *hv_sorted = nullptr, *hidx = nullptr, *hidx2 = nullptr, *hidx3 = nullptr;
  uint64_t *key_lo = nullptr, *key_lo_s = nullptr, *key_hi = nullptr, *key_hi_s = nullptr;
  float* hstiff = nullptr;
  HIP_TRY(pool.get(&hv, 4 * (size_t)H));
  HIP_TRY(pool.get(&hv_sorted, 4 * (size_t)H));
  HIP_TRY(pool.get(&hidx, H));
  HIP_TRY(pool.get(&hidx2, H));
  HIP_TRY(pool.get(&hidx3, H));
  HIP_TRY(pool.get(&key_lo, H));
  HIP_TRY(pool.get(&key_lo_s, H));
  HIP_TRY(pool.get(&key_hi, H));
  HIP_TRY(pool.get(&key_hi_s, H));
  HIP_TRY(pool.get(&hstiff, H));
  k_hinge_emit<<<grid_of(ne), 256, 0, s>>>(ekeys_s, evals_s, hflags, hpos, ne, tris, rest, hv, key_lo, hidx);
  // lexicographic order of [v0, v1, v2, v3]: stable sorts by (v2, v3), then by (v0, v1)
  TRY_RC(sort_pairs(pool, key_lo, key_lo_s, hidx, hidx2, (int)H, s));
  k_hinge_key_hi<<<grid_of(H), 256, 0, s>>>(hidx2, hv, (int)H, key_hi);
  TRY_RC(sort_pairs(pool, key_hi, key_hi_s, hidx2, hidx3, (int)H, s));
  k_hinge_gather<<<grid_of(H), 256, 0, s>>>(hidx3, hv, (int)H, rest, bend_wale, bend_course, hv_sorted, hstiff);
  HIP_TRY(hipGetLastError());

  // element geometry and lumped masses
  float* geom = nullptr;
  double *area = nullptr, *mass = nullptr;
  uint32_t *vk = nullptr, *vv = nullptr, *vk_s = nullptr, *vv_s = nullptr;
  HIP_TRY(pool.get(&geom, 7 * (size_t)T));
  HIP_TRY(pool.get(&area, T));
  HIP_TRY(pool.get(&mass, V));
  HIP_TRY(pool.get(&vk, ne));
  HIP_TRY(pool.get(&vv, ne));
  HIP_TRY(pool.get(&vk_s, ne));
  HIP_TRY(pool.get(&vv_s, ne));
  k_elements<<<grid_of(T), 256, 0, s>>>(tris, T, rest, geom, area, vk, vv);
  TRY_RC(sort_pairs(pool, vk, vk_s, vv, vv_s, ne, s));
  k_vertex_mass<<<grid_of(V), 256, 0, s>>>(vk_s, vv_s, (uint32_t)ne, V, area, density, mass);
  HIP_TRY(hipGetLastError());

  // pattern and block contributions
  const size_t np = (size_t)R + 9 * (size_t)T + 16 * (size_t)H;
  uint64_t *pk = nullptr, *pk_s = nullptr;
  uint32_t *pv = nullptr, *pv_s = nullptr, *bflags = nullptr, *bpos = nullptr, *cnt = nullptr;
  HIP_TRY(pool.get(&pk, np));
  HIP_TRY(pool.get(&pk_s, np));
  HIP_TRY(pool.get(&pv, np));
  HIP_TRY(pool.get(&pv_s, np));
  HIP_TRY(pool.get(&cnt, 1));
  k_block_pairs<<<grid_of(np), 256, 0, s>>>(tris, hv_sorted, row_of, T, (int)H, (int)R, pk, pv);
  TRY_RC(sort_pairs(pool, pk, pk_s, pv, pv_s, (int)np, s));
  k_count_below<<<1, 64, 0, s>>>(pk_s, (uint32_t)np, NO_KEY, cnt);
  HIP_TRY(hipGetLastError());
  uint32_t nvalid = 0;
  HIP_TRY(hipMemcpyAsync(&nvalid, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(pool.get(&bflags, nvalid));
  HIP_TRY(pool.get(&bpos, nvalid));
  k_block_flags<<<grid_of(nvalid), 256, 0, s>>>(pk_s, (int)nvalid, bflags);
  TRY_RC(exclusive_sum(pool, bflags, bpos, (int)nvalid, s));
  uint32_t B = 0;
  TRY_RC(flagged_total(bflags, bpos, (int)nvalid, s, &B));
  uint32_t *col_idx = nullptr, *bcp = nullptr, *row_ptr = nullptr, *diag = nullptr;
  HIP_TRY(pool.get(&col_idx, B));
  HIP_TRY(pool.get(&bcp, (size_t)B + 1));
  HIP_TRY(pool.get(&row_ptr, (size_t)R + 1));
  HIP_TRY(pool.get(&diag, R));
  k_block_emit<<<grid_of(nvalid), 256, 0, s>>>(pk_s, bflags, bpos, (int)nvalid, col_idx, bcp, row_ptr, diag);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(bcp + B, &nvalid, sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(row_ptr + R, &B, sizeof(uint32_t), hipMemcpyHostToDevice, s));

  // force contributions
  const size_t nf = (size_t)R + 3 * (size_t)T + 4 * (size_t)H;
  uint32_t *fk = nullptr, *fv = nullptr, *fcp = nullptr, *fci = nullptr;
  HIP_TRY(pool.get(&fk, nf));
  HIP_TRY(pool.get(&fv, nf));
  HIP_TRY(pool.get(&fcp, (size_t)R + 1));
  k_force_pairs<<<grid_of(nf), 256, 0, s>>>(tris, hv_sorted, row_of, T, (int)H, (int)R, fk, fv);
  HIP_TRY(hipGetLastError());
  uint32_t nfv = 0;
  TRY_RC(lists_from_pairs(pool, fk, fv, nf, (int)R, s, fcp, &fci, &nfv));

  // back to the host
  out.R = (int)R;
  out.H = (int)H;
  out.B = (int)B;
  out.hinge_vertices.resize(4 * (size_t)H);
  out.hinge_stiffness.resize(H);
  out.elem_geom.resize(7 * (size_t)T);
  out.mass.resize(V);
  out.row_vertex.resize(R);
  out.row_ptr.resize((size_t)R + 1);
  out.col_idx.resize(B);
  out.diag_idx.resize(R);
  out.block_contrib_ptr.resize((size_t)B + 1);
  out.block_contrib_idx.resize(nvalid);
  out.force_contrib_ptr.resize((size_t)R + 1);
  out.force_contrib_idx.resize(nfv);
  auto down = [&](auto& vec, const void* src) -> int {
    if (!vec.empty())
      HIP_TRY(hipMemcpyAsync(vec.data(), src, vec.size() * sizeof(vec[0]), hipMemcpyDeviceToHost, s));
    return DRAPE_OK;
  };
  TRY_RC(down(out.hinge_vertices, hv_sorted));
  TRY_RC(down(out.hinge_stiffness, hstiff));
  TRY_RC(down(out.elem_geom, geom));
  TRY_RC(down(out.mass, mass));
  TRY_RC(down(out.row_vertex, row_vertex));
  TRY_RC(down(out.row_ptr, row_ptr));
  TRY_RC(down(out.col_idx, col_idx));
  TRY_RC(down(out.diag_idx, diag));
  TRY_RC(down(out.block_contrib_ptr, bcp));
  TRY_RC(down(out.block_contrib_idx, pv_s));
  TRY_RC(down(out.force_contrib_ptr, fcp));
  TRY_RC(down(out.force_contrib_idx, fci));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

int build_maps_device(Context& c, const std::vector<uint32_t>& row_ptr_h, const std::vector<uint32_t>& col_idx_h,
                      const std::vector<int>& row_of_h) {
  hipStream_t s = c.stream;
  const int T = c.T, H = c.H, R = c.R;
  const size_t np = (size_t)R + 9 * (size_t)T + 16 * (size_t)H;
  const size_t nf = (size_t)R + 3 * (size_t)T + 4 * (size_t)H;
  if (np >= (size_t)INT32_MAX || 9 * (size_t)T + 16 * (size_t)H + R >= (size_t)NO_ROW) {
    set_error("scatter maps: the problem is too large for 32-bit sort counts");
    return DRAPE_E_INVALID;
  }
  Pool pool;
  uint32_t *rp = nullptr, *ci = nullptr, *pk = nullptr, *pv = nullptr, *fk = nullptr, *fv = nullptr;
  int32_t* row_of = nullptr;
  int* bad = nullptr;
  HIP_TRY(pool.get(&rp, (size_t)R + 1));
  HIP_TRY(pool.get(&ci, (size_t)c.B));
  HIP_TRY(pool.get(&row_of, (size_t)c.V));
  HIP_TRY(pool.get(&pk, np));
  HIP_TRY(pool.get(&pv, np));
  HIP_TRY(pool.get(&fk, nf));
  HIP_TRY(pool.get(&fv, nf));
  HIP_TRY(pool.get(&bad, 1));
  HIP_TRY(hipMemcpyAsync(rp, row_ptr_h.data(), ((size_t)R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(ci, col_idx_h.data(), (size_t)c.B * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(row_of, row_of_h.data(), (size_t)c.V * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), s));
  k_map_pairs<<<grid_of(np), 256, 0, s>>>(c.tri, c.hin, row_of, rp, ci, T, H, R, pk, pv, bad);
  k_map_forces<<<grid_of(nf), 256, 0, s>>>(c.tri, c.hin, row_of, T, H, R, fk, fv);
  HIP_TRY(hipGetLastError());
  int bad_h = 0;
  HIP_TRY(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (bad_h) {
    set_error("an element couples rows outside the pattern");
    return DRAPE_E_INVALID;
  }
  HIP_TRY(hipMalloc((void**)&c.bc_ptr, ((size_t)c.B + 1) * sizeof(uint32_t)));
  HIP_TRY(hipMalloc((void**)&c.fc_ptr, ((size_t)R + 1) * sizeof(uint32_t)));
  uint32_t *bci = nullptr, *fci = nullptr, nb = 0, nfc = 0;
  TRY_RC(lists_from_pairs(pool, pk, pv, np, c.B, s, c.bc_ptr, &bci, &nb));
  TRY_RC(lists_from_pairs(pool, fk, fv, nf, R, s, c.fc_ptr, &fci, &nfc));
  HIP_TRY(hipMalloc((void**)&c.bc_idx, std::max<size_t>(nb, 1) * sizeof(uint32_t)));
  HIP_TRY(hipMalloc((void**)&c.fc_idx, std::max<size_t>(nfc, 1) * sizeof(uint32_t)));
  HIP_TRY(hipMemcpyAsync(c.bc_idx, bci, (size_t)nb * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(c.fc_idx, fci, (size_t)nfc * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  // the block-sorted layout: every record's slot in its kind's region (each record is in at most one list) and
  // each block's first slot per kind
  {
    const uint32_t m_end = 9 * (uint32_t)T, h_end = m_end + 16 * (uint32_t)H;
    uint32_t *fm = nullptr, *fh = nullptr, *fr = nullptr, *sm = nullptr, *sh = nullptr, *sr = nullptr;
    for (uint32_t** q : {&fm, &fh, &fr, &sm, &sh, &sr}) HIP_TRY(pool.get(q, (size_t)nb + 1));
    k_kind_flags<<<grid_of((size_t)nb + 1), 256, 0, s>>>(c.bc_idx, nb, m_end, h_end, fm, fh, fr);
    TRY_RC(exclusive_sum(pool, fm, sm, (int)nb + 1, s));
    TRY_RC(exclusive_sum(pool, fh, sh, (int)nb + 1, s));
    TRY_RC(exclusive_sum(pool, fr, sr, (int)nb + 1, s));
    uint32_t totals[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&totals[0], sm + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&totals[1], sh + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c.hinge0 = totals[0];
    c.row0 = totals[0] + totals[1];
    HIP_TRY(hipMalloc((void**)&c.rec_pos, np * sizeof(uint32_t)));
    HIP_TRY(hipMalloc((void**)&c.kind_ptr, 3 * ((size_t)c.B + 1) * sizeof(uint32_t)));
    HIP_TRY(hipMemsetAsync(c.rec_pos, 0xff, np * sizeof(uint32_t), s));
    if (nb > 0)
      k_record_positions<<<grid_of(nb), 256, 0, s>>>(c.bc_idx, nb, m_end, h_end, sm, sh, sr, c.hinge0, c.row0,
                                                     c.rec_pos);
    k_kind_ptrs<<<grid_of((size_t)c.B + 1), 256, 0, s>>>(c.bc_ptr, c.B, sm, sh, sr, c.kind_ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
  }
  c.block_contribs = nb;
  c.force_contribs = nfc;
  return DRAPE_OK;
}

}  // namespace drape
