// Internal declarations shared by the HIP kernels and the host side of libdrape_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/drape_hip.h"

namespace drape {

constexpr int WG = 256;          // rows per workgroup and per reduction group (pcg.rs:216 WORKGROUP)
constexpr int SLICE = 64;        // rows per sliced-BSR slice: one wavefront
constexpr int CHUNK_PLAIN = 32;  // PCG iterations between readbacks without replacement (solver.ts:145-149)
constexpr uint32_t NO_SLOT = 0xffffffffu;

// Scalar CG state on the device (state.wgsl:4-9), rz double-buffered by iteration parity so that every
// workgroup of a kernel reads the previous value while workgroup 0 writes the next one.
struct PcgState {
  float rz[2];
  float pq, alpha, beta, rr, rr0, tol2, true_rr, above2;
  int32_t iteration, status, max_iterations, replace_pending;
  float diverge2;  // rr > diverge2 rr0 is a breakdown (two-level solves; +inf otherwise)
};

// Two-level preconditioner data (coarse space) -- see twolevel.hip.
struct TwoLevel;
// Training buffers of the learned prolongation -- see gnn_train.hip.
struct GnnTrain;
// Self-contact broad phase, candidates and pair terms -- see selfcontact.hip.
struct SelfContact;

struct Context {
  // sizes
  int V = 0, T = 0, H = 0, R = 0, B = 0, K = 0;  // K colliders
  int G = 0;                                     // reduction groups = ceil(R / 256)
  int S = 0;                                     // slices = ceil(R / 64)
  bool linear = false;                           // plain BSR system (no mesh)
  drape_rules rules{};
  drape_problem problem{};                       // pointers not kept; sizes and constants only
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_spmv0 = nullptr, ev_spmv1 = nullptr;

  // host copies needed by the Newton loop
  std::vector<uint32_t> row_vertex;
  std::vector<float> unit_mass, colliders_host, row_mass_host;
  std::vector<uint32_t> row_ptr_host, col_idx_host;
  std::vector<float> rest_host;   // rest positions [3V]

  // geometry (device)
  float4* pos = nullptr;      // [V]
  float4* xhat = nullptr;     // [V]
  float4* dir = nullptr;      // [V]
  int4* tri = nullptr;        // [T] v0 v1 v2 -
  float4* geom0 = nullptr;    // [T] g0x g0y g1x g1y
  float4* geom1 = nullptr;    // [T] g2x g2y area -
  int4* hin = nullptr;        // [H]
  float* hstiff = nullptr;    // [H]
  uint32_t* rvert = nullptr;  // [R]
  float* rmass = nullptr;     // [R] M/h^2
  float* colliders = nullptr; // [4 + 8K]
  float* mu = nullptr;        // [1]
  float* unit_mass_dev = nullptr; // [R] vertex masses
  float4* org = nullptr;      // [V] origin x_n of the Newton solve / accepted PTC state
  float* material = nullptr;  // [4] wale course coupling shear

  // element-local outputs (reference layout): grad [9T + 12H + 3R], hess [81T + 144H + 9R]
  float* lgrad = nullptr;
  float* lhess = nullptr;
  // scatter maps sorted by row (e3_generate.rs:111-152 order): per CSR block / per row
  uint32_t* bc_ptr = nullptr;
  uint32_t* bc_idx = nullptr;
  uint32_t* fc_ptr = nullptr;
  uint32_t* fc_idx = nullptr;
  // The block-sorted layout of the element-local Hessian blocks for the Newton path: three regions, the
  // membrane blocks, then the hinge blocks (from hinge0), then the row blocks (from row0), each in bc_idx order
  // (by BSR block, then by list position). rec_pos: the slot of every block (record id as in bc_idx: 9e + 3k + l,
  // 9T + 16h + 4k + l, 9T + 16H + row), NO_SLOT for a block in no BSR block (a pinned vertex's); kind_ptr
  // [3 (B + 1)]: per BSR block the first slot of its row, membrane and hinge blocks within their regions. The
  // element kernels write each block to its slot, so each kernel fills whole lines of its own region and the
  // assembly reads each BSR block's three runs contiguously, in list order (row, membranes, hinges).
  uint32_t* rec_pos = nullptr;
  uint32_t* kind_ptr = nullptr;
  uint32_t hinge0 = 0, row0 = 0;
  bool hess_sorted = false;       // lhess holds the block-sorted records (else the reference's element layout)

  // sliced BSR (SELL-64): per slice pointer (in block slots of 64 lanes) and width
  uint32_t* slice_ptr = nullptr;  // [S+1]
  uint32_t* sell_col = nullptr;   // [slots * 64]
  float* sell_val = nullptr;      // [slots * 64 * 9]: for slot base q, entry e of lane l at 9q*64... see spmv
  uint32_t* row_cnt = nullptr;    // [G*256] blocks of the row at each physical position (0 beyond R)
  uint32_t* perm = nullptr;       // [G*256] physical position -> logical row (~0u beyond R)
  uint32_t* diag_slot = nullptr;  // [R] sell_val base index of the row's diagonal block
  uint32_t* block_dst = nullptr;  // [B] sell_val base index of CSR block b (entry e at + e*64)
  uint32_t* csr_row_ptr = nullptr;// [R+1]
  uint32_t* csr_col = nullptr;    // [B]
  size_t sell_slots = 0;

  // PCG vectors (float4 per row, w = 0)
  float4 *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr;
  float* dinv = nullptr;          // [9R]
  float* partials = nullptr;      // [2 * 3 * G]: pq | (rz, rr) | replace/true
  PcgState* state = nullptr;
  PcgState* state_host = nullptr; // pinned
  int32_t* poll_host = nullptr;   // pinned [2]: PCG status copies polled while the next iterations run
  hipEvent_t ev_poll[2] = {nullptr, nullptr};
  double* refine = nullptr;       // FP64 refinement vectors [3R] x2
  float4* b_saved = nullptr;      // the right-hand side while the refinement loop solves for corrections
  float4* z1 = nullptr;           // V-cycle: pre-smoothed correction [G*256]
  float4* r1 = nullptr;           // V-cycle: residual after pre-smoothing [G*256]
  float4* vel = nullptr;          // [V] velocities of the dynamic loop (allocated on first use)

  // reductions for the Newton loop (FP64)
  double* dpart = nullptr;        // [2 * max groups]
  double* dout = nullptr;         // [8]
  double* dout_host = nullptr;    // pinned
  float* bounds = nullptr;        // [R]
  float* dists = nullptr;         // [R]
  float* fpart = nullptr;         // [4 * groups]
  int32_t* ipart = nullptr;       // [2 * groups]

  bool curvature_dropped = false;
  TwoLevel* two_level = nullptr;
  GnnTrain* gnn_train = nullptr;  // training buffers of the learned prolongation (allocated on first use)
  SelfContact* self_contact = nullptr;  // self-contact state (allocated on first use under rules.self_contact)
  uint64_t system_version = 0;    // bumped by every assembly: the coarse operator follows it
  uint64_t shift_version = 0;     // bumped when mu or the step size of M/h^2 changes
  float shift_mu = 0.0f;
  double shift_h = 0.0;
  int64_t fallbacks = 0;          // two-level solves redone with block-Jacobi (non-SPD A_c, or fallback = 1)
  int64_t truncations = 0;        // solves returned truncated at negative curvature (status 4)
  int64_t breakdowns = 0;         // solves that reported a breakdown (status 3) to their caller
  int64_t read_bytes = 0;         // device-to-host bytes read back by the solve loops
  int64_t block_contribs = 0;     // entries of bc_idx and fc_idx
  int64_t force_contribs = 0;

  // barrier stiffness: the start value (colliders[2] at creation) and the value of each step of the last
  // drape_simulate
  double kappa0 = 0.0;
  std::vector<double> kappa_history;
  std::vector<float> step_distance;  // smallest vertex-collider distance after each step of the last drape_simulate

  // Newton iterates kept by a probe (positions after each accepted update, [count][3V]) when keep_iterates
  bool keep_iterates = false;
  std::vector<float> iterates;

  // accepted PTC states kept for the caller: every state when keep_states, else the start and the last
  bool keep_states = false;
  std::vector<float> kept_states;  // [count][3V]

  // timing of the last PCG solve
  double last_spmv_ms = 0, last_solve_ms = 0;
  int last_spmv_launches = 0;
  bool time_spmv = false;
};

// error handling
void set_error(const std::string& message);
#define HIP_TRY(expr)                                                                           \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) {                                                                     \
      ::drape::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                    \
      return DRAPE_E_DEVICE;                                                                    \
    }                                                                                           \
  } while (0)

#define TRY_RC(expr)           \
  do {                         \
    int _rc = (expr);          \
    if (_rc != DRAPE_OK) return _rc; \
  } while (0)

// Synchronous device-to-host read on the context's stream, counted in read_bytes (the readbacks the
// reference's timed backend counts, timed-backend.ts).
inline int read_back(Context& c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  c.read_bytes += (int64_t)bytes;
  return DRAPE_OK;
}

// ---- device problem builder (builder.hip) -----------------------------------------------------
// A sheet's topology built on the device from its rest positions (f64 x y per vertex), triangles and
// free-vertex mask: sorted oriented hinges and their stiffness, element geometry, f64 lumped masses of
// every vertex (density times area), the rows, the BSR pattern with diagonal positions, and the
// per-block and per-row contribution lists in the reference's units and order.
struct MeshTopology {
  int R = 0, H = 0, B = 0;
  std::vector<uint32_t> hinge_vertices, row_vertex, row_ptr, col_idx, diag_idx;
  std::vector<uint32_t> block_contrib_ptr, block_contrib_idx, force_contrib_ptr, force_contrib_idx;
  std::vector<float> hinge_stiffness, elem_geom;
  std::vector<double> mass;
};
int build_topology_device(int V, const double* rest_xy, int T, const uint32_t* tris, const uint8_t* free_mask,
                          double density, double bend_wale, double bend_course, MeshTopology& out);
// The context's scatter maps (bc_ptr/bc_idx, fc_ptr/fc_idx) for its pattern, built on the device.
int build_maps_device(Context& c, const std::vector<uint32_t>& row_ptr, const std::vector<uint32_t>& col_idx,
                      const std::vector<int>& row_of);

// ---- kernel launchers (elements.hip) --------------------------------------------------------
// element + row kernels into lgrad/lhess: lhess in the block-sorted order of the assembly, or with `reference`
// in the reference's element-local layout (81T + 144H + 9R, for the kernel checks)
int launch_evaluate(Context& c, bool exact, bool reference = false);
int launch_assemble(Context& c);                    // scatter maps -> sell_val and b
int launch_rhs_norm(Context& c, double* norm2);     // FP64 |b|^2 (synchronises)
int launch_slope(Context& c, double* slope);        // FP64 -sum b.x (synchronises)
int launch_scatter_direction(Context& c, const float4* rows_src);  // dir[vertex(row)] = rows_src[row]
int launch_advance(Context& c, float alpha);        // pos[v] += alpha * dir[v] on rows
int launch_energy_change(Context& c, float alpha, double* sum, double* abs_sum);
int launch_max_step(Context& c, float* alpha, int32_t* row, float* min_distance, int32_t* min_row);
int launch_min_distance(Context& c, float* min_distance);  // smallest vertex-collider distance at pos
// out = {max|solution row|, max|pos - org| on rows, max|pos| on rows, non-finite count}
int launch_row_maxes(Context& c, const float4* org, float out[4]);
int launch_towards(Context& c);                     // dir = x_hat - pos (all vertices)
int launch_step_all(Context& c, float alpha);       // pos += alpha dir (all vertices)
int launch_ptc_prepare(Context& c, const float4* base, double h);  // x_hat, M/h^2 of a PTC attempt
int spmv_timing(Context& c, int launches, double* avg_ms);
int dir_spmv_timing(Context& c, int launches, double* avg_ms);  // k_pcg_dir_spmv alone
int copy_timing(size_t bytes, int launches, double* avg_ms);  // known-size copy kernel
int launch_set_float(Context& c, float* dst, float v);  // device scalar write by kernel argument
int launch_dyn_predict(Context& c, const float4* vel, double h);  // x_hat = x + h v + h^2 g on the free rows
int launch_dyn_velocity(Context& c, float4* vel, double h);       // v = (pos - org) / h

// ---- self-contact (selfcontact.hip) ---------------------------------------------------------
// All no-ops unless rules.self_contact. evaluate: broad phase within dhat at the current positions and the
// pair terms; assemble: their gradient into b and diagonal blocks into the matrix (after launch_assemble's
// kernels); view: the coupling the SpMV adds (left null without terms); step_bound / energy_change: over
// the pairs that can come within dhat along the current direction (found once per direction; moved() marks
// positions or direction as changed).
void self_contact_destroy(SelfContact* sc);
int self_contact_evaluate(Context& c);
int self_contact_assemble(Context& c);
void self_contact_view(const Context& c, const uint32_t** cptr, const uint32_t** cinc, const int4** crow,
                       const float4** cdat, const float4** ccoef);
void self_contact_moved(Context& c);
int self_contact_step_bound(Context& c, float* alpha, int32_t* row);
int self_contact_energy_change(Context& c, float alpha, double* sum, double* abs_sum);
int self_contact_pairs(Context& c, int32_t* n_pt, int32_t* n_ee, int32_t* vertices, double* distances);
void self_contact_set_kappa(Context& c, double kappa);  // follows the colliders' kappa

// ---- PCG (pcg.hip) --------------------------------------------------------------------------
int pcg_solve(Context& c, double tolerance, int max_iterations, drape_pcg_stats* stats);

// ---- two-level preconditioner (twolevel.hip) ------------------------------------------------
void two_level_destroy(TwoLevel* t);
int two_level_create(Context& c, int columns);           // coarse layout of a grid sheet, bilinear weights
int two_level_set_weights(Context& c, const float* w);  // [4R] parent-slot weights (NULL: bilinear)
int two_level_prepare(Context& c, bool* ok);            // Galerkin + factorisation when the system changed
int two_level_before_assembly(Context& c);              // an assembly waits until a side factorisation has read A
int two_level_launch_deferred(Context& c);              // queue a decided side factorisation (after the solve's first launches)
int two_level_apply(Context& c, float* partials, int guard);  // z = M^-1 r and partial (r.z, r.r)
// One two-level PCG iteration after its SpMV: alpha from the p.q partials, x += alpha p, r -= alpha q and
// z = M^-1 r with partial (r.z, r.r), in the restriction, the coarse solve and the prolongation launches.
int two_level_iterate(Context& c, const float* partial_pq, const float4* p, float* partial_rz, int it);
// The coarse correction's parts for the V-cycle: yc = B^T B P^T r (guarded like two_level_apply) and the
// device arrays the prolongation reads.
struct CoarseView {
  const int32_t* ppos;    // [4R] position of each parent slot's node in yc (the coarse factor's ordering), -1 if none
  const float* w;         // [4R] prolongation weights
  const float4* lever;    // [4R] rotation levers of the slots (null without rotation unknowns)
  const float* yc;        // the coarse solution, 3 (or 6 with rotations) unknowns per coarse node
};
CoarseView two_level_view(const Context& c);
int two_level_dof(const Context& c);  // unknowns per coarse node: 3, or 6 with rotations
int two_level_rotations(Context& c, int32_t* dof, float* levers, int64_t* replaced);
int two_level_restrict_solve(Context& c, const float4* r, int guard, bool fp32 = false);  // fp32: FP32 B always
void two_level_observe(Context& c, int iterations, double relative_residual);
int two_level_galerkin(Context& c, double* out);       // A_c of the assembled system, row-major, to the host
int two_level_inverse_factor(Context& c, float* out);  // B = L^-1 of the current factorisation, to the host
int two_level_ordering(Context& c, int32_t* order);    // the natural unknown at each position of the factor
int two_level_gemv_timing(Context& c, int which, int launches, double* avg_ms, double* bytes);
int two_level_factor_timing(Context& c, int launches, double* avg_ms, double* flops);
int two_level_set_model(Context& c, const float* theta, int count);  // learned prolongation (NULL: none)
// Factorises the assembled system with the current weights (the network is not run), unconditionally.
int two_level_factor_weights(Context& c, bool* ok);
float* two_level_weights_device(Context& c);  // [4R] current weights on the device
int two_level_weights(Context& c, float* out);         // [4R] weights of the current factorisation
int two_level_info(Context& c, int32_t* coarse_rows, int32_t* parent, float* weights, int64_t* factorisations,
                   int64_t* failures, double* factor_ms);

// ---- learned prolongation (gnn.hip) ---------------------------------------------------------
struct GnnArgs {
  hipStream_t stream;
  int R;
  const float* sell_val;
  const uint32_t* block_dst;
  const uint32_t* diag_slot;
  const uint32_t* row_ptr;  // CSR pattern [R + 1]
  const uint32_t* col;      // [B]
  const float2* fxy;        // [R] position inside the coarse cell
  const float* wbil;        // [4R] bilinear weights (0 on empty slots)
  const float* theta;       // parameters
  float *scale, *features, *edge, *h0, *h1, *msg;  // scratch: [R], [13R], [B], [32R] x 3
  float* w;                 // [4R] output weights
  float* sigma;             // [4R] output rotation scales (the rotation head; null: the network has none)
};
constexpr int GNN_FEATURES = 13, GNN_WIDTH = 32, GNN_LAYERS = 3;
// Parameters of the network: input layer, three layers and the weight head; the rotation head (Wr [4 x 32],
// br [4] after them) adds 4 x 32 + 4.
int gnn_param_count(bool rotation_head = false);
int gnn_infer(const GnnArgs& a);
int gnn_features(const GnnArgs& a);  // scale, features and normalised edge weights only
// gnn_infer with every activation kept: hs [4][32R] (h0..h3), ms [3][32R] (the messages of the three
// layers); the weights are bit-identical to gnn_infer's.
int gnn_forward_keep(const GnnArgs& a, float* hs, float* ms);
GnnArgs two_level_gnn_args(Context& c);  // the context's network inputs, scratch and weight buffers

// ---- training of the learned prolongation (gnn_train.hip) -----------------------------------
void gnn_train_destroy(GnnTrain* g);
int gnn_train_set_vectors(Context& c, const float* vectors, int count);
int gnn_train_make_vectors(Context& c, int count, int sweeps, float omega, uint64_t seed, float* out);
int gnn_train_gradient(Context& c, const float* params, int count, double* loss, float* grad);

// ---- host state (context.cpp) -----------------------------------------------------------------
int set_kappa(Context& c, float kappa);  // the colliders' barrier stiffness, host copy and device

// ---- helpers --------------------------------------------------------------------------------
inline int blocks_for(int n, int per = WG) { return (n + per - 1) / per; }

// Records the regularisation mu and the step size h of the mass term (h <= 0: unchanged); a change
// makes the next two-level solve refactor its coarse operator.
inline void note_shift(Context& c, float mu, double h) {
  const bool changed = mu != c.shift_mu || (h > 0 && h != c.shift_h);
  c.shift_mu = mu;
  if (h > 0) c.shift_h = h;
  if (changed) ++c.shift_version;
}

}  // namespace drape
