// Context creation and the C ABI entry points that wrap single device operations.
//
// drape_create turns the reference's problem arrays into the MI355X layout: float4 vertex records for
// single 16-byte gathers, per-block scatter lists sorted by row (the reference's contribution order,
// e3_generate.rs:118-152, so assembly stays deterministic and atomic-free), and a sliced BSR matrix with
// one wavefront per 64-row slice.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../internal.h"

namespace drape {

static thread_local std::string g_error;
void set_error(const std::string& message) { g_error = message; }

template <class T> static int alloc(T** p, size_t n) {
  if (n == 0) n = 1;
  HIP_TRY(hipMalloc((void**)p, n * sizeof(T)));
  return DRAPE_OK;
}
template <class T> static int upload(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n == 0) return DRAPE_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return DRAPE_OK;
}
#define TRY(expr)                  \
  do {                             \
    int _rc = (expr);              \
    if (_rc != DRAPE_OK) return _rc; \
  } while (0)

int set_kappa(Context& c, float kappa) {
  c.colliders_host[2] = kappa;
  self_contact_set_kappa(c, kappa);
  HIP_TRY(hipMemcpyAsync(c.colliders + 2, &c.colliders_host[2], sizeof(float), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

static void free_context(Context* c) {
  if (!c) return;
  // first: a factorisation still running on the two-level side stream reads the matrix and the GNN inputs
  if (c->two_level) two_level_destroy(c->two_level);
  c->two_level = nullptr;
  void* ptrs[] = {c->pos, c->xhat, c->dir, c->tri, c->geom0, c->geom1, c->hin, c->hstiff, c->rvert, c->rmass,
                  c->colliders, c->mu, c->unit_mass_dev, c->org, c->material, c->lgrad, c->lhess, c->bc_ptr,
                  c->bc_idx, c->fc_ptr, c->fc_idx, c->rec_pos, c->kind_ptr, c->slice_ptr, c->sell_col, c->sell_val, c->row_cnt, c->perm, c->diag_slot,
                  c->block_dst, c->csr_row_ptr, c->csr_col, c->b, c->x, c->r, c->z, c->p, c->p2, c->q, c->dinv,
                  c->partials, c->state, c->refine, c->b_saved, c->z1, c->r1, c->vel, c->dpart, c->dout, c->bounds, c->dists, c->fpart, c->ipart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->state_host) (void)hipHostFree(c->state_host);
  if (c->poll_host) (void)hipHostFree(c->poll_host);
  if (c->dout_host) (void)hipHostFree(c->dout_host);
  if (c->gnn_train) gnn_train_destroy(c->gnn_train);
  if (c->self_contact) self_contact_destroy(c->self_contact);
  for (hipEvent_t e : {c->ev0, c->ev1, c->ev_spmv0, c->ev_spmv1, c->ev_poll[0], c->ev_poll[1]})
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Sliced BSR layout and the per-block destination of the CSR order. Inside each 256-row reduction group
// the rows are placed by decreasing block count (stable), so that the 64-row slices carry little padding;
// the kernels keep every reduction in logical row order through the permutation.
struct SellLayout {
  std::vector<uint32_t> block_dst, slice_ptr, sell_col, cnt, perm, diag_slot;
};

static int build_sell(Context& c, const uint32_t* row_ptr, const uint32_t* col_idx, SellLayout& L) {
  const int R = c.R;
  const int G = blocks_for(R);
  const int P = G * WG;  // physical positions
  c.S = P / SLICE;
  L.perm.assign(P, 0xffffffffu);
  L.cnt.assign(P, 0);
  std::vector<uint32_t> phys_of(R);
  for (int g = 0; g < G; ++g) {
    std::vector<int> rows;
    for (int r = g * WG; r < std::min(R, (g + 1) * WG); ++r) rows.push_back(r);
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) {
      return row_ptr[a + 1] - row_ptr[a] > row_ptr[b + 1] - row_ptr[b];
    });
    for (size_t t = 0; t < rows.size(); ++t) {
      const int ph = g * WG + (int)t;
      L.perm[ph] = (uint32_t)rows[t];
      L.cnt[ph] = row_ptr[rows[t] + 1] - row_ptr[rows[t]];
      phys_of[rows[t]] = (uint32_t)ph;
    }
  }
  L.slice_ptr.assign(c.S + 1, 0);
  for (int sl = 0; sl < c.S; ++sl) {
    uint32_t width = 0;
    for (int l = 0; l < SLICE; ++l) width = std::max(width, L.cnt[sl * SLICE + l]);
    L.slice_ptr[sl + 1] = L.slice_ptr[sl] + width;
  }
  c.sell_slots = L.slice_ptr[c.S];
  if (c.sell_slots * 9 * SLICE >= (size_t)UINT32_MAX) {
    set_error("matrix too large for 32-bit slot indices");
    return DRAPE_E_INVALID;
  }
  L.sell_col.assign(c.sell_slots * SLICE, 0);
  L.block_dst.assign(c.B, 0);
  L.diag_slot.assign(R, 0);
  for (int row = 0; row < R; ++row) {
    const uint32_t ph = phys_of[row];
    const int sl = (int)(ph / SLICE), l = (int)(ph % SLICE);
    bool has_diag = false;
    for (uint32_t b = row_ptr[row]; b < row_ptr[row + 1]; ++b) {
      if (b > row_ptr[row] && col_idx[b] <= col_idx[b - 1]) {
        set_error("col_idx must be strictly increasing within each row");
        return DRAPE_E_INVALID;
      }
      if (col_idx[b] >= (uint32_t)R) {
        set_error("col_idx out of range");
        return DRAPE_E_INVALID;
      }
      const size_t q = L.slice_ptr[sl] + (b - row_ptr[row]);
      L.sell_col[q * SLICE + l] = col_idx[b];
      L.block_dst[b] = (uint32_t)(q * 9 * SLICE + l);
      if (col_idx[b] == (uint32_t)row) {
        L.diag_slot[row] = L.block_dst[b];
        has_diag = true;
      }
    }
    if (!has_diag) {
      set_error("row " + std::to_string(row) + " has no diagonal block");
      return DRAPE_E_INVALID;
    }
  }
  return DRAPE_OK;
}

// Allocations shared by mesh and linear contexts.
static int alloc_solver(Context& c, const std::vector<uint32_t>& row_ptr, const std::vector<uint32_t>& col_idx) {
  SellLayout L;
  TRY(build_sell(c, row_ptr.data(), col_idx.data(), L));
  c.G = blocks_for(c.R);
  hipStream_t s = c.stream;
  TRY(alloc(&c.slice_ptr, L.slice_ptr.size()));
  TRY(alloc(&c.sell_col, L.sell_col.size()));
  TRY(alloc(&c.sell_val, c.sell_slots * 9 * SLICE));
  TRY(alloc(&c.row_cnt, L.cnt.size()));
  TRY(alloc(&c.perm, L.perm.size()));
  TRY(alloc(&c.diag_slot, L.diag_slot.size()));
  TRY(alloc(&c.block_dst, L.block_dst.size()));
  TRY(alloc(&c.csr_row_ptr, row_ptr.size()));
  TRY(alloc(&c.csr_col, col_idx.size()));
  TRY(upload(c.slice_ptr, L.slice_ptr.data(), L.slice_ptr.size(), s));
  TRY(upload(c.sell_col, L.sell_col.data(), L.sell_col.size(), s));
  HIP_TRY(hipMemsetAsync(c.sell_val, 0, c.sell_slots * 9 * SLICE * sizeof(float), s));
  TRY(upload(c.row_cnt, L.cnt.data(), L.cnt.size(), s));
  TRY(upload(c.perm, L.perm.data(), L.perm.size(), s));
  TRY(upload(c.diag_slot, L.diag_slot.data(), L.diag_slot.size(), s));
  TRY(upload(c.block_dst, L.block_dst.data(), L.block_dst.size(), s));
  TRY(upload(c.csr_row_ptr, row_ptr.data(), row_ptr.size(), s));
  TRY(upload(c.csr_col, col_idx.data(), col_idx.size(), s));
  const size_t Rp = (size_t)c.G * WG;
  for (float4** v : {&c.b, &c.x, &c.r, &c.z, &c.p, &c.p2, &c.q}) {
    TRY(alloc(v, Rp));
    HIP_TRY(hipMemsetAsync(*v, 0, Rp * sizeof(float4), s));
  }
  TRY(alloc(&c.dinv, 9 * Rp));
  TRY(alloc(&c.partials, 8 * (size_t)c.G + 16));
  TRY(alloc(&c.state, 1));
  HIP_TRY(hipHostMalloc((void**)&c.state_host, sizeof(PcgState)));
  HIP_TRY(hipHostMalloc((void**)&c.poll_host, 2 * sizeof(int32_t)));
  HIP_TRY(hipHostMalloc((void**)&c.dout_host, 16 * sizeof(double)));
  TRY(alloc(&c.dout, 16));
  TRY(alloc(&c.fpart, 4 * (size_t)c.G + 16));
  TRY(alloc(&c.ipart, 2 * (size_t)c.G + 16));
  TRY(alloc(&c.refine, 6 * Rp));
  TRY(alloc(&c.b_saved, Rp));
  TRY(alloc(&c.z1, Rp));
  TRY(alloc(&c.r1, Rp));
  c.row_ptr_host = row_ptr;
  c.col_idx_host = col_idx;
  HIP_TRY(hipEventCreate(&c.ev0));
  HIP_TRY(hipEventCreate(&c.ev1));
  HIP_TRY(hipEventCreate(&c.ev_spmv0));
  HIP_TRY(hipEventCreate(&c.ev_spmv1));
  HIP_TRY(hipEventCreateWithFlags(&c.ev_poll[0], hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&c.ev_poll[1], hipEventDisableTiming));
  // the uploads above read host vectors that go out of scope on return
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

static int create_mesh_context(const drape_problem* pr, const drape_rules* rules, Context& c) {
  c.V = pr->vertices; c.T = pr->triangles; c.H = pr->hinges; c.R = pr->rows; c.B = pr->blocks;
  if (c.V <= 0 || c.R <= 0 || c.B < c.R || c.T < 0 || c.H < 0) {
    set_error("invalid problem sizes");
    return DRAPE_E_INVALID;
  }
  if (!pr->rest_positions || !pr->row_vertex || !pr->row_ptr || !pr->col_idx || !pr->row_unit_mass ||
      (c.T > 0 && (!pr->elem_vertices || !pr->elem_geom)) || (c.H > 0 && (!pr->hinge_vertices || !pr->hinge_stiffness))) {
    set_error("missing problem array");
    return DRAPE_E_INVALID;
  }
  c.problem = *pr;
  c.rules = *rules;
  std::vector<uint32_t> row_ptr(pr->row_ptr, pr->row_ptr + c.R + 1), col_idx(pr->col_idx, pr->col_idx + c.B);
  if (row_ptr[0] != 0 || row_ptr[c.R] != (uint32_t)c.B) {
    set_error("row_ptr does not span col_idx");
    return DRAPE_E_INVALID;
  }
  std::vector<int> row_of(c.V, -1);
  for (int row = 0; row < c.R; ++row) {
    uint32_t v = pr->row_vertex[row];
    if (v >= (uint32_t)c.V || row_of[v] >= 0) {
      set_error("row_vertex out of range or repeated");
      return DRAPE_E_INVALID;
    }
    row_of[v] = row;
  }
  c.row_vertex.assign(pr->row_vertex, pr->row_vertex + c.R);
  c.unit_mass.assign(pr->row_unit_mass, pr->row_unit_mass + c.R);
  // colliders
  c.K = 0;
  if (pr->colliders && pr->collider_floats >= 4) {
    c.K = (int)pr->colliders[0];
    if (c.K < 0 || pr->collider_floats < 4 + 8 * c.K) {
      set_error("colliders array shorter than its count");
      return DRAPE_E_INVALID;
    }
    c.colliders_host.assign(pr->colliders, pr->colliders + 4 + 8 * c.K);
    c.colliders_host[3] = 0.0f;
    c.kappa0 = c.colliders_host[2];
  }
  // element records
  std::vector<int4> tri(c.T), hin(c.H);
  std::vector<float4> g0(c.T), g1(c.T), pos(c.V);
  for (int e = 0; e < c.T; ++e) {
    const uint32_t* v = pr->elem_vertices + 3 * e;
    for (int k = 0; k < 3; ++k)
      if (v[k] >= (uint32_t)c.V) { set_error("elem_vertices out of range"); return DRAPE_E_INVALID; }
    tri[e] = make_int4((int)v[0], (int)v[1], (int)v[2], 0);
    const float* g = pr->elem_geom + 7 * e;
    g0[e] = make_float4(g[0], g[1], g[2], g[3]);
    g1[e] = make_float4(g[4], g[5], g[6], 0.0f);
  }
  for (int h = 0; h < c.H; ++h) {
    const uint32_t* v = pr->hinge_vertices + 4 * h;
    for (int k = 0; k < 4; ++k)
      if (v[k] >= (uint32_t)c.V) { set_error("hinge_vertices out of range"); return DRAPE_E_INVALID; }
    hin[h] = make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
  }
  for (int v = 0; v < c.V; ++v)
    pos[v] = make_float4(pr->rest_positions[3 * v], pr->rest_positions[3 * v + 1], pr->rest_positions[3 * v + 2], 0.0f);
  c.rest_host.assign(pr->rest_positions, pr->rest_positions + 3 * (size_t)c.V);
  const int T = c.T, H = c.H, R = c.R;

  HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  hipStream_t s = c.stream;
  TRY(alloc_solver(c, row_ptr, col_idx));
  TRY(alloc(&c.pos, c.V)); TRY(alloc(&c.xhat, c.V)); TRY(alloc(&c.dir, c.V)); TRY(alloc(&c.org, c.V));
  TRY(alloc(&c.tri, c.T)); TRY(alloc(&c.geom0, c.T)); TRY(alloc(&c.geom1, c.T));
  TRY(alloc(&c.hin, c.H)); TRY(alloc(&c.hstiff, c.H));
  TRY(alloc(&c.rvert, c.R)); TRY(alloc(&c.rmass, c.R)); TRY(alloc(&c.unit_mass_dev, c.R));
  TRY(alloc(&c.colliders, 4 + 8 * (size_t)c.K)); TRY(alloc(&c.mu, 1)); TRY(alloc(&c.material, 4));
  TRY(alloc(&c.lgrad, 9 * (size_t)T + 12 * (size_t)H + 3 * (size_t)R));
  TRY(alloc(&c.lhess, 81 * (size_t)T + 144 * (size_t)H + 9 * (size_t)R));
  const int maxg = blocks_for(T) + blocks_for(H) + blocks_for(R) + 4;
  TRY(alloc(&c.dpart, 2 * (size_t)maxg));
  TRY(upload(c.pos, pos.data(), pos.size(), s));
  TRY(upload(c.xhat, pos.data(), pos.size(), s));
  TRY(upload(c.org, pos.data(), pos.size(), s));
  HIP_TRY(hipMemsetAsync(c.dir, 0, (size_t)c.V * sizeof(float4), s));
  TRY(upload(c.tri, tri.data(), tri.size(), s));
  TRY(upload(c.geom0, g0.data(), g0.size(), s));
  TRY(upload(c.geom1, g1.data(), g1.size(), s));
  TRY(upload(c.hin, hin.data(), hin.size(), s));
  TRY(upload(c.hstiff, pr->hinge_stiffness, (size_t)c.H, s));
  TRY(upload(c.rvert, pr->row_vertex, (size_t)c.R, s));
  TRY(upload(c.unit_mass_dev, pr->row_unit_mass, (size_t)c.R, s));
  if (c.K > 0) TRY(upload(c.colliders, c.colliders_host.data(), c.colliders_host.size(), s));
  else HIP_TRY(hipMemsetAsync(c.colliders, 0, 4 * sizeof(float), s));
  TRY(launch_set_float(c, c.mu, 0.0f));
  note_shift(c, 0.0f, pr->time_step);
  const float mat[4] = {(float)pr->wale, (float)pr->course, (float)pr->coupling, (float)pr->shear};
  TRY(upload(c.material, mat, 4, s));
  HIP_TRY(hipStreamSynchronize(s));
  // scatter maps sorted by row: per CSR block the local 3x3 blocks (units of 9 floats) in the order
  // row term, membranes (9e + 3k + l), hinges (9T + 16h + 4k + l); per row the local gradients
  // (units of 3 floats): row term, membranes (3e + k), hinges (3T + 4h + k) -- built on the device
  TRY(build_maps_device(c, row_ptr, col_idx, row_of));
  // M/h^2 of the problem's time step
  c.row_mass_host.resize(R);
  const double h = pr->time_step > 0 ? pr->time_step : 1.0 / 240.0;
  for (int row = 0; row < R; ++row) c.row_mass_host[row] = (float)((double)c.unit_mass[row] / (h * h));
  TRY(upload(c.rmass, c.row_mass_host.data(), (size_t)R, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

}  // namespace drape

using namespace drape;

extern "C" {

void drape_default_rules(drape_rules* r) {
  std::memset(r, 0, sizeof *r);
  // newton / pcg sections of problem.json (e3_generate.rs:12-25)
  r->newton_tolerance = 1e-6;
  r->newton_max_iterations = 30;
  r->armijo = 1e-4;
  r->max_halvings = 30;
  r->pcg_tolerance = 1e-7;
  r->pcg_max_iterations = 20000;
  r->replace_every = 50;
  r->replace_above = 1e-2;
  // robust (e4_rules.rs:6-57)
  r->eta_start = 1e-2; r->eta_min = 1e-7; r->eta_max = 1e-2; r->eta_gamma = 0.9; r->eta_safeguard = 0.1;
  r->update_stop = 1e-4; r->resolution_factor = 2.0; r->mu_start = 1e-3; r->mu_cap = 1e3;
  // ptc (e7_ptc.rs:25-28, e7_run.rs:12-33)
  r->h0 = 1.0 / 240.0; r->h_max = 1.0; r->h_min_factor = 0.1; r->growth = 2.0; r->shrink = 0.5;
  r->inner_tolerance = 0.1; r->inner_max_iterations = 10; r->max_rejections = 12;
  r->equilibrium_f32 = 1e-3; r->stagnation_window = 10; r->stagnation_factor = 0.95; r->gravity_load = 0.0;
  // RobustSettings of run-e7.ts:22 (E6 rung D)
  r->floor_stop = 1; r->forcing = 1; r->regularize = 1; r->noise_aware = 1;
  r->exact_tangent = 1;
  r->refine_fp64 = 0;
  r->preconditioner = DRAPE_PRECOND_BLOCK_JACOBI;
  r->coarse_reuse = 3;
  r->fallback = 2;
  r->coarse_shift = 0.0;
  r->smoother = 0.0;
  r->coarse_half = 2;
  r->refresh_rate = 2.0;
  r->refresh_on_breakdown = 1;
  r->self_contact = 0;
  r->coarse_async = 0;
  r->block_jacobi_truncate = 0;
  r->adaptive_kappa = 0;
  r->kappa_distance = 0.1;
  r->kappa_max_factor = 100.0;
  r->coarse_factor = 0;
  r->coarse_split = 1;
  r->mu_carry = 0;
  r->coarse_rotations = 0;
  r->coarse_pivot_floor = 0.0;
}

const char* drape_last_error(void) { return g_error.c_str(); }

int drape_create(const drape_problem* problem, const drape_rules* rules, drape_ctx** out) {
  if (!problem || !out) {
    set_error("null argument");
    return DRAPE_E_INVALID;
  }
  drape_rules defaults;
  drape_default_rules(&defaults);
  Context* c = new Context();
  int rc = create_mesh_context(problem, rules ? rules : &defaults, *c);
  if (rc == DRAPE_OK && c->rules.preconditioner == DRAPE_PRECOND_TWO_LEVEL)
    rc = two_level_create(*c, problem->grid_columns);
  if (rc != DRAPE_OK) {
    free_context(c);
    *out = nullptr;
    return rc;
  }
  *out = reinterpret_cast<drape_ctx*>(c);
  return DRAPE_OK;
}

int drape_linear_create(int32_t rows, int32_t blocks, const uint32_t* row_ptr, const uint32_t* col_idx,
                        const float* vals, const float* rhs, const drape_rules* rules, drape_ctx** out) {
  if (!row_ptr || !col_idx || !vals || !rhs || !out || rows <= 0 || blocks < rows) {
    set_error("invalid linear system");
    return DRAPE_E_INVALID;
  }
  Context* c = new Context();
  c->linear = true;
  c->R = rows;
  c->B = blocks;
  if (rules) c->rules = *rules;
  else drape_default_rules(&c->rules);
  auto fail = [&](int rc) {
    free_context(c);
    *out = nullptr;
    return rc;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(DRAPE_E_DEVICE);
  }
  std::vector<uint32_t> rp(row_ptr, row_ptr + rows + 1), ci(col_idx, col_idx + blocks);
  if (rp[0] != 0 || rp[rows] != (uint32_t)blocks) {
    set_error("row_ptr does not span col_idx");
    return fail(DRAPE_E_INVALID);
  }
  int rc = alloc_solver(*c, rp, ci);
  if (rc != DRAPE_OK) return fail(rc);
  rc = alloc(&c->dpart, 2 * (size_t)c->G + 32);
  if (rc != DRAPE_OK) return fail(rc);
  // scatter vals into the sliced layout on the host, then upload
  std::vector<uint32_t> block_dst(blocks);
  if (hipMemcpy(block_dst.data(), c->block_dst, blocks * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(DRAPE_E_DEVICE);
  std::vector<float> sell(c->sell_slots * 9 * SLICE, 0.0f);
  for (int b = 0; b < blocks; ++b)
    for (int e = 0; e < 9; ++e) sell[block_dst[b] + SLICE * e] = vals[9 * (size_t)b + e];
  std::vector<float4> b4((size_t)c->G * WG, make_float4(0, 0, 0, 0));
  for (int r = 0; r < rows; ++r) b4[r] = make_float4(rhs[3 * r], rhs[3 * r + 1], rhs[3 * r + 2], 0.0f);
  if (hipMemcpy(c->sell_val, sell.data(), sell.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->b, b4.data(), b4.size() * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("upload failed");
    return fail(DRAPE_E_DEVICE);
  }
  c->system_version = 1;
  *out = reinterpret_cast<drape_ctx*>(c);
  return DRAPE_OK;
}

void drape_destroy(drape_ctx* ctx) { free_context(reinterpret_cast<Context*>(ctx)); }

static Context* ctx_of(drape_ctx* p) { return reinterpret_cast<Context*>(p); }

static int upload_positions(Context& c, float4* dst, const float* src) {
  std::vector<float4> v(c.V);
  for (int i = 0; i < c.V; ++i) v[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0f);
  HIP_TRY(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  return DRAPE_OK;
}

int drape_set_state(drape_ctx* p, const float* positions, const float* x_hat, double h, float mu) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (positions) TRY(upload_positions(*c, c->pos, positions));
  if (positions) self_contact_moved(*c);
  if (x_hat) TRY(upload_positions(*c, c->xhat, x_hat));
  if (h > 0) {
    for (int row = 0; row < c->R; ++row) c->row_mass_host[row] = (float)((double)c->unit_mass[row] / (h * h));
    HIP_TRY(hipMemcpyAsync(c->rmass, c->row_mass_host.data(), c->R * sizeof(float), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(c->mu, &mu, sizeof(float), hipMemcpyHostToDevice, c->stream));
  note_shift(*c, mu, h);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_get_positions(drape_ctx* p, float* out) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !out) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  std::vector<float4> v(c->V);
  HIP_TRY(hipMemcpyAsync(v.data(), c->pos, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->V; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
  return DRAPE_OK;
}

int drape_set_barrier_curvature(drape_ctx* p, int32_t drop) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (c->K == 0) return DRAPE_OK;
  c->colliders_host[3] = drop ? 1.0f : 0.0f;
  c->curvature_dropped = drop != 0;
  HIP_TRY(hipMemcpyAsync(c->colliders + 3, &c->colliders_host[3], sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_get_contributions(drape_ctx* p, int64_t* sizes, uint32_t* block_contrib_ptr, uint32_t* block_contrib_idx,
                            uint32_t* force_contrib_ptr, uint32_t* force_contrib_idx) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !sizes) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  sizes[0] = c->block_contribs;
  sizes[1] = c->force_contribs;
  hipStream_t s = c->stream;
  if (block_contrib_ptr)
    HIP_TRY(hipMemcpyAsync(block_contrib_ptr, c->bc_ptr, ((size_t)c->B + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (block_contrib_idx)
    HIP_TRY(hipMemcpyAsync(block_contrib_idx, c->bc_idx, (size_t)c->block_contribs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (force_contrib_ptr)
    HIP_TRY(hipMemcpyAsync(force_contrib_ptr, c->fc_ptr, ((size_t)c->R + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (force_contrib_idx)
    HIP_TRY(hipMemcpyAsync(force_contrib_idx, c->fc_idx, (size_t)c->force_contribs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return DRAPE_OK;
}

int drape_set_colliders(drape_ctx* p, int32_t enabled) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (c->K == 0) return DRAPE_OK;
  c->colliders_host[0] = enabled ? (float)c->K : 0.0f;
  HIP_TRY(hipMemcpyAsync(c->colliders, &c->colliders_host[0], sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_set_kappa(drape_ctx* p, double kappa) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !(kappa > 0) || !std::isfinite(kappa)) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  if (c->K == 0) return DRAPE_OK;
  return set_kappa(*c, (float)kappa);
}

int drape_kappa(drape_ctx* p, double* kappa) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !kappa) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  *kappa = c->K > 0 ? (double)c->colliders_host[2] : 0.0;
  return DRAPE_OK;
}

int drape_kappa_history(drape_ctx* p, double* kappa, float* min_distance, int32_t cap, int32_t* count) {
  Context* c = ctx_of(p);
  if (!c || c->linear || cap < 0) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  const int32_t n = (int32_t)c->kappa_history.size();
  if (count) *count = n;
  for (int32_t i = 0; i < n && i < cap; ++i) {
    if (kappa) kappa[i] = c->kappa_history[i];
    if (min_distance) min_distance[i] = i < (int32_t)c->step_distance.size() ? c->step_distance[i] : INFINITY;
  }
  return DRAPE_OK;
}

int drape_set_row_mass(drape_ctx* p, const float* row_mass) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !row_mass) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  std::copy(row_mass, row_mass + c->R, c->row_mass_host.begin());
  HIP_TRY(hipMemcpyAsync(c->rmass, c->row_mass_host.data(), c->R * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  ++c->shift_version;
  c->shift_h = 0.0;
  return DRAPE_OK;
}

int drape_set_tangent(drape_ctx* p, int32_t exact) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  c->rules.exact_tangent = exact != 0;
  return DRAPE_OK;
}

int drape_device_name(char* name, int32_t cap) {
  if (!name || cap <= 0) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  std::snprintf(name, (size_t)cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return DRAPE_OK;
}

int drape_residual(drape_ctx* p, float* norm, double* norm_squared) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  TRY(launch_evaluate(*c, c->rules.exact_tangent != 0));
  TRY(launch_assemble(*c));
  double n2 = 0.0;
  TRY(launch_rhs_norm(*c, &n2));
  if (norm_squared) *norm_squared = n2;
  if (norm) *norm = (float)std::sqrt(n2);
  return DRAPE_OK;
}

static int solution_to_host(Context& c, float* out) {
  std::vector<float4> v(c.R);
  HIP_TRY(hipMemcpyAsync(v.data(), c.x, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  for (int i = 0; i < c.R; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
  return DRAPE_OK;
}

int drape_pcg(drape_ctx* p, double tolerance, int32_t max_iterations, drape_pcg_stats* stats, float* solution) {
  Context* c = ctx_of(p);
  if (!c || !stats || max_iterations <= 0 || !(tolerance > 0)) { set_error("invalid PCG arguments"); return DRAPE_E_INVALID; }
  TRY(pcg_solve(*c, tolerance, max_iterations, stats));
  if (solution) TRY(solution_to_host(*c, solution));
  return DRAPE_OK;
}

int drape_slope(drape_ctx* p, double* slope) {
  Context* c = ctx_of(p);
  if (!c || !slope) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  return launch_slope(*c, slope);
}

int drape_set_direction(drape_ctx* p, const float* rows) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (rows) {
    std::vector<float4> v((size_t)c->G * WG, make_float4(0, 0, 0, 0));
    for (int i = 0; i < c->R; ++i) v[i] = make_float4(rows[3 * i], rows[3 * i + 1], rows[3 * i + 2], 0.0f);
    HIP_TRY(hipMemcpyAsync(c->q, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    TRY(launch_scatter_direction(*c, c->q));
  } else {
    TRY(launch_scatter_direction(*c, c->x));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_energy_change(drape_ctx* p, float alpha, double* sum, double* abs_sum) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !sum || !abs_sum) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  return launch_energy_change(*c, alpha, sum, abs_sum);
}

int drape_max_step(drape_ctx* p, float* alpha, int32_t* vertex, float* min_distance) {
  Context* c = ctx_of(p);
  if (!c || c->linear || !alpha) { set_error("invalid argument"); return DRAPE_E_INVALID; }
  int32_t row = -1, min_row = -1;
  float a = 1.0f, d = INFINITY;
  TRY(launch_max_step(*c, &a, &row, &d, &min_row));
  *alpha = a;
  if (vertex) *vertex = (a < 1.0f && row >= 0) ? (int32_t)c->row_vertex[row] : -1;
  if (min_distance) *min_distance = d;
  return DRAPE_OK;
}

int drape_eval_elements(drape_ctx* p, const float* positions, float* grad, float* hess) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (positions) TRY(upload_positions(*c, c->pos, positions));
  TRY(launch_evaluate(*c, c->rules.exact_tangent != 0, true));  // the reference's element-local layout
  const size_t ng = 9 * (size_t)c->T + 12 * (size_t)c->H + 3 * (size_t)c->R;
  const size_t nh = 81 * (size_t)c->T + 144 * (size_t)c->H + 9 * (size_t)c->R;
  if (grad) HIP_TRY(hipMemcpyAsync(grad, c->lgrad, ng * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (hess) HIP_TRY(hipMemcpyAsync(hess, c->lhess, nh * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_get_system(drape_ctx* p, float* rhs, float* vals) {
  Context* c = ctx_of(p);
  if (!c) { set_error("null context"); return DRAPE_E_INVALID; }
  if (rhs) {
    std::vector<float4> v(c->R);
    HIP_TRY(hipMemcpyAsync(v.data(), c->b, v.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->R; ++i) { rhs[3 * i] = v[i].x; rhs[3 * i + 1] = v[i].y; rhs[3 * i + 2] = v[i].z; }
  }
  if (vals) {
    std::vector<float> sell(c->sell_slots * 9 * SLICE);
    std::vector<uint32_t> block_dst(c->B);
    HIP_TRY(hipMemcpyAsync(sell.data(), c->sell_val, sell.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(block_dst.data(), c->block_dst, block_dst.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int b = 0; b < c->B; ++b)
      for (int e = 0; e < 9; ++e) vals[9 * (size_t)b + e] = sell[block_dst[b] + SLICE * e];
  }
  return DRAPE_OK;
}

int drape_set_preconditioner(drape_ctx* p, int32_t kind, const float* weights) {
  Context* c = ctx_of(p);
  if (!c || (kind != DRAPE_PRECOND_BLOCK_JACOBI && kind != DRAPE_PRECOND_TWO_LEVEL)) {
    set_error("invalid preconditioner");
    return DRAPE_E_INVALID;
  }
  if (kind == DRAPE_PRECOND_TWO_LEVEL) {
    if (c->linear) {
      set_error("the two-level preconditioner needs a mesh context");
      return DRAPE_E_INVALID;
    }
    if (!c->two_level) TRY(two_level_create(*c, c->problem.grid_columns));
    TRY(two_level_set_weights(*c, weights));
  } else if (weights) {
    set_error("weights apply to the two-level preconditioner only");
    return DRAPE_E_INVALID;
  }
  c->rules.preconditioner = kind;
  return DRAPE_OK;
}

int drape_two_level_layout(drape_ctx* p, int32_t* coarse_rows, int32_t* parent, float* bilinear, int64_t* counters) {
  Context* c = ctx_of(p);
  if (!c) { set_error("null context"); return DRAPE_E_INVALID; }
  int64_t f = 0, failed = 0;
  double ms = 0.0;
  TRY(two_level_info(*c, coarse_rows, parent, bilinear, &f, &failed, &ms));
  if (counters) {
    counters[0] = f;
    counters[1] = failed;
    counters[2] = c->fallbacks;
    counters[3] = (int64_t)(1000.0 * ms);
  }
  return DRAPE_OK;
}

int drape_two_level_rotations(drape_ctx* p, int32_t* dof, float* levers, int64_t* replaced) {
  Context* c = ctx_of(p);
  if (!c) { set_error("null context"); return DRAPE_E_INVALID; }
  if (!c->two_level) { set_error("no two-level layout on this context"); return DRAPE_E_INVALID; }
  return two_level_rotations(*c, dof, levers, replaced);
}

int drape_two_level_coarse(drape_ctx* p, double* coarse) {
  Context* c = ctx_of(p);
  if (!c || !coarse) { set_error("null argument"); return DRAPE_E_INVALID; }
  if (!c->two_level) { set_error("no two-level layout on this context"); return DRAPE_E_INVALID; }
  return two_level_galerkin(*c, coarse);
}

int drape_two_level_inverse_factor(drape_ctx* p, float* factor) {
  Context* c = ctx_of(p);
  if (!c || !factor) { set_error("null argument"); return DRAPE_E_INVALID; }
  if (!c->two_level) { set_error("no two-level layout on this context"); return DRAPE_E_INVALID; }
  return two_level_inverse_factor(*c, factor);
}

int drape_two_level_ordering(drape_ctx* p, int32_t* order) {
  Context* c = ctx_of(p);
  if (!c || !order) { set_error("null argument"); return DRAPE_E_INVALID; }
  if (!c->two_level) { set_error("no two-level layout on this context"); return DRAPE_E_INVALID; }
  return two_level_ordering(*c, order);
}

int drape_set_prolongation_model(drape_ctx* p, const float* params, int32_t count) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  if (!c->two_level) TRY(two_level_create(*c, c->problem.grid_columns));
  return two_level_set_model(*c, params, count);
}

int drape_self_contact_pairs(drape_ctx* p, int32_t* n_pt, int32_t* n_ee, int32_t* vertices, double* distances) {
  Context* c = ctx_of(p);
  if (!c || c->linear) { set_error("not a mesh context"); return DRAPE_E_INVALID; }
  return self_contact_pairs(*c, n_pt, n_ee, vertices, distances);
}

static Context* training_context(drape_ctx* p) {
  Context* c = ctx_of(p);
  if (!c || c->linear) {
    set_error("not a mesh context");
    return nullptr;
  }
  if (!c->two_level && two_level_create(*c, c->problem.grid_columns) != DRAPE_OK) return nullptr;
  return c;
}

int drape_training_vectors_set(drape_ctx* p, const float* vectors, int32_t count) {
  Context* c = training_context(p);
  if (!c) return DRAPE_E_INVALID;
  return gnn_train_set_vectors(*c, vectors, count);
}

int drape_training_vectors_make(drape_ctx* p, int32_t count, int32_t sweeps, float omega, uint64_t seed,
                                float* out) {
  Context* c = training_context(p);
  if (!c) return DRAPE_E_INVALID;
  return gnn_train_make_vectors(*c, count, sweeps, omega, seed, out);
}

int drape_prolongation_gradient(drape_ctx* p, const float* params, int32_t count, double* loss, float* grad) {
  Context* c = training_context(p);
  if (!c) return DRAPE_E_INVALID;
  return gnn_train_gradient(*c, params, count, loss, grad);
}

int drape_two_level_weights(drape_ctx* p, float* weights) {
  Context* c = ctx_of(p);
  if (!c || !weights) { set_error("null argument"); return DRAPE_E_INVALID; }
  if (!c->two_level) { set_error("no two-level layout on this context"); return DRAPE_E_INVALID; }
  return two_level_weights(*c, weights);
}

int drape_solution_fp64(drape_ctx* p, double* solution) {
  Context* c = ctx_of(p);
  if (!c || !solution) { set_error("null argument"); return DRAPE_E_INVALID; }
  if (!c->rules.refine_fp64) { set_error("the FP64 solution exists only with rules.refine_fp64"); return DRAPE_E_STATE; }
  HIP_TRY(hipMemcpyAsync(solution, c->refine, 3 * (size_t)c->R * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DRAPE_OK;
}

int drape_copy_timing(int64_t bytes, int32_t launches, double* avg_ms) {
  if (bytes <= 0 || !avg_ms) { set_error("invalid copy timing arguments"); return DRAPE_E_INVALID; }
  return copy_timing((size_t)bytes, launches, avg_ms);
}

int drape_pcg_timing(drape_ctx* p, double* spmv_ms, double* solve_ms, int32_t* spmv_launches) {
  Context* c = ctx_of(p);
  if (!c) { set_error("null context"); return DRAPE_E_INVALID; }
  const int launches = spmv_launches && *spmv_launches > 0 ? *spmv_launches : 200;
  double avg = 0.0;
  TRY(spmv_timing(*c, launches, &avg));
  if (spmv_ms) *spmv_ms = avg;
  if (solve_ms) *solve_ms = c->last_solve_ms;
  if (spmv_launches) *spmv_launches = launches;
  return DRAPE_OK;
}

int drape_kernel_timing(drape_ctx* p, int32_t kernel, int32_t launches, double* avg_ms, double* bytes) {
  Context* c = ctx_of(p);
  if (!c || !avg_ms || !bytes || launches <= 0 || kernel < 0 || kernel > 5) {
    set_error("kernel timing: invalid argument");
    return DRAPE_E_INVALID;
  }
  const double R = c->R, B = c->B;
  switch (kernel) {
    case 0:
      TRY(spmv_timing(*c, launches, avg_ms));
      *bytes = 40.0 * B + 28.0 * R + 4.0;
      return DRAPE_OK;
    case 1:
      TRY(dir_spmv_timing(*c, launches, avg_ms));
      *bytes = 40.0 * B + 52.0 * R + 4.0;
      return DRAPE_OK;
    default:
      if (!c->two_level) {
        set_error("kernel timing: the coarse solve needs the two-level preconditioner");
        return DRAPE_E_INVALID;
      }
      if (kernel == 5) return two_level_factor_timing(*c, launches, avg_ms, bytes);
      return two_level_gemv_timing(*c, kernel - 2, launches, avg_ms, bytes);
  }
}

}  // extern "C"
