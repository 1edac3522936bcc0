// C ABI of the MI355X Schur LM step: problem handle, structure stage
// (DoubleSection::structureReady), kernel sequences, LM / GN drivers.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "nelson_amd.h"
#include "structure.h"
#include "structure_dev.h"

using namespace nsn;

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(NSN_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                                  \
  } while (0)

namespace {

constexpr int kTB = 64;          // Cholesky tile
constexpr int kCholMaxWG = 256;  // persistent workgroups of the factorisation
// mixed precision: at most this many refinement steps per increment, each
// run while the FP64 residual's largest entry is above kIrTol times the
// gradient's
constexpr int kIrMaxSteps = 2;
constexpr double kIrTol = 1e-10;
// workgroups of a triangular solve: the chain and at most 255 destination
// owners, one per CU, so that all of them are resident together whenever one
// of them owns several destinations
constexpr int kTrsvMaxWG = 256;
// the S assembly splits tiles into smaller parts until it has at least this
// many workgroups (four per CU) or the parts are single camera blocks
constexpr int kMinSParts = 1024;
// LM / GN sub-iterations queued between two host synchronisations (at most)
constexpr int kLmReplaysPerSync = 8;

// Device allocations of one owner.  A failed allocation returns nullptr and
// sets `failed`, so a sequence of allocations is checked once at its end.
// fail_after >= 0 makes the allocation with that index fail (the
// NSN_TEST_FAIL_ALLOC environment variable of nsn_problem_create), so that
// tests can reach every failure path.
struct DevBuf {
  std::vector<void*> ptrs;
  bool failed = false;
  long fail_after = -1;
  ~DevBuf() { reset(); }
  void reset() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    failed = false;
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (fail_after == 0 || hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();  // an out-of-memory error is not left for a later check to find
      if (fail_after >= 0) fail_after = -1;
      failed = true;
      return nullptr;
    }
    if (fail_after > 0) fail_after--;
    ptrs.push_back(p);
    return (T*)p;
  }
};

template <typename T>
int upload(T* dst, const std::vector<T>& src) {
  if (src.empty()) return NSN_OK;
  HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return NSN_OK;
}

}  // namespace

struct nsn_problem {
  int device = 0;
  hipStream_t st = nullptr;
  int C = 0, P = 0, E = 0, NU = 0, NV = 0, nW = 0, Ec = 0, prec = NSN_FP64;
  double f = 0, cx = 0, cy = 0, huber = 0;
  TileView tv{};
  TileProfile prof;  // the host copy of tv's envelope
  double factor_flops = 0.0;
  int nsb = 0;
  long long npairs = 0;
  // host structure
  std::vector<int> cam_var, pt_var, var_cam_user, var_pt_user, w_cam, w_pt, w_ref;
  // variable camera in solver order <-> interface index (camera order among
  // the variable cameras); identities unless the cameras are reordered
  std::vector<int> var_cam_api, api_cam_var;
  DevBuf mem;
  BAView view{};
  Scratch scr{};
  double *cams = nullptr, *pts = nullptr, *cams_bak = nullptr, *pts_bak = nullptr;
  void *U = nullptr, *bU = nullptr, *V = nullptr, *bV = nullptr, *W = nullptr, *Vinv = nullptr, *Y = nullptr,
       *Sd = nullptr, *bS = nullptr, *tiles = nullptr, *inv = nullptr, *presub = nullptr,
       *part = nullptr;
  double *xU = nullptr, *xV = nullptr;
  double* yV = nullptr;  // y_q = L_q^-1 (-bV_q) per variable point (k_vinv_y)
  int* chol_tasks = nullptr;  // worker task queue of the factorisation (chol_queue)
  int n_chol_tasks = 0;
  int n_part_tiles = 0;  // its chunk scratch tiles
  int *tile_order = nullptr, *sb_i = nullptr, *sb_j = nullptr, *sb_ptr = nullptr, *sb_pairs = nullptr;
  // parts of the S assembly: tile, local rectangle (rows lo, hi, cols lo, hi), block list
  int *part_tile = nullptr, *part_rect = nullptr, *part_ptr = nullptr, *part_blk = nullptr;
  int n_s_parts = 0;
  int *flags = nullptr, *counters = nullptr;
  int nflags = 0, NR = 0;
  double* parts = nullptr;  // triangular-solve destination sums, two of NR x TB
  // edge permutations and observation arrays (nsn_set_observations)
  std::vector<int> pe_order, ce_order;
  double *d_pe_obs = nullptr, *d_pe_w = nullptr, *d_ce_obs = nullptr, *d_ce_w = nullptr;
  // adjoint work buffers (allocated on the first nsn_adjoint)
  DevBuf adj_mem;
  bool zero_fresh = true;  // zeroU / zeroV still to be cleared after their allocation
  double *gU = nullptr, *gV = nullptr, *lamU = nullptr, *lamV = nullptr, *tU = nullptr, *tV = nullptr,
         *bpU = nullptr, *bpV = nullptr, *bmU = nullptr, *bmV = nullptr, *dU = nullptr, *dV = nullptr,
         *zeroU = nullptr, *zeroV = nullptr, *dobs = nullptr, *dw = nullptr, *adj_eps = nullptr;
  // iterative refinement (mixed precision)
  int ir_steps = 0;
  double ir_tol = 0.0;  // refinement stops once max |r| <= ir_tol * max |b| (0: never early)
  double *rU = nullptr, *rV = nullptr, *bU64 = nullptr, *bV64 = nullptr, *rS = nullptr, *zS = nullptr,
         *zU = nullptr;
  DevCtl* ctl = nullptr;
  DevCtl* hctl = nullptr;  // pinned mirror
  DevLog* log = nullptr;
  int log_cap = 0;
  bool linearized = false, stepped = false;
  double last_abs = 0.0;
  int chain_steps = 0;  // the factorisation's critical path in diagonal steps
  int chol_ahead = 0;   // its worker queue's look-ahead
  hipEvent_t ev[8] = {};
  // the LM sub-iteration captured once per problem (a GN iteration is
  // launched directly: measured on the ladybug workload, replaying it as a
  // graph was 2 % slower than its queued launches)
  hipGraphExec_t lm_graph = nullptr;
  bool lm_graph_tried = false;
  std::unique_ptr<DevBuf> pri_mem;  // unary priors (nsn_set_priors)
  // a block problem (nsn_problem_create_blocks) has no edges: its Hessian is
  // always imported; off-diagonal U blocks (U-U edges) are added to S
  bool blocks_only = false;
  bool imported = false;  // the current Hessian came from an import
  int n_uoff = 0;
  int *uoff_i = nullptr, *uoff_j = nullptr;
  // the same blocks on the host, known before the structure stage (the tile
  // envelope of S contains them)
  std::vector<int> uoff_i_host, uoff_j_host;
  int* tprof = nullptr;  // device copy of the tile envelope: first[NT], row_off[NT]
  void* Uoff = nullptr;

  ~nsn_problem() {
    drop_graphs();
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (hctl) (void)hipHostFree(hctl);
    if (st) (void)hipStreamDestroy(st);
  }

  size_t esz() const { return prec == NSN_MIXED ? sizeof(float) : sizeof(double); }

  // the captured graph holds the launch sequence and its arguments: it is
  // dropped when either changes
  void drop_graphs() {
    if (lm_graph) (void)hipGraphExecDestroy(lm_graph);
    lm_graph = nullptr;
    lm_graph_tried = false;
  }
};

TraceRange::TraceRange(const char* name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

namespace {

// ---- kernel sequences -----------------------------------------------------
template <typename T>
void seq_lin_full(nsn_problem* h, int gate) {
  launch_reset_lin(h->scr, h->ctl, gate, h->st);
  launch_lin_points<T>(h->view, true, (T*)h->V, (T*)h->bV, (T*)h->W, h->scr, h->ctl, gate, h->st);
  launch_lin_cams<T>(h->view, (T*)h->U, (T*)h->bU, h->scr, h->ctl, gate, h->st);
  launch_finalize_lin(h->scr, h->ctl, gate, 1, h->st);
}

template <typename T>
void seq_cost(nsn_problem* h, int gate) {
  launch_lin_points<T>(h->view, false, nullptr, nullptr, nullptr, h->scr, h->ctl, gate, h->st);
  launch_finalize_lin(h->scr, h->ctl, gate, 0, h->st);
}

// xU = (S^-1 (rhsU - W V^-1 rhsV)), xV = V^-1 (rhsV - W^T xU) with the
// current factorisation (xV += ... with add); the increment statistics of the
// back-substitution use bden when given; accU, when given, gets xU added
// (the refinement's camera update, done by the back-substitution kernel)
template <typename T>
void seq_schur_solve(nsn_problem* h, int gate, const double* rhsU, const double* rhsV, double* xU, double* xV,
                     int add, const double* bden, double* accU = nullptr) {
  hipStream_t st = h->st;
  // the two solves' solutions and destination sums start unset (reduction)
  double* far_f = h->parts;
  double* far_b = h->parts + (size_t)h->NR * kTB;
  double* const unset[4] = {h->zS, xU, far_f, far_b};
  launch_ir_reduce<T>(h->view, (const T*)h->W, (const T*)h->Vinv, rhsU, rhsV, h->rS, unset, h->tv.n, h->ctl, gate,
                      st);
  launch_trsv<T>(h->tv, (const T*)h->tiles, (const T*)h->inv, false, h->rS, h->zS, far_f, h->ctl, gate, kTrsvMaxWG,
                 st);
  launch_trsv<T>(h->tv, (const T*)h->tiles, (const T*)h->inv, true, h->zS, xU, far_b, h->ctl, gate, kTrsvMaxWG, st);
  launch_backsub<T>(h->view, (const T*)h->W, (const T*)h->Vinv, (const T*)h->bV, xU, xV, h->scr, h->ctl, gate, st,
                    rhsV, add, bden, accU, accU ? h->tv.n : 0);
}

// computeIncrement: V^-1, Z, S, factorisation, solves, back-substitution
template <typename T>
void seq_schur(nsn_problem* h, int gate, int checks, bool timed) {
  hipStream_t st = h->st;
  launch_begin_step(h->ctl, gate, st);
  if (timed) (void)hipEventRecord(h->ev[0], st);
  launch_vinv_y<T>(h->view, (const T*)h->V, (const T*)h->bV, (const T*)h->W, (T*)h->Vinv, (T*)h->Y, h->yV, h->ctl,
                   gate, st);
  if (timed) (void)hipEventRecord(h->ev[1], st);
  launch_s_diag<T>(h->view, (const T*)h->U, (const T*)h->bU, (const T*)h->Y, h->yV, (T*)h->Sd, (T*)h->bS, h->ctl,
                   gate, st);
  launch_s_tiles<T>(h->tv, h->n_s_parts, h->tile_order, h->part_tile, h->part_rect, h->part_ptr, h->part_blk, h->sb_i,
                    h->sb_j, h->sb_ptr, h->sb_pairs, (const T*)h->Sd, (const T*)h->bS, (const T*)h->Y,
                    (T*)h->tiles, h->ctl, gate, st);
  launch_add_uoff<T>(h->tv, h->n_uoff, h->uoff_i, h->uoff_j, (const T*)h->Uoff, (T*)h->tiles, h->ctl, gate, st);
  if (timed) (void)hipEventRecord(h->ev[2], st);
  // also sets the backward solve's solution and destination sums unset
  launch_chol_reset<T>(h->flags, h->nflags, h->counters, h->xU, h->parts, h->tv.n, h->ctl, gate, st);
  {
    int* pre_sub = h->flags + h->tv.ntiles;
    launch_chol<T>(h->tv, (T*)h->tiles, (T*)h->inv, (T*)h->presub, (T*)h->part, h->chol_tasks, h->n_chol_tasks,
                   h->flags, pre_sub, pre_sub + h->tv.NT, pre_sub + 2 * h->tv.NT, h->counters, h->ctl, gate,
                   kCholMaxWG, st);
  }
  if (timed) (void)hipEventRecord(h->ev[3], st);
  launch_trsv<T>(h->tv, (const T*)h->tiles, (const T*)h->inv, true, nullptr, h->xU, h->parts, h->ctl, gate,
                 kTrsvMaxWG, st);
  if (timed) (void)hipEventRecord(h->ev[4], st);
  launch_backsub<T>(h->view, (const T*)h->W, (const T*)h->Vinv, (const T*)h->bV, h->xU, h->xV, h->scr, h->ctl,
                    gate, st);
  if (timed) (void)hipEventRecord(h->ev[5], st);
  // the FP64 refinement residual is evaluated from the edges, which an
  // imported Hessian does not come from: imported systems are not refined
  const int ir = (h->prec == NSN_MIXED && !h->imported) ? h->ir_steps : 0;
  // a refinement step runs its solve while the residual it starts with is
  // above the tolerance (k_ir_decide, on the device); the first residual
  // also gives the FP64 gradient of the gain ratio
  const int gir = gate == GATE_ALWAYS ? GATE_IR_ANY : GATE_IR;
  for (int it = 0; it < ir; it++) {
    // x += M^-1 (-b - (H + D) x), M the FP32 Schur solve
    launch_ir_residual(h->view, h->xU, h->xV, h->rU, h->rV, h->bU64, h->bV64, h->ctl, gir, st);
    launch_ir_decide(h->ctl, gir, h->ir_tol, st);
    seq_schur_solve<T>(h, gir, h->rU, h->rV, h->zU, h->xV, 1, h->bV64, h->xU);
  }
  if (ir > 0) launch_xu_stats<double>(h->NU, h->xU, h->bU64, h->scr, h->ctl, gate, st);
  else launch_xu_stats<T>(h->NU, h->xU, (const T*)h->bU, h->scr, h->ctl, gate, st);
  launch_step_finish(h->scr, h->ctl, gate, checks, st);
  if (timed) (void)hipEventRecord(h->ev[6], st);
}

// one LM sub-iteration (LevenbergMarquardt.hpp:67-137) as a fixed sequence
template <typename T>
void seq_lm_subiter(nsn_problem* h) {
  seq_lin_full<T>(h, GATE_NEED_LIN);
  launch_copy_state(h->view, h->cams_bak, h->pts_bak, h->cams, h->pts, h->ctl, GATE_NEED_LIN, h->st);
  launch_lm_after_lin(h->ctl, h->log, h->st);
  seq_schur<T>(h, GATE_NOT_DONE, 1, false);
  launch_oplus(h->view, h->xU, h->xV, h->ctl, GATE_SOLVE_OK, h->st);
  seq_cost<T>(h, GATE_SOLVE_OK);
  launch_lm_control(h->ctl, h->log, h->st);
  launch_copy_state(h->view, h->cams, h->pts, h->cams_bak, h->pts_bak, h->ctl, GATE_ROLLBACK, h->st);
}

template <typename T>
void seq_gn_iter(nsn_problem* h) {
  seq_schur<T>(h, GATE_NOT_DONE, 2, false);
  launch_oplus(h->view, h->xU, h->xV, h->ctl, GATE_SOLVE_OK, h->st);
  seq_lin_full<T>(h, GATE_SOLVE_OK);
  launch_gn_after_step(h->ctl, h->log, h->st);
}

int sync_ctl(nsn_problem* h) {
  HIP_TRY(hipMemcpyAsync(h->hctl, h->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, h->st));
  HIP_TRY(hipStreamSynchronize(h->st));
  return NSN_OK;
}

int push_ctl(nsn_problem* h) {
  HIP_TRY(hipMemcpyAsync(h->ctl, h->hctl, sizeof(DevCtl), hipMemcpyHostToDevice, h->st));
  HIP_TRY(hipStreamSynchronize(h->st));
  return NSN_OK;
}

int check_device_errors(nsn_problem* h) {
  if (h->hctl->wait_timeout) return fail(NSN_EDEVICE, "a bounded device wait expired in the tile Cholesky");
  return NSN_OK;
}

// ---- structure stage ------------------------------------------------------

// the tile envelope of [S; bS^T]: tile row i starts at the first tile
// column that holds a non-zero block of its rows (the S blocks (ci, cj),
// ci < cj, of variable cameras sharing a point, at rows 6 cj, columns 6 ci;
// the extra U-U blocks); the last tile row, which holds the right-hand side,
// is dense.  A tile row with no block left of its diagonal tile starts a
// diagonal chain of the factorisation (its step depends on no earlier
// column: independent parts of a nested-dissection order); every other row
// keeps the two sub-diagonal tiles of its chain (the chain and the
// triangular solves read them).  NSN_DENSE_S=1 in the environment when a
// problem is created forces a dense layout (for comparisons: the tiles
// outside the envelope are exact zeros, so both layouts give the same
// results bit for bit).
static void tile_envelope(int NT, const std::vector<int>& sb_i, const std::vector<int>& sb_j,
                          const std::vector<int>& xi, const std::vector<int>& xj, TileProfile& p) {
  const char* de = std::getenv("NSN_DENSE_S");
  const bool dense = de != nullptr && de[0] != '\0' && de[0] != '0';
  std::vector<int> nz(NT);  // first tile column with a non-zero block
  for (int i = 0; i < NT; i++) nz[i] = i;
  auto add = [&](int ci, int cj) {
    const int tc = (6 * ci) / kTB;
    for (int ti = (6 * cj) / kTB; ti <= (6 * cj + 5) / kTB && ti < NT; ti++) nz[ti] = std::min(nz[ti], tc);
  };
  for (size_t b = 0; b < sb_i.size(); b++)
    if (sb_i[b] != sb_j[b]) add(std::min(sb_i[b], sb_j[b]), std::max(sb_i[b], sb_j[b]));
  for (size_t b = 0; b < xi.size(); b++) add(std::min(xi[b], xj[b]), std::max(xi[b], xj[b]));
  if (NT > 0) nz[NT - 1] = 0;
  if (dense) std::fill(nz.begin(), nz.end(), 0);
  p.first.assign(NT, 0);
  int cs = 0;  // start of the current chain
  for (int i = 0; i < NT; i++) {
    if (nz[i] == i) cs = i;
    p.first[i] = std::min(nz[i], std::max(i - 2, cs));
  }
  p.chain_start = chol_chain_groups(NT, p.first.data(), kCholMaxChains);
  p.row_off.assign(NT, 0);
  long long off = 0;
  for (int i = 0; i < NT; i++) {
    p.row_off[i] = (int)off;
    off += i - p.first[i] + 1;
  }
  p.ntiles = (int)off;
}

// the tile Cholesky's operations over the envelope: each tile update
// C -= A B^T is 2 TB^3 (TB^3 on a diagonal tile), each off-diagonal tile's
// triangular solve TB^3, each diagonal tile's factorisation TB^3/3
static double tile_factor_flops(const TileProfile& prof) {
  const int NT = (int)prof.first.size();
  const double t3 = (double)kTB * kTB * kTB;
  double f = 0.0;
  for (int i = 0; i < NT; i++)
    for (int j = prof.first[i]; j <= i; j++) {
      const int k0 = std::max(prof.first[i], prof.first[j]);
      const double terms = std::max(0, j - k0);
      f += i == j ? terms * t3 + t3 / 3.0 : terms * 2.0 * t3 + t3;
    }
  return f;
}

// the factorisation's critical path in diagonal steps: the largest column
// key of the queue (chol_keys) plus one; NT for a single chain
static int tile_chain_steps(const TileProfile& prof) {
  const int NT = (int)prof.first.size();
  if (NT == 0) return 0;
  const std::vector<int> key = chol_keys(NT, prof.first.data(), prof.chain_start);
  return *std::max_element(key.begin(), key.end()) + 1;
}

int build_structure(nsn_problem* h, const nsn_problem_desc* d) {
  const int C = h->C, P = h->P, E = h->E;
  HostStructure hs;
  int rc0 = device_structure(d, hs, true, h->st);
  if (rc0) return rc0;
  h->cam_var = hs.cam_var;
  h->pt_var = hs.pt_var;
  h->var_cam_user = hs.var_cam_user;
  h->var_pt_user = hs.var_pt_user;
  h->var_cam_api = hs.var_cam_api;
  h->api_cam_var.assign(hs.NU, 0);
  for (int u = 0; u < hs.NU; u++) h->api_cam_var[hs.var_cam_api[u]] = u;
  h->w_cam = hs.w_cam;
  h->w_pt = hs.w_pt;
  h->w_ref = hs.w_ref;
  h->NU = hs.NU;
  h->NV = hs.NV;
  h->nW = hs.nW;
  h->Ec = hs.Ec;
  h->npairs = hs.npairs;
  h->nsb = hs.nsb;
  const int NU = h->NU, NV = h->NV, nW = h->nW;
  std::vector<int>& pe = hs.pe;
  std::vector<int>& ce = hs.ce;
  const std::vector<int>&pt_ptr = hs.pt_ptr, &pe_cam = hs.pe_cam, &pe_wblk = hs.pe_wblk, &wptr = hs.wptr,
  &cam_ptr = hs.cam_ptr, &ce_pt = hs.ce_pt, &cw_ptr = hs.cw_ptr, &cw_idx = hs.cw_idx, &sb_i = hs.sb_i,
  &sb_j = hs.sb_j, &sb_ptr = hs.sb_ptr, &sb_pairs = hs.sb_pairs;
  const std::vector<double>&pe_obs = hs.pe_obs, &pe_w = hs.pe_w, &ce_obs = hs.ce_obs, &ce_w = hs.ce_w;
  std::vector<int> cw_pt(nW);  // the point of each camera-major block entry
  for (int i = 0; i < nW; i++) cw_pt[i] = hs.w_pt[cw_idx[i]];
  // tiles of [S; bS^T] over the envelope of S (tile_envelope)
  const int n = 6 * NU;
  TileView tv{};
  tv.n = n;
  tv.TB = kTB;
  tv.NT = (n + 1 + kTB - 1) / kTB;
  if (tv.NT > kMaxTileRows)
    return fail(NSN_EINVAL, "S of order %d has %d tile rows; the triangular solves keep at most %d", n, tv.NT,
                kMaxTileRows);
  TileProfile& prof = h->prof;
  tile_envelope(tv.NT, sb_i, sb_j, h->uoff_i_host, h->uoff_j_host, prof);
  tv.ntiles = prof.ntiles;
  h->tv = tv;
  h->factor_flops = tile_factor_flops(prof);
  h->chain_steps = tile_chain_steps(prof);
  std::vector<int> tile_order(2 * (size_t)tv.ntiles);
  for (int i = 0; i < tv.NT; i++)
    for (int j = prof.first[i]; j <= i; j++) {
      const long long t = prof.slot(i, j);
      tile_order[2 * t] = i;
      tile_order[2 * t + 1] = j;
    }
  // parts of the tiles for the S assembly: rectangles of a tile split at
  // camera boundaries every g cameras in rows and columns (a large g: one part
  // per tile).  A small S with many terms per block gets small parts, so the
  // assembly spreads over the device instead of running in a few workgroups.
  auto n_seg = [](int a, int len, int u) { return (a + len - 1) / u - a / u + 1; };
  auto count_parts = [&](int g) {
    const int u = 6 * g;
    long long np = 0;
    for (int t = 0; t < tv.ntiles; t++)
      np += (long long)n_seg(tile_order[2 * t] * kTB, kTB, u) * n_seg(tile_order[2 * t + 1] * kTB, kTB, u);
    return np;
  };
  int g = 1 << 20;
  for (int cand : {1 << 20, 4, 2, 1}) {
    g = cand;
    if (count_parts(cand) >= kMinSParts) break;
  }
  const int u = 6 * g;
  std::vector<int> part_base(tv.ntiles + 1, 0), part_tile, part_rect;
  for (int t = 0; t < tv.ntiles; t++) {
    const int r0 = tile_order[2 * t] * kTB, c0 = tile_order[2 * t + 1] * kTB;
    const int nr = n_seg(r0, kTB, u), nc = n_seg(c0, kTB, u);
    part_base[t + 1] = part_base[t] + nr * nc;
    for (int a = 0; a < nr; a++)
      for (int b = 0; b < nc; b++) {
        const int rlo = std::max(r0, (r0 / u + a) * u), rhi = std::min(r0 + kTB, (r0 / u + a + 1) * u);
        const int clo = std::max(c0, (c0 / u + b) * u), chi = std::min(c0 + kTB, (c0 / u + b + 1) * u);
        part_tile.push_back(t);
        part_rect.insert(part_rect.end(), {rlo - r0, rhi - r0, clo - c0, chi - c0});
      }
  }
  const int nparts = part_base[tv.ntiles];
  auto for_parts = [&](int sb, auto&& fn) {
    const int ci = sb_i[sb], cj = sb_j[sb];
    const int ti0 = (6 * cj) / kTB, ti1 = (6 * cj + 5) / kTB;
    const int tj0 = (6 * ci) / kTB, tj1 = (6 * ci + 5) / kTB;
    for (int ti = ti0; ti <= ti1; ti++)
      for (int tj = tj0; tj <= tj1; tj++)
        if (ti >= tj) {
          const long long t = prof.slot(ti, tj);
          const int r0 = ti * kTB, c0 = tj * kTB;
          const int sr = std::max(6 * cj, r0) / u - r0 / u, sc = std::max(6 * ci, c0) / u - c0 / u;
          fn(part_base[t] + sr * n_seg(c0, kTB, u) + sc);
        }
  };
  std::vector<int> part_ptr(nparts + 1, 0);
  for (int sb = 0; sb < h->nsb; sb++) for_parts(sb, [&](int q) { part_ptr[q + 1]++; });
  for (int q = 0; q < nparts; q++) part_ptr[q + 1] += part_ptr[q];
  std::vector<int> part_blk(std::max(part_ptr[nparts], 1));
  {
    std::vector<int> fill(part_ptr.begin(), part_ptr.end() - 1);
    for (int sb = 0; sb < h->nsb; sb++) for_parts(sb, [&](int q) { part_blk[fill[q]++] = sb; });
  }
  h->n_s_parts = nparts;
  h->NR = (n + kTB - 1) / kTB;
  const int nchain = (int)prof.chain_start.size() - 1;
  // the task count and the chunk tiles do not depend on the look-ahead
  const int n_tasks0 = chol_queue(tv.NT, 0, prof.first.data(), kCholMaxChains).ntasks;
  const int ahead = chol_pick_lookahead(tv.NT, prof.first.data(), kCholMaxChains,
                                        std::min(n_tasks0, kCholMaxWG - nchain));
  h->chol_ahead = ahead;
  CholQueue cq = chol_queue(tv.NT, ahead, prof.first.data(), kCholMaxChains);
  std::vector<int>& chol_tasks = cq.task;
  if (chol_tasks.empty()) chol_tasks.assign(kCholTaskInts, 0);
  h->n_chol_tasks = cq.ntasks;
  h->n_part_tiles = cq.nparts;
  h->tv.nchain = nchain;
  // flags of the factorisation: ready (ntiles), pre_sub / pre_diag (NT
  // each), the chunk tiles (one each)
  h->nflags = tv.ntiles + 2 * tv.NT + cq.nparts;

  // ---- device allocation ----
  DevBuf& m = h->mem;
  const size_t es = h->esz();
  auto a_int = [&](size_t k) { return m.alloc<int>(k); };
  auto a_dbl = [&](size_t k) { return m.alloc<double>(k); };
  auto a_T = [&](size_t k) { return (void*)m.alloc<char>(k * es); };
  int* d_cam_var = a_int(C);
  int* d_pt_var = a_int(P);
  int* d_vcu = a_int(NU);
  int* d_vpu = a_int(NV);
  int* d_pt_ptr = a_int(P + 1);
  int* d_pe_cam = a_int(E);
  double* d_pe_obs = a_dbl(2 * (size_t)E);
  double* d_pe_w = a_dbl(E);
  int* d_pe_wblk = a_int(E);
  int* d_pe_edge = a_int(E);
  int* d_cam_ptr = a_int(NU + 1);
  int* d_ce_pt = a_int(h->Ec);
  double* d_ce_obs = a_dbl(2 * (size_t)h->Ec);
  double* d_ce_w = a_dbl(h->Ec);
  int* d_wptr = a_int(NV + 1);
  int* d_w_cam = a_int(nW);
  int* d_w_pt = a_int(nW);
  int* d_cw_ptr = a_int(NU + 1);
  int* d_cw_idx = a_int(nW);
  int* d_cw_pt = a_int(nW);
  h->cams = a_dbl(12 * (size_t)C);
  h->pts = a_dbl(3 * (size_t)P);
  h->cams_bak = a_dbl(12 * (size_t)C);
  h->pts_bak = a_dbl(3 * (size_t)P);
  h->U = a_T(21 * (size_t)NU);
  h->bU = a_T(6 * (size_t)NU);
  h->V = a_T(6 * (size_t)NV);
  h->bV = a_T(3 * (size_t)NV);
  h->W = a_T(18 * (size_t)nW);
  h->Vinv = a_T(6 * (size_t)NV);
  h->Y = a_T(18 * (size_t)nW);
  h->yV = a_dbl(3 * (size_t)NV);
  h->Sd = a_T(21 * (size_t)NU);
  h->bS = a_T(6 * (size_t)NU);
  h->tiles = a_T((size_t)tv.ntiles * kTB * kTB);
  h->tprof = a_int(2 * (size_t)tv.NT + prof.chain_start.size());
  h->inv = a_T((size_t)tv.NT * kTB * kTB);
  h->presub = a_T((size_t)tv.NT * kTB * kTB);
  h->part = a_T((size_t)std::max(h->n_part_tiles, 1) * kTB * kTB);
  h->xU = a_dbl(6 * (size_t)NU);
  h->parts = a_dbl(2 * (size_t)h->NR * kTB);  // the triangular solves' destination sums
  h->rU = a_dbl(6 * (size_t)NU);
  h->rV = a_dbl(3 * (size_t)NV);
  h->bU64 = a_dbl(6 * (size_t)NU);
  h->bV64 = a_dbl(3 * (size_t)NV);
  h->rS = a_dbl(6 * (size_t)NU);
  h->zS = a_dbl(6 * (size_t)NU);
  h->zU = a_dbl(6 * (size_t)NU);
  h->xV = a_dbl(3 * (size_t)NV);
  h->tile_order = a_int(tile_order.size());
  h->chol_tasks = a_int(chol_tasks.size());
  h->part_tile = a_int(part_tile.size());
  h->part_rect = a_int(part_rect.size());
  h->part_ptr = a_int(part_ptr.size());
  h->part_blk = a_int(part_blk.size());
  h->sb_i = a_int(sb_i.size());
  h->sb_j = a_int(sb_j.size());
  h->sb_ptr = a_int(sb_ptr.size());
  h->sb_pairs = a_int(sb_pairs.size());
  h->flags = a_int(h->nflags);
  h->counters = a_int(8);
  const int nblk_pts = std::max(1, (int)((std::max((long long)P * kLinLanes, (long long)NV) + 255) / 256));
  const int nblk_cams = std::max(1, (6 * NU + 255) / 256);
  h->scr.part_chi2 = a_dbl(nblk_pts);
  h->scr.part_xv = a_dbl(2 * (size_t)nblk_pts);
  h->scr.part_xu = a_dbl(2 * (size_t)nblk_cams);
  h->scr.maxes = m.alloc<unsigned long long>(4);
  h->scr.nblk_pts = (int)(((long long)P * kLinLanes + 255) / 256);
  h->scr.nblk_cams = (6 * NU + 255) / 256;
  h->scr.nblk_xv = (int)(((long long)NV * kLinLanes + 255) / 256);
  h->ctl = m.alloc<DevCtl>(1);
  h->log_cap = 4096;
  h->log = m.alloc<DevLog>(h->log_cap);
  if (m.failed)
    return fail(NSN_ENOMEM, "device allocation failed (%.1f MB of tiles, %.1f MB of W blocks)",
                (double)tv.ntiles * kTB * kTB * es / 1e6, 18.0 * nW * es / 1e6);
  int rc;
  if ((rc = upload(d_cam_var, h->cam_var)) || (rc = upload(d_pt_var, h->pt_var)) ||
      (rc = upload(d_vcu, h->var_cam_user)) || (rc = upload(d_vpu, h->var_pt_user)) ||
      (rc = upload(d_pt_ptr, pt_ptr)) || (rc = upload(d_pe_cam, pe_cam)) || (rc = upload(d_pe_obs, pe_obs)) ||
      (rc = upload(d_pe_w, pe_w)) || (rc = upload(d_pe_wblk, pe_wblk)) || (rc = upload(d_pe_edge, pe)) || (rc = upload(d_cam_ptr, cam_ptr)) ||
      (rc = upload(d_ce_pt, ce_pt)) || (rc = upload(d_ce_obs, ce_obs)) || (rc = upload(d_ce_w, ce_w)) ||
      (rc = upload(d_wptr, wptr)) || (rc = upload(d_w_cam, h->w_cam)) || (rc = upload(d_w_pt, h->w_pt)) || (rc = upload(d_cw_ptr, cw_ptr)) ||
      (rc = upload(d_cw_idx, cw_idx)) || (rc = upload(d_cw_pt, cw_pt)) || (rc = upload(h->tile_order, tile_order)) || (rc = upload(h->chol_tasks, chol_tasks)) ||
      (rc = upload(h->part_tile, part_tile)) || (rc = upload(h->part_rect, part_rect)) ||
      (rc = upload(h->part_ptr, part_ptr)) || (rc = upload(h->part_blk, part_blk)) ||
      (rc = upload(h->sb_i, sb_i)) || (rc = upload(h->sb_j, sb_j)) || (rc = upload(h->sb_ptr, sb_ptr)) ||
      (rc = upload(h->sb_pairs, sb_pairs)))
    return rc;
  {
    std::vector<int> tp(prof.first);
    tp.insert(tp.end(), prof.row_off.begin(), prof.row_off.end());
    tp.insert(tp.end(), prof.chain_start.begin(), prof.chain_start.end());
    if ((rc = upload(h->tprof, tp))) return rc;
    h->tv.first = h->tprof;
    h->tv.row_off = h->tprof + tv.NT;
    h->tv.chain_start = h->tprof + 2 * (size_t)tv.NT;
  }
  BAView& v = h->view;
  v.C = C; v.P = P; v.E = E; v.NU = NU; v.NV = NV; v.nW = nW; v.w_dup = hs.w_dup ? 1 : 0;
  v.f = h->f; v.cx = h->cx; v.cy = h->cy; v.huber = h->huber;
  v.cam_var = d_cam_var; v.pt_var = d_pt_var; v.var_cam_user = d_vcu; v.var_pt_user = d_vpu;
  v.pt_ptr = d_pt_ptr; v.pe_cam = d_pe_cam; v.pe_obs = d_pe_obs; v.pe_w = d_pe_w; v.pe_wblk = d_pe_wblk; v.pe_edge = d_pe_edge;
  v.cam_ptr = d_cam_ptr; v.ce_pt = d_ce_pt; v.ce_obs = d_ce_obs; v.ce_w = d_ce_w;
  v.wptr = d_wptr; v.w_cam = d_w_cam; v.w_pt = d_w_pt; v.cw_ptr = d_cw_ptr; v.cw_idx = d_cw_idx; v.cw_pt = d_cw_pt;
  v.cams = h->cams; v.pts = h->pts;
  v.intr = nullptr;
  if (d->intrinsics) {
    double* d_intr = a_dbl(3 * (size_t)C);
    if (m.failed) return fail(NSN_ENOMEM, "device allocation failed (intrinsics)");
    if ((rc = upload(d_intr, std::vector<double>(d->intrinsics, d->intrinsics + 3 * (size_t)C)))) return rc;
    v.intr = d_intr;
  }
  h->pe_order = std::move(pe);
  h->ce_order = std::move(ce);
  h->d_pe_obs = d_pe_obs; h->d_pe_w = d_pe_w; h->d_ce_obs = d_ce_obs; h->d_ce_w = d_ce_w;
  return NSN_OK;
}

DevCtl make_ctl(int log_cap) {
  DevCtl c;
  std::memset(&c, 0, sizeof(c));
  c.log_cap = log_cap;
  return c;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" int nsn_device_count(int* n) {
  if (!n) return fail(NSN_EINVAL, "null argument");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) { *n = 0; return fail(NSN_EHIP, "hipGetDeviceCount: %s", hipGetErrorString(e)); }
  *n = c;
  return NSN_OK;
}

extern "C" int nsn_structure_check(const nsn_problem_desc* d, long long* n_diff, double* host_ms,
                                   double* device_ms) {
  TraceRange trace_range("nsn_structure_check");
  clear_error();
  if (!d || !n_diff) return fail(NSN_EINVAL, "nsn_structure_check: null argument");
  if (d->n_cameras < 1 || d->n_points < 1 || d->n_edges < 0 || (d->n_edges > 0 && (!d->cam_idx || !d->pt_idx || !d->obs)))
    return fail(NSN_EINVAL, "nsn_structure_check: bad sizes or missing edge arrays");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NSN_EHIP, "nsn_structure_check: no HIP device");
  if (d->device < 0 || d->device >= ndev) return fail(NSN_EINVAL, "nsn_structure_check: bad device %d", d->device);
  HIP_TRY(hipSetDevice(d->device));
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HostStructure a, b;
  auto t0 = std::chrono::steady_clock::now();
  int rc = host_structure(d, a, true);
  auto t1 = std::chrono::steady_clock::now();
  if (!rc) rc = device_structure(d, b, true, st);
  auto t2 = std::chrono::steady_clock::now();
  (void)hipStreamDestroy(st);
  if (rc) return rc;
  const char* what = nullptr;
  *n_diff = structure_diff(a, b, &what);
  if (host_ms) *host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (device_ms) *device_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  if (*n_diff) fail(NSN_OK, "device structure differs from the host's in %lld fields, first: %s", *n_diff, what);
  return NSN_OK;
}

static int problem_create_impl(const nsn_problem_desc* d, nsn_problem** out, long fail_after,
                               const std::vector<int>* uoff_i = nullptr, const std::vector<int>* uoff_j = nullptr);

extern "C" int nsn_problem_create(const nsn_problem_desc* d, nsn_problem** out) {
  TraceRange trace_range("nsn_problem_create");
  clear_error();
  long fa = -1;
  if (const char* e = std::getenv("NSN_TEST_FAIL_ALLOC")) fa = std::strtol(e, nullptr, 10);
  return problem_create_impl(d, out, fa);
}

static int problem_create_impl(const nsn_problem_desc* d, nsn_problem** out, long fail_after,
                               const std::vector<int>* uoff_i, const std::vector<int>* uoff_j) {
  if (!d || !out) return fail(NSN_EINVAL, "nsn_problem_create: null argument");
  *out = nullptr;
  if (d->n_cameras < 1 || d->n_points < 1 || d->n_edges < 0 || (d->n_edges > 0 && (!d->cam_idx || !d->pt_idx || !d->obs)))
    return fail(NSN_EINVAL, "nsn_problem_create: bad sizes or missing edge arrays");
  if (d->precision != NSN_FP64 && d->precision != NSN_MIXED)
    return fail(NSN_EINVAL, "nsn_problem_create: unknown precision %d", d->precision);
  if (d->intrinsics)
    for (int c = 0; c < d->n_cameras; c++)
      if (!std::isfinite(d->intrinsics[3 * c]) || d->intrinsics[3 * c] == 0.0 ||
          !std::isfinite(d->intrinsics[3 * c + 1]) || !std::isfinite(d->intrinsics[3 * c + 2]))
        return fail(NSN_EINVAL, "nsn_problem_create: camera %d: intrinsics must be finite with f != 0", c);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(NSN_EHIP, "nsn_problem_create: no HIP device");
  if (d->device < 0 || d->device >= ndev) return fail(NSN_EINVAL, "nsn_problem_create: bad device %d", d->device);
  HIP_TRY(hipSetDevice(d->device));
  std::unique_ptr<nsn_problem> h(new nsn_problem());
  h->device = d->device;
  h->C = d->n_cameras; h->P = d->n_points; h->E = d->n_edges; h->prec = d->precision;
  h->ir_steps = d->precision == NSN_MIXED ? kIrMaxSteps : 0;
  h->ir_tol = kIrTol;
  h->f = d->focal; h->cx = d->cx; h->cy = d->cy; h->huber = d->huber_delta;
  h->mem.fail_after = fail_after;
  HIP_TRY(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  for (auto& e : h->ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipHostMalloc((void**)&h->hctl, sizeof(DevCtl), hipHostMallocDefault));
  if (uoff_i && uoff_j) {
    h->uoff_i_host = *uoff_i;
    h->uoff_j_host = *uoff_j;
  }
  int rc = build_structure(h.get(), d);
  if (rc) return rc;
  *h->hctl = make_ctl(h->log_cap);
  if ((rc = push_ctl(h.get()))) return rc;
  *out = h.release();
  return NSN_OK;
}

extern "C" int nsn_problem_create_blocks(const nsn_block_desc* d, nsn_problem** out) {
  TraceRange trace_range("nsn_problem_create_blocks");
  clear_error();
  if (!d || !out) return fail(NSN_EINVAL, "nsn_problem_create_blocks: null argument");
  *out = nullptr;
  if (d->n_u < 1 || d->n_v < 0 || d->n_w < 0 || d->n_uoff < 0 || (d->n_w > 0 && (!d->w_u || !d->w_v)) ||
      (d->n_uoff > 0 && (!d->uoff_i || !d->uoff_j)))
    return fail(NSN_EINVAL, "nsn_problem_create_blocks: bad sizes or missing block arrays");
  if (d->precision != NSN_FP64 && d->precision != NSN_MIXED)
    return fail(NSN_EINVAL, "nsn_problem_create_blocks: unknown precision %d", d->precision);
  std::vector<long long> seen;
  seen.reserve(std::max(d->n_w, d->n_uoff));
  for (int k = 0; k < d->n_w; k++) {
    if (d->w_u[k] < 0 || d->w_u[k] >= d->n_u || d->w_v[k] < 0 || d->w_v[k] >= d->n_v)
      return fail(NSN_EINVAL, "nsn_problem_create_blocks: W block %d at (%d, %d) out of range", k, d->w_u[k], d->w_v[k]);
    seen.push_back((long long)d->w_u[k] * d->n_v + d->w_v[k]);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(NSN_EINVAL, "nsn_problem_create_blocks: a W block is listed twice");
  seen.clear();
  for (int k = 0; k < d->n_uoff; k++) {
    if (d->uoff_i[k] < 0 || d->uoff_j[k] >= d->n_u || d->uoff_i[k] >= d->uoff_j[k])
      return fail(NSN_EINVAL, "nsn_problem_create_blocks: U block %d at (%d, %d) is not strictly upper", k,
                  d->uoff_i[k], d->uoff_j[k]);
    seen.push_back((long long)d->uoff_i[k] * d->n_u + d->uoff_j[k]);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(NSN_EINVAL, "nsn_problem_create_blocks: an off-diagonal U block is listed twice");
  // the problem is built as a section whose edges are the W blocks: every
  // U and V parameter variable; with no V parameter, one fixed placeholder
  // point keeps the point arrays non-empty (it has no block)
  const int P = std::max(d->n_v, 1);
  std::vector<double> obs(2 * (size_t)std::max(d->n_w, 1), 0.0);
  std::vector<unsigned char> pt_fixed(P, 0);
  if (d->n_v == 0) pt_fixed[0] = 1;
  nsn_problem_desc bd{};
  bd.n_cameras = d->n_u;
  bd.n_points = P;
  bd.n_edges = d->n_w;
  bd.cam_idx = d->w_u;
  bd.pt_idx = d->w_v;
  bd.obs = obs.data();
  bd.pt_fixed = pt_fixed.data();
  bd.focal = 1.0;
  bd.precision = d->precision;
  bd.device = d->device;
  // the U-U blocks, stored in UID order (row-major (i, j)); the tile
  // envelope of S is laid out with them
  std::vector<int> ui(d->n_uoff), uj(d->n_uoff);
  {
    std::vector<int> ord(d->n_uoff);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
      return d->uoff_i[a] != d->uoff_i[b] ? d->uoff_i[a] < d->uoff_i[b] : d->uoff_j[a] < d->uoff_j[b];
    });
    for (int k = 0; k < d->n_uoff; k++) {
      ui[k] = d->uoff_i[ord[k]];
      uj[k] = d->uoff_j[ord[k]];
    }
  }
  nsn_problem* h = nullptr;
  int rc = problem_create_impl(&bd, &h, -1, &ui, &uj);
  if (rc) return rc;
  std::unique_ptr<nsn_problem> hp(h);
  h->blocks_only = true;
  if (d->n_uoff > 0) {
    DevBuf& m = h->mem;
    h->uoff_i = m.alloc<int>(d->n_uoff);
    h->uoff_j = m.alloc<int>(d->n_uoff);
    h->Uoff = m.alloc<char>(36 * (size_t)d->n_uoff * h->esz());
    if (m.failed) return fail(NSN_ENOMEM, "nsn_problem_create_blocks: device allocation failed");
    if ((rc = upload(h->uoff_i, ui)) || (rc = upload(h->uoff_j, uj))) return rc;
    h->n_uoff = d->n_uoff;
  }
  *out = hp.release();
  return NSN_OK;
}

extern "C" int nsn_problem_destroy(nsn_problem* h) {
  if (!h) return NSN_OK;
  (void)hipSetDevice(h->device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  delete h;
  return NSN_OK;
}

extern "C" int nsn_problem_info_get(nsn_problem* h, nsn_problem_info* info) {
  if (!h || !info) return fail(NSN_EINVAL, "null argument");
  info->n_var_cameras = h->NU;
  info->n_var_points = h->NV;
  info->n_wblocks = h->nW;
  info->n_s = 6 * h->NU;
  info->n_s_blocks = h->nsb;
  info->n_s_pairs = h->npairs;
  info->n_tiles = h->tv.ntiles;
  info->n_tiles_dense = h->tv.NT * (h->tv.NT + 1) / 2;
  info->factor_flops = h->factor_flops;
  info->n_chains = h->tv.nchain;
  info->chain_steps = h->chain_steps;
  info->lookahead = h->chol_ahead;
  info->chunk_tiles = h->n_part_tiles;
  return NSN_OK;
}

extern "C" int nsn_problem_plan(const nsn_problem_desc* d, nsn_problem_info* info, int* cam_order) {
  if (!d || !info) return fail(NSN_EINVAL, "null argument");
  HostStructure hs;
  int rc = host_structure(d, hs, true);
  if (rc) return rc;
  const int n = 6 * hs.NU;
  const int NT = (n + 1 + kTB - 1) / kTB;
  if (NT > kMaxTileRows)
    return fail(NSN_EINVAL, "S of order %d has %d tile rows; the triangular solves keep at most %d", n, NT,
                kMaxTileRows);
  TileProfile prof;
  tile_envelope(NT, hs.sb_i, hs.sb_j, {}, {}, prof);
  std::memset(info, 0, sizeof(*info));
  info->n_var_cameras = hs.NU;
  info->n_var_points = hs.NV;
  info->n_wblocks = hs.nW;
  info->n_s = n;
  info->n_s_blocks = hs.nsb;
  info->n_s_pairs = hs.npairs;
  info->n_tiles = prof.ntiles;
  info->n_tiles_dense = NT * (NT + 1) / 2;
  info->factor_flops = tile_factor_flops(prof);
  info->n_chains = (int)prof.chain_start.size() - 1;
  info->chain_steps = tile_chain_steps(prof);
  const CholQueue cq0 = chol_queue(NT, 0, prof.first.data(), kCholMaxChains);
  info->lookahead = chol_pick_lookahead(NT, prof.first.data(), kCholMaxChains,
                                        std::min(cq0.ntasks, kCholMaxWG - info->n_chains));
  info->chunk_tiles = cq0.nparts;
  if (cam_order)
    for (int u = 0; u < hs.NU; u++) cam_order[u] = hs.var_cam_user[u];
  return NSN_OK;
}

extern "C" int nsn_set_state_rt(nsn_problem* h, const double* cams_rt, const double* pts) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h || !cams_rt || !pts) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpy(h->cams, cams_rt, sizeof(double) * 12 * (size_t)h->C, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->pts, pts, sizeof(double) * 3 * (size_t)h->P, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->cams_bak, cams_rt, sizeof(double) * 12 * (size_t)h->C, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->pts_bak, pts, sizeof(double) * 3 * (size_t)h->P, hipMemcpyHostToDevice));
  h->linearized = false;
  h->stepped = false;
  return NSN_OK;
}

extern "C" int nsn_set_state(nsn_problem* h, const double* cams, const double* pts) {
  if (!h || !cams || !pts) return fail(NSN_EINVAL, "null argument");
  std::vector<double> rt(12 * (size_t)h->C);
  for (int c = 0; c < h->C; c++) {
    rot_exp(cams + 6 * c, &rt[12 * c]);
    for (int i = 0; i < 3; i++) rt[12 * c + 9 + i] = cams[6 * c + 3 + i];
  }
  return nsn_set_state_rt(h, rt.data(), pts);
}

extern "C" int nsn_get_state_rt(nsn_problem* h, double* cams_rt, double* pts) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->st));
  if (cams_rt) HIP_TRY(hipMemcpy(cams_rt, h->cams, sizeof(double) * 12 * (size_t)h->C, hipMemcpyDeviceToHost));
  if (pts) HIP_TRY(hipMemcpy(pts, h->pts, sizeof(double) * 3 * (size_t)h->P, hipMemcpyDeviceToHost));
  return NSN_OK;
}

extern "C" int nsn_get_state(nsn_problem* h, double* cams, double* pts) {
  if (!h) return fail(NSN_EINVAL, "null argument");
  std::vector<double> rt(12 * (size_t)h->C);
  int rc = nsn_get_state_rt(h, rt.data(), pts);
  if (rc) return rc;
  if (cams)
    for (int c = 0; c < h->C; c++) {
      rot_log(&rt[12 * c], cams + 6 * c);
      for (int i = 0; i < 3; i++) cams[6 * c + 3 + i] = rt[12 * c + 9 + i];
    }
  return NSN_OK;
}

extern "C" int nsn_backup(nsn_problem* h) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  launch_copy_state(h->view, h->cams_bak, h->pts_bak, h->cams, h->pts, h->ctl, GATE_ALWAYS, h->st);
  HIP_TRY(hipStreamSynchronize(h->st));
  return NSN_OK;
}

extern "C" int nsn_rollback(nsn_problem* h) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  launch_copy_state(h->view, h->cams, h->pts, h->cams_bak, h->pts_bak, h->ctl, GATE_ALWAYS, h->st);
  HIP_TRY(hipStreamSynchronize(h->st));
  return NSN_OK;
}

static int reset_ctl_primitive(nsn_problem* h) {
  // primitives run with the gates open: done = 0, need_lin = 1
  int rc = sync_ctl(h);
  if (rc) return rc;
  h->hctl->done = 0;
  h->hctl->mode = 0;
  return push_ctl(h);
}

extern "C" int nsn_linearize(nsn_problem* h, int with_hessian, double* chi2) {
  TraceRange trace_range("nsn_linearize");
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  h->imported = false;
  HIP_TRY(hipSetDevice(h->device));
  int rc = reset_ctl_primitive(h);
  if (rc) return rc;
  if (h->prec == NSN_MIXED) {
    if (with_hessian) seq_lin_full<float>(h, GATE_ALWAYS); else seq_cost<float>(h, GATE_ALWAYS);
  } else {
    if (with_hessian) seq_lin_full<double>(h, GATE_ALWAYS); else seq_cost<double>(h, GATE_ALWAYS);
  }
  HIP_TRY(hipGetLastError());
  if ((rc = sync_ctl(h))) return rc;
  if (with_hessian) h->linearized = true;
  if (chi2) *chi2 = h->hctl->chi2;
  return NSN_OK;
}

extern "C" int nsn_max_abs_b(nsn_problem* h, double* out) {
  if (!h || !out) return fail(NSN_EINVAL, "null argument");
  if (!h->linearized) return fail(NSN_ESTATE, "nsn_max_abs_b before nsn_linearize(with_hessian=1)");
  int rc = sync_ctl(h);
  if (rc) return rc;
  *out = h->hctl->maxb;
  return NSN_OK;
}

extern "C" int nsn_max_abs_h_diag(nsn_problem* h, double* out) {
  if (!h || !out) return fail(NSN_EINVAL, "null argument");
  if (!h->linearized) return fail(NSN_ESTATE, "nsn_max_abs_h_diag before nsn_linearize(with_hessian=1)");
  int rc = sync_ctl(h);
  if (rc) return rc;
  *out = h->hctl->maxdiag;
  return NSN_OK;
}

template <typename T>
static int export_hessian_t(nsn_problem* h, double* U, double* V, double* W, int* Wi, int* Wj, double* bU,
                            double* bV) {
  const int NU = h->NU, NV = h->NV, nW = h->nW;
  std::vector<T> hu(21 * (size_t)NU), hv(6 * (size_t)NV), hw(18 * (size_t)nW), hbu(6 * (size_t)NU),
      hbv(3 * (size_t)NV);
  HIP_TRY(hipStreamSynchronize(h->st));
  if (U) HIP_TRY(hipMemcpy(hu.data(), h->U, hu.size() * sizeof(T), hipMemcpyDeviceToHost));
  if (V) HIP_TRY(hipMemcpy(hv.data(), h->V, hv.size() * sizeof(T), hipMemcpyDeviceToHost));
  if (W) HIP_TRY(hipMemcpy(hw.data(), h->W, hw.size() * sizeof(T), hipMemcpyDeviceToHost));
  if (bU) HIP_TRY(hipMemcpy(hbu.data(), h->bU, hbu.size() * sizeof(T), hipMemcpyDeviceToHost));
  if (bV) HIP_TRY(hipMemcpy(hbv.data(), h->bV, hbv.size() * sizeof(T), hipMemcpyDeviceToHost));
  // per-camera arrays in interface order (camera order)
  if (U)
    for (int u = 0; u < NU; u++)
      for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
          const int i = std::min(a, b), j = std::max(a, b);
          U[36 * (size_t)h->var_cam_api[u] + 6 * a + b] = (double)hu[21 * (size_t)u + i * 6 - i * (i - 1) / 2 + (j - i)];
        }
  if (V)
    for (int q = 0; q < NV; q++) {
      const T* s = &hv[6 * (size_t)q];
      const double full[9] = {(double)s[0], (double)s[1], (double)s[2], (double)s[1], (double)s[3],
                              (double)s[4], (double)s[2], (double)s[4], (double)s[5]};
      std::memcpy(V + 9 * (size_t)q, full, sizeof(full));
    }
  for (int r = 0; r < nW; r++) {
    const int k = h->w_ref[r];
    if (W)
      for (int i = 0; i < 18; i++) W[18 * (size_t)r + i] = (double)hw[18 * (size_t)k + i];
    if (Wi) Wi[r] = h->var_cam_api[h->w_cam[k]];
    if (Wj) Wj[r] = h->w_pt[k];
  }
  if (bU)
    for (int u = 0; u < NU; u++)
      for (int i = 0; i < 6; i++) bU[6 * (size_t)h->var_cam_api[u] + i] = (double)hbu[6 * (size_t)u + i];
  if (bV) for (size_t i = 0; i < hbv.size(); i++) bV[i] = (double)hbv[i];
  return NSN_OK;
}

extern "C" int nsn_export_hessian(nsn_problem* h, double* U, double* V, double* W, int* Wi, int* Wj, double* bU,
                                  double* bV, double* chi2) {
  if (!h) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  int rc = h->prec == NSN_MIXED ? export_hessian_t<float>(h, U, V, W, Wi, Wj, bU, bV)
                                : export_hessian_t<double>(h, U, V, W, Wi, Wj, bU, bV);
  if (rc) return rc;
  if (chi2) {
    if ((rc = sync_ctl(h))) return rc;
    *chi2 = h->hctl->chi2_lin;
  }
  return NSN_OK;
}

template <typename T>
static int import_hessian_t(nsn_problem* h, const double* U, const double* V, const double* W, const double* bU,
                            const double* bV) {
  const int NU = h->NU, NV = h->NV, nW = h->nW;
  std::vector<T> hu(21 * (size_t)NU), hv(6 * (size_t)NV), hw(18 * (size_t)nW), hbu(6 * (size_t)NU),
      hbv(3 * (size_t)NV);
  // per-camera arrays in interface order (camera order)
  for (int u = 0; u < NU; u++) {
    const size_t a = h->var_cam_api[u];
    for (int i = 0; i < 6; i++)
      for (int j = i; j < 6; j++) hu[21 * (size_t)u + i * 6 - i * (i - 1) / 2 + (j - i)] = (T)U[36 * a + 6 * i + j];
  }
  for (int q = 0; q < NV; q++) {
    const double* s = V + 9 * (size_t)q;
    const double up[6] = {s[0], s[1], s[2], s[4], s[5], s[8]};
    for (int i = 0; i < 6; i++) hv[6 * (size_t)q + i] = (T)up[i];
  }
  for (int r = 0; r < nW; r++) {
    const int k = h->w_ref[r];
    for (int i = 0; i < 18; i++) hw[18 * (size_t)k + i] = (T)W[18 * (size_t)r + i];
  }
  for (int u = 0; u < NU; u++)
    for (int i = 0; i < 6; i++) hbu[6 * (size_t)u + i] = (T)bU[6 * (size_t)h->var_cam_api[u] + i];
  for (size_t i = 0; i < hbv.size(); i++) hbv[i] = (T)bV[i];
  int rc;
  if ((rc = upload((T*)h->U, hu)) || (rc = upload((T*)h->V, hv)) || (rc = upload((T*)h->W, hw)) ||
      (rc = upload((T*)h->bU, hbu)) || (rc = upload((T*)h->bV, hbv)))
    return rc;
  return NSN_OK;
}

template <typename T>
static int import_uoff_t(nsn_problem* h, const double* Uoff) {
  std::vector<T> hu(36 * (size_t)h->n_uoff);
  for (size_t i = 0; i < hu.size(); i++) hu[i] = (T)Uoff[i];
  HIP_TRY(hipMemcpy(h->Uoff, hu.data(), hu.size() * sizeof(T), hipMemcpyHostToDevice));
  return NSN_OK;
}

extern "C" int nsn_import_hessian_blocks(nsn_problem* h, const double* U, const double* Uoff, const double* V,
                                         const double* W, const double* bU, const double* bV, double chi2) {
  TraceRange trace_range("nsn_import_hessian_blocks");
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (h->n_uoff > 0) {
    if (!Uoff) return fail(NSN_EINVAL, "nsn_import_hessian_blocks: the problem has off-diagonal U blocks, Uoff is NULL");
    HIP_TRY(hipSetDevice(h->device));
    const int rc = h->prec == NSN_MIXED ? import_uoff_t<float>(h, Uoff) : import_uoff_t<double>(h, Uoff);
    if (rc) return rc;
  }
  return nsn_import_hessian(h, U, V, W, bU, bV, chi2);
}

extern "C" int nsn_import_hessian(nsn_problem* h, const double* U, const double* V, const double* W, const double* bU,
                                  const double* bV, double chi2) {
  TraceRange trace_range("nsn_import_hessian");
  if (!h || !U || !V || !W || !bU || !bV) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  int rc = h->prec == NSN_MIXED ? import_hessian_t<float>(h, U, V, W, bU, bV)
                                : import_hessian_t<double>(h, U, V, W, bU, bV);
  if (rc) return rc;
  h->imported = true;
  if ((rc = sync_ctl(h))) return rc;
  double mb = 0, md = 0;
  for (size_t i = 0; i < 6 * (size_t)h->NU; i++) mb = std::max(mb, std::fabs(bU[i]));
  for (size_t i = 0; i < 3 * (size_t)h->NV; i++) mb = std::max(mb, std::fabs(bV[i]));
  for (int u = 0; u < h->NU; u++)
    for (int i = 0; i < 6; i++) md = std::max(md, std::fabs(U[36 * (size_t)u + 7 * i]));
  for (int q = 0; q < h->NV; q++)
    for (int i = 0; i < 3; i++) md = std::max(md, std::fabs(V[9 * (size_t)q + 4 * i]));
  h->hctl->chi2 = chi2;
  h->hctl->chi2_lin = chi2;
  h->hctl->maxb = mb;
  h->hctl->maxdiag = md;
  if ((rc = push_ctl(h))) return rc;
  h->linearized = true;
  return NSN_OK;
}

extern "C" int nsn_schur_init(nsn_problem* h) {
  if (!h) return fail(NSN_EINVAL, "null argument");
  // the symbolic work (pair lists, tiles) is done by nsn_problem_create, so
  // init only checks the call order (SolverCholeskySchur.hpp:44-62)
  if (!h->linearized) return fail(NSN_ESTATE, "nsn_schur_init before nsn_linearize(with_hessian=1)");
  return NSN_OK;
}

extern "C" int nsn_schur_step(nsn_problem* h, double rel, double abs_, double* dxu, double* dxv, double* sq_norm) {
  TraceRange trace_range("nsn_schur_step");
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (!h->linearized) return fail(NSN_ESTATE, "nsn_schur_step before nsn_linearize(with_hessian=1)");
  HIP_TRY(hipSetDevice(h->device));
  int rc = reset_ctl_primitive(h);
  if (rc) return rc;
  launch_set_damping(h->ctl, rel, abs_, h->st);
  if (h->prec == NSN_MIXED) seq_schur<float>(h, GATE_ALWAYS, 0, true);
  else seq_schur<double>(h, GATE_ALWAYS, 0, true);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_ctl(h))) return rc;
  if ((rc = check_device_errors(h))) return rc;
  h->stepped = true;
  h->last_abs = abs_;
  const bool ok = h->hctl->solve_ok != 0;
  if (dxu) {
    if (ok) {
      std::vector<double> x(6 * (size_t)h->NU);
      HIP_TRY(hipMemcpy(x.data(), h->xU, sizeof(double) * x.size(), hipMemcpyDeviceToHost));
      for (int u = 0; u < h->NU; u++)  // interface order
        std::memcpy(dxu + 6 * (size_t)h->var_cam_api[u], x.data() + 6 * (size_t)u, 6 * sizeof(double));
    } else {
      std::memset(dxu, 0, sizeof(double) * 6 * (size_t)h->NU);
    }
  }
  if (dxv) {
    if (ok) HIP_TRY(hipMemcpy(dxv, h->xV, sizeof(double) * 3 * (size_t)h->NV, hipMemcpyDeviceToHost));
    else std::memset(dxv, 0, sizeof(double) * 3 * (size_t)h->NV);
  }
  if (sq_norm) *sq_norm = ok ? h->hctl->incsq : 0.0;
  if (!ok) {
    launch_zero<double>(h->xU, 6LL * h->NU, h->st);
    launch_zero<double>(h->xV, 3LL * h->NV, h->st);
    HIP_TRY(hipStreamSynchronize(h->st));
    return NSN_NOT_PD;
  }
  return NSN_OK;
}

extern "C" int nsn_phase_times(nsn_problem* h, double* ms, int n) {
  if (!h || !ms || n < 5) return fail(NSN_EINVAL, "nsn_phase_times needs 5 slots");
  if (!h->stepped) return fail(NSN_ESTATE, "no Schur step has run");
  HIP_TRY(hipEventSynchronize(h->ev[6]));
  float t = 0;
  const int pairs[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {5, 6}};
  for (int i = 0; i < std::min(n, 6); i++) {
    HIP_TRY(hipEventElapsedTime(&t, h->ev[pairs[i][0]], h->ev[pairs[i][1]]));
    ms[i] = t;
  }
  return NSN_OK;
}

extern "C" int nsn_set_priors(nsn_problem* h, int n_cam, const int* cam_ids, const double* cam_p0,
                              const double* cam_w, int n_pt, const int* pt_ids, const double* pt_p0,
                              const double* pt_w) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (n_cam < 0 || n_pt < 0) return fail(NSN_EINVAL, "nsn_set_priors: negative count");
  if ((n_cam > 0 && (!cam_ids || !cam_p0 || !cam_w)) || (n_pt > 0 && (!pt_ids || !pt_p0 || !pt_w)))
    return fail(NSN_EINVAL, "nsn_set_priors: missing arrays");
  std::vector<int> cidx(h->C, -1), pidx(h->P, -1);
  for (int k = 0; k < n_cam; k++) {
    const int c = cam_ids[k];
    if (c < 0 || c >= h->C) return fail(NSN_EINVAL, "nsn_set_priors: camera prior %d on camera %d out of range", k, c);
    if (cidx[c] >= 0) return fail(NSN_EINVAL, "nsn_set_priors: two priors on camera %d", c);
    for (int i = 0; i < 6; i++)
      if (!std::isfinite(cam_p0[6 * (size_t)k + i]) || !std::isfinite(cam_w[6 * (size_t)k + i]))
        return fail(NSN_EINVAL, "nsn_set_priors: camera prior %d is not finite", k);
    cidx[c] = k;
  }
  for (int k = 0; k < n_pt; k++) {
    const int q = pt_ids[k];
    if (q < 0 || q >= h->P) return fail(NSN_EINVAL, "nsn_set_priors: point prior %d on point %d out of range", k, q);
    if (pidx[q] >= 0) return fail(NSN_EINVAL, "nsn_set_priors: two priors on point %d", q);
    for (int i = 0; i < 3; i++)
      if (!std::isfinite(pt_p0[3 * (size_t)k + i]) || !std::isfinite(pt_w[3 * (size_t)k + i]))
        return fail(NSN_EINVAL, "nsn_set_priors: point prior %d is not finite", k);
    pidx[q] = k;
  }
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->st));  // the old prior arrays may be in use
  std::unique_ptr<DevBuf> m(new DevBuf());
  BAView& v = h->view;
  v.cam_prior = nullptr; v.cam_prior_p = nullptr; v.cam_prior_w = nullptr;
  v.pt_prior = nullptr; v.pt_prior_p = nullptr; v.pt_prior_w = nullptr;
  int rc;
  if (n_cam > 0) {
    int* di = m->alloc<int>(h->C);
    double* dp = m->alloc<double>(6 * (size_t)n_cam);
    double* dw = m->alloc<double>(6 * (size_t)n_cam);
    if (m->failed) return fail(NSN_ENOMEM, "nsn_set_priors: device allocation failed");
    if ((rc = upload(di, cidx)) || (rc = upload(dp, std::vector<double>(cam_p0, cam_p0 + 6 * (size_t)n_cam))) ||
        (rc = upload(dw, std::vector<double>(cam_w, cam_w + 6 * (size_t)n_cam))))
      return rc;
    v.cam_prior = di; v.cam_prior_p = dp; v.cam_prior_w = dw;
  }
  if (n_pt > 0) {
    int* di = m->alloc<int>(h->P);
    double* dp = m->alloc<double>(3 * (size_t)n_pt);
    double* dw = m->alloc<double>(3 * (size_t)n_pt);
    if (m->failed) return fail(NSN_ENOMEM, "nsn_set_priors: device allocation failed");
    if ((rc = upload(di, pidx)) || (rc = upload(dp, std::vector<double>(pt_p0, pt_p0 + 3 * (size_t)n_pt))) ||
        (rc = upload(dw, std::vector<double>(pt_w, pt_w + 3 * (size_t)n_pt))))
      return rc;
    v.pt_prior = di; v.pt_prior_p = dp; v.pt_prior_w = dw;
  }
  h->pri_mem = std::move(m);
  h->drop_graphs();  // the captured iterations hold the old view
  h->linearized = false;
  h->stepped = false;
  return NSN_OK;
}

extern "C" int nsn_set_refinement(nsn_problem* h, int steps) {
  return nsn_set_refinement_tol(h, steps, 0.0);
}

extern "C" int nsn_set_refinement_tol(nsn_problem* h, int max_steps, double tol) {
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (max_steps < 0 || max_steps > 8)
    return fail(NSN_EINVAL, "%s: steps must be in 0..8, got %d", __func__, max_steps);
  if (!(tol >= 0.0) || !std::isfinite(tol)) return fail(NSN_EINVAL, "%s: tol must be finite and >= 0", __func__);
  // the captured iterations hold the old sequence and tolerance
  if (max_steps != h->ir_steps || tol != h->ir_tol) h->drop_graphs();
  h->ir_steps = max_steps;
  h->ir_tol = tol;
  return NSN_OK;
}

extern "C" int nsn_refinement_count(nsn_problem* h, int* steps) {
  if (!h || !steps) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpyAsync(h->hctl, h->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, h->st));
  HIP_TRY(hipStreamSynchronize(h->st));
  *steps = h->hctl->ir_count;
  return NSN_OK;
}

extern "C" int nsn_oplus(nsn_problem* h) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (!h->stepped) return fail(NSN_ESTATE, "nsn_oplus before nsn_schur_step");
  HIP_TRY(hipSetDevice(h->device));
  launch_oplus(h->view, h->xU, h->xV, h->ctl, GATE_ALWAYS, h->st);
  HIP_TRY(hipStreamSynchronize(h->st));
  return NSN_OK;
}

extern "C" int nsn_rho(nsn_problem* h, double mu, double old_chi2, double* rho) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h || !rho) return fail(NSN_EINVAL, "null argument");
  if (!h->stepped) return fail(NSN_ESTATE, "nsn_rho before nsn_schur_step");
  int rc = sync_ctl(h);
  if (rc) return rc;
  double den = h->hctl->den;
  if (mu != h->last_abs) {
    // x^T (mu x - b) for a mu other than the step's damping
    const int nu = 6 * h->NU, nv = 3 * h->NV;
    std::vector<double> x(nu + nv), b(nu + nv);
    HIP_TRY(hipMemcpy(x.data(), h->xU, sizeof(double) * nu, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(x.data() + nu, h->xV, sizeof(double) * nv, hipMemcpyDeviceToHost));
    if (h->prec == NSN_MIXED && h->ir_steps > 0) {
      HIP_TRY(hipMemcpy(b.data(), h->bU64, sizeof(double) * nu, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(b.data() + nu, h->bV64, sizeof(double) * nv, hipMemcpyDeviceToHost));
    } else if (h->prec == NSN_MIXED) {
      std::vector<float> bf(nu + nv);
      HIP_TRY(hipMemcpy(bf.data(), h->bU, sizeof(float) * nu, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(bf.data() + nu, h->bV, sizeof(float) * nv, hipMemcpyDeviceToHost));
      for (int i = 0; i < nu + nv; i++) b[i] = bf[i];
    } else {
      HIP_TRY(hipMemcpy(b.data(), h->bU, sizeof(double) * nu, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(b.data() + nu, h->bV, sizeof(double) * nv, hipMemcpyDeviceToHost));
    }
    double du = 0, dv = 0;
    for (int i = 0; i < nu; i++) du += x[i] * (mu * x[i] - b[i]);
    for (int i = nu; i < nu + nv; i++) dv += x[i] * (mu * x[i] - b[i]);
    den = du + dv;
  }
  *rho = (old_chi2 - h->hctl->chi2) / den;
  return NSN_OK;
}

extern "C" int nsn_set_observations(nsn_problem* h, const double* obs, const double* weight) {
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  const int E = h->E, Ec = h->Ec;
  if (obs) {
    std::vector<double> po(2 * (size_t)E), co(2 * (size_t)Ec);
    for (int s = 0; s < E; s++) {
      po[2 * (size_t)s] = obs[2 * (size_t)h->pe_order[s]];
      po[2 * (size_t)s + 1] = obs[2 * (size_t)h->pe_order[s] + 1];
    }
    for (int s = 0; s < Ec; s++) {
      co[2 * (size_t)s] = obs[2 * (size_t)h->ce_order[s]];
      co[2 * (size_t)s + 1] = obs[2 * (size_t)h->ce_order[s] + 1];
    }
    int rc;
    if ((rc = upload(h->d_pe_obs, po)) || (rc = upload(h->d_ce_obs, co))) return rc;
  }
  if (weight) {
    std::vector<double> pw(E), cw(Ec);
    for (int s = 0; s < E; s++) pw[s] = weight[h->pe_order[s]];
    for (int s = 0; s < Ec; s++) cw[s] = weight[h->ce_order[s]];
    int rc;
    if ((rc = upload(h->d_pe_w, pw)) || (rc = upload(h->d_ce_w, cw))) return rc;
  }
  h->linearized = false;
  h->stepped = false;
  return NSN_OK;
}

namespace {

// work vectors of the adjoint: seven of the camera order (6 NU), seven of the
// point order (3 NV), the per-edge outputs and eps
struct AdjBufs {
  double *gU, *lamU, *tU, *bpU, *bmU, *dU, *zeroU;
  double *gV, *lamV, *tV, *bpV, *bmV, *dV, *zeroV;
  double *dobs, *dw, *eps;
};

// the adjoint's work vectors of a handle, allocated on its first nsn_adjoint
int alloc_adjoint(nsn_problem* h, AdjBufs& b) {
  const size_t nu = 6 * (size_t)h->NU, nv = 3 * (size_t)h->NV;
  double** us[] = {&h->gU, &h->lamU, &h->tU, &h->bpU, &h->bmU, &h->dU, &h->zeroU};
  double** vs[] = {&h->gV, &h->lamV, &h->tV, &h->bpV, &h->bmV, &h->dV, &h->zeroV};
  if (!h->gU) {
    DevBuf& m = h->adj_mem;
    for (double** p : us) *p = m.alloc<double>(nu);
    for (double** p : vs) *p = m.alloc<double>(nv);
    h->dobs = m.alloc<double>(2 * (size_t)h->E);
    h->dw = m.alloc<double>(h->E);
    h->adj_eps = m.alloc<double>(1);
    if (m.failed) {  // nothing half-allocated stays behind for the next call
      m.reset();
      for (double** p : us) *p = nullptr;
      for (double** p : vs) *p = nullptr;
      h->dobs = h->dw = h->adj_eps = nullptr;
      return fail(NSN_ENOMEM, "nsn_adjoint: device allocation failed");
    }
  }
  b = AdjBufs{h->gU, h->lamU, h->tU, h->bpU, h->bmU, h->dU, h->zeroU, h->gV, h->lamV, h->tV, h->bpV,
              h->bmV, h->dV, h->zeroV, h->dobs, h->dw, h->adj_eps};
  return NSN_OK;
}

// gradient b of the cost at the current state perturbed by oplus(t),
// FP64, into (bu, bv); the state is restored afterwards
void grad_at_offset(nsn_problem* h, const AdjBufs& b, const double* tu, const double* tv, double* bu, double* bv) {
  hipStream_t st = h->st;
  launch_copy_state(h->view, h->cams_bak, h->pts_bak, h->cams, h->pts, h->ctl, GATE_ALWAYS, st);
  launch_oplus(h->view, tu, tv, h->ctl, GATE_ALWAYS, st);
  launch_ir_residual(h->view, b.zeroU, b.zeroV, h->rU, h->rV, bu, bv, h->ctl, GATE_ALWAYS, st);
  launch_copy_state(h->view, h->cams, h->pts, h->cams_bak, h->pts_bak, h->ctl, GATE_ALWAYS, st);
}

template <typename T>
int adjoint_t(nsn_problem* h, const AdjBufs& b, const double* g_cams, const double* g_pts, int steps,
              double* dL_dobs, double* dL_dw) {
  hipStream_t st = h->st;
  const int NU = h->NU, NV = h->NV;
  const long long nu = 6LL * NU, nv = 3LL * NV;
  int rc;
  if ((rc = reset_ctl_primitive(h))) return rc;
  // the Gauss-Newton Hessian at x* and its Schur factorisation (undamped)
  seq_lin_full<T>(h, GATE_ALWAYS);
  launch_set_damping(h->ctl, 0.0, 0.0, st);
  seq_schur<T>(h, GATE_ALWAYS, 0, false);
  if ((rc = sync_ctl(h))) return rc;
  if ((rc = check_device_errors(h))) return rc;
  if (!h->hctl->solve_ok)
    return fail(NSN_NOT_PD, "nsn_adjoint: the Gauss-Newton Hessian at this state is not positive definite");
  // loss gradient in variable order
  std::vector<double> gu(nu), gv(nv);
  for (int u = 0; u < NU; u++)
    for (int i = 0; i < 6; i++) gu[6 * (size_t)u + i] = g_cams ? g_cams[6 * (size_t)h->var_cam_user[u] + i] : 0.0;
  for (int q = 0; q < NV; q++)
    for (int i = 0; i < 3; i++) gv[3 * (size_t)q + i] = g_pts ? g_pts[3 * (size_t)h->var_pt_user[q] + i] : 0.0;
  if ((rc = upload(b.gU, gu)) || (rc = upload(b.gV, gv))) return rc;
  // lambda = H^-1 g with the Schur factorisation
  seq_schur_solve<T>(h, GATE_ALWAYS, b.gU, b.gV, b.lamU, b.lamV, 0, nullptr);
  // refinement against the exact Hessian: lambda += H^-1 (g - B lambda), with
  // B lambda = (b(x + eps lambda) - b(x - eps lambda)) / (2 eps), the central
  // difference of the FP64 gradient along lambda
  // (all on the device: eps = 1e-6 / max |lambda| lives in adj_eps)
  for (int k = 0; k < steps; k++) {
    launch_adj_eps(b.lamU, nu, b.lamV, nv, b.eps, st);
    launch_adj_offset(b.lamU, nu, b.lamV, nv, b.eps, 1.0, b.tU, b.tV, st);
    grad_at_offset(h, b, b.tU, b.tV, b.bpU, b.bpV);
    launch_adj_offset(b.lamU, nu, b.lamV, nv, b.eps, -1.0, b.tU, b.tV, st);
    grad_at_offset(h, b, b.tU, b.tV, b.bmU, b.bmV);
    launch_adj_resid(b.gU, b.bmU, b.bpU, nu, b.gV, b.bmV, b.bpV, nv, b.eps, st);
    seq_schur_solve<T>(h, GATE_ALWAYS, b.bpU, b.bpV, b.dU, b.dV, 0, nullptr);
    launch_axpby(nu, 1.0, b.lamU, 1.0, b.dU, b.lamU, st);
    launch_axpby(nv, 1.0, b.lamV, 1.0, b.dV, b.lamV, st);
  }
  launch_adjoint_edges(h->view, b.lamU, b.lamV, b.dobs, b.dw, st);
  HIP_TRY(hipGetLastError());
  if ((rc = sync_ctl(h))) return rc;
  if ((rc = check_device_errors(h))) return rc;
  if (dL_dobs) HIP_TRY(hipMemcpy(dL_dobs, b.dobs, sizeof(double) * 2 * (size_t)h->E, hipMemcpyDeviceToHost));
  if (dL_dw) HIP_TRY(hipMemcpy(dL_dw, b.dw, sizeof(double) * (size_t)h->E, hipMemcpyDeviceToHost));
  // the factorisation and increment buffers now hold the adjoint's values
  h->stepped = false;
  return NSN_OK;
}

}  // namespace

extern "C" int nsn_adjoint(nsn_problem* h, const double* g_cams, const double* g_pts, int hessian_steps,
                           double* dL_dobs, double* dL_dw) {
  TraceRange trace_range("nsn_adjoint");
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (hessian_steps < 0 || hessian_steps > 16)
    return fail(NSN_EINVAL, "nsn_adjoint: hessian_steps must be in 0..16, got %d", hessian_steps);
  HIP_TRY(hipSetDevice(h->device));
  AdjBufs b;
  int rc = alloc_adjoint(h, b);
  if (rc) return rc;
  if (h->zero_fresh) {
    HIP_TRY(hipMemset(h->zeroU, 0, sizeof(double) * std::max<size_t>(6 * (size_t)h->NU, 1)));
    HIP_TRY(hipMemset(h->zeroV, 0, sizeof(double) * std::max<size_t>(3 * (size_t)h->NV, 1)));
    h->zero_fresh = false;
  }
  return h->prec == NSN_MIXED ? adjoint_t<float>(h, b, g_cams, g_pts, hessian_steps, dL_dobs, dL_dw)
                              : adjoint_t<double>(h, b, g_cams, g_pts, hessian_steps, dL_dobs, dL_dw);
}

// ---- the adjoint without allocation ---------------------------------------
// For callers that cannot allocate device memory at the point of the call
// (graph capture, serving loops): the adjoint's work vectors live in this
// file-scope device array instead of the handle's lazily allocated buffers.
// Each device holds its own copy for the life of the process; the host reads
// its address once per device with hipGetSymbolAddress and passes that
// address, never the array's name, to the kernels.  Calls on one device share
// the array, so calls that overlap on one device race.
constexpr size_t kAdjScratchDoubles = size_t(4) << 20;  // 32 MiB
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kAdjScratchDoubles];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device `dev`
int scratch_address(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(NSN_EINVAL, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NSN_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return NSN_OK;
}

// doubles of the adjoint's work vectors, each rounded up to 256 bytes
size_t adj_scratch_need(const nsn_problem* h) {
  auto r = [](size_t n) { return (std::max<size_t>(n, 1) + 31) / 32 * 32; };
  return 7 * r(6 * (size_t)h->NU) + 7 * r(3 * (size_t)h->NV) + r(2 * (size_t)h->E) + r(h->E) + r(1);
}

}  // namespace

extern "C" int nsn_adjoint_noalloc(nsn_problem* h, const double* g_cams, const double* g_pts, int hessian_steps,
                                   double* dL_dobs, double* dL_dw) {
  TraceRange trace_range("nsn_adjoint_noalloc");
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h) return fail(NSN_EINVAL, "null argument");
  if (hessian_steps < 0 || hessian_steps > 16)
    return fail(NSN_EINVAL, "nsn_adjoint_noalloc: hessian_steps must be in 0..16, got %d", hessian_steps);
  const size_t need = adj_scratch_need(h);
  if (need > kAdjScratchDoubles)
    return fail(NSN_EINVAL,
                "nsn_adjoint_noalloc: the problem needs %zu doubles of work space (7 x 6 NU + 7 x 3 NV + 3 E), "
                "more than the %zu of the static scratch array; use nsn_adjoint",
                need, kAdjScratchDoubles);
  HIP_TRY(hipSetDevice(h->device));
  double* base = nullptr;
  int rc = scratch_address(h->device, &base);
  if (rc) return rc;
  auto r = [](size_t n) { return (std::max<size_t>(n, 1) + 31) / 32 * 32; };
  const size_t nu = 6 * (size_t)h->NU, nv = 3 * (size_t)h->NV;
  double* q = base;
  auto take = [&](size_t n) { double* p = q; q += r(n); return p; };
  AdjBufs b;
  b.gU = take(nu); b.lamU = take(nu); b.tU = take(nu); b.bpU = take(nu); b.bmU = take(nu); b.dU = take(nu);
  b.zeroU = take(nu);
  b.gV = take(nv); b.lamV = take(nv); b.tV = take(nv); b.bpV = take(nv); b.bmV = take(nv); b.dV = take(nv);
  b.zeroV = take(nv);
  b.dobs = take(2 * (size_t)h->E); b.dw = take(h->E); b.eps = take(1);
  // the array keeps the previous call's contents: the zero vectors are
  // cleared by a fill kernel on every call
  launch_zero<double>(b.zeroU, (long long)nu, h->st);
  launch_zero<double>(b.zeroV, (long long)nv, h->st);
  return h->prec == NSN_MIXED ? adjoint_t<float>(h, b, g_cams, g_pts, hessian_steps, dL_dobs, dL_dw)
                              : adjoint_t<double>(h, b, g_cams, g_pts, hessian_steps, dL_dobs, dL_dw);
}

// captures one launch sequence on the problem's stream into an executable
// graph; null when capture is switched off (NSN_NO_GRAPH) or fails, and the
// caller then launches the sequence directly
static hipGraphExec_t capture(nsn_problem* h, void (*seq)(nsn_problem*)) {
  static const bool no_graph = std::getenv("NSN_NO_GRAPH") != nullptr;
  if (no_graph) return nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    seq(h);
    if (hipStreamEndCapture(h->st, &g) == hipSuccess && g) {
      if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipGetLastError();
  return exec;
}

template <typename T>
static int run_lm(nsn_problem* h, const nsn_lm_settings* s, nsn_lm_result* res, nsn_log_entry* log, int log_cap) {
  h->imported = false;  // the drivers linearise from the edges
  DevCtl c = make_ctl(h->log_cap);
  c.mode = 1;
  c.need_lin = 1;
  c.maxNumIt = s->maxNumIt; c.maxNumSubIt = s->maxNumSubIt; c.minNumIt = s->minNumIt;
  c.epsBVector = s->epsBVector; c.epsIncSquare = s->epsIncSquare; c.epsChi2 = s->epsChi2;
  c.tau = s->tau; c.penaltyV = s->penaltyV; c.initV = s->initV;
  c.maxScaleFactor = s->maxScaleFactor; c.minScaleFactor = s->minScaleFactor;
  *h->hctl = c;
  int rc = push_ctl(h);
  if (rc) return rc;
  if (!h->lm_graph_tried) {
    h->lm_graph_tried = true;
    h->lm_graph = capture(h, seq_lm_subiter<T>);
  }
  const long long max_sub = (long long)std::max(1, s->maxNumIt) * std::max(1, s->maxNumSubIt) + 4;
  auto t0 = std::chrono::steady_clock::now();
  // sub-iterations are queued in batches of 1, 2, 4 and then 8 between host
  // synchronisations: every kernel is gated on the device's done flag, so the
  // replays queued after the solve ends do nothing, and a short solve wastes
  // few of them
  long long queued = 0;
  for (int batch = 1; queued < max_sub; batch = std::min(2 * batch, kLmReplaysPerSync)) {
    for (int b = 0; b < batch && queued < max_sub; b++, queued++) {
      if (h->lm_graph) HIP_TRY(hipGraphLaunch(h->lm_graph, h->st));
      else seq_lm_subiter<T>(h);
    }
    if ((rc = sync_ctl(h))) return rc;
    if (h->hctl->done || h->hctl->wait_timeout) break;
  }
  auto t1 = std::chrono::steady_clock::now();
  HIP_TRY(hipGetLastError());
  if ((rc = check_device_errors(h))) return rc;
  const DevCtl& r = *h->hctl;
  res->reason = r.done ? r.reason : NSN_MAX_ITERATION_REACHED;
  res->iterations = r.iter;
  res->sub_iterations = r.total_solves;
  res->n_log = std::min(r.n_log, h->log_cap);
  res->final_chi2 = r.chi2;
  res->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (log && log_cap > 0) {
    const int n = std::min(res->n_log, log_cap);
    std::vector<DevLog> dl(n);
    if (n > 0) HIP_TRY(hipMemcpy(dl.data(), h->log, sizeof(DevLog) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { log[i].it = dl[i].it; log[i].subit = dl[i].subit; log[i].chi2 = dl[i].chi2; log[i].mu = dl[i].mu; }
  }
  h->linearized = true;
  h->stepped = true;
  h->last_abs = r.abs_;
  return NSN_OK;
}

extern "C" int nsn_lm_solve(nsn_problem* h, const nsn_lm_settings* s, nsn_lm_result* res, nsn_log_entry* log,
                            int log_cap) {
  TraceRange trace_range("nsn_lm_solve");
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h || !s || !res) return fail(NSN_EINVAL, "null argument");
  if (s->maxNumSubIt < 1) return fail(NSN_EINVAL, "maxNumSubIt must be >= 1");
  HIP_TRY(hipSetDevice(h->device));
  return h->prec == NSN_MIXED ? run_lm<float>(h, s, res, log, log_cap) : run_lm<double>(h, s, res, log, log_cap);
}

template <typename T>
static int run_gn(nsn_problem* h, const nsn_gn_settings* s, nsn_lm_result* res, nsn_log_entry* log, int log_cap) {
  h->imported = false;  // the drivers linearise from the edges
  DevCtl c = make_ctl(h->log_cap);
  c.mode = 2;
  c.maxNumIt = s->maxNumIt; c.minNumIt = s->minNumIt;
  c.epsBVector = s->epsBVector; c.epsIncSquare = s->epsIncSquare; c.epsChi2 = s->epsChi2;
  c.gn_rel = s->relLambda; c.gn_abs = s->absLambda;
  *h->hctl = c;
  int rc = push_ctl(h);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  seq_lin_full<T>(h, GATE_ALWAYS);
  launch_gn_after_lin(h->ctl, h->log, h->st);
  // iterations queued in batches between synchronisations, as in run_lm
  const int max_it = std::max(1, s->maxNumIt) + 2;
  int queued = 0;
  for (int batch = 1; queued < max_it; batch = std::min(2 * batch, kLmReplaysPerSync)) {
    if ((rc = sync_ctl(h))) return rc;
    if (h->hctl->done || h->hctl->wait_timeout) break;
    for (int b = 0; b < batch && queued < max_it; b++, queued++) seq_gn_iter<T>(h);
  }
  if ((rc = sync_ctl(h))) return rc;
  auto t1 = std::chrono::steady_clock::now();
  HIP_TRY(hipGetLastError());
  if ((rc = check_device_errors(h))) return rc;
  const DevCtl& r = *h->hctl;
  res->reason = r.done ? r.reason : NSN_MAX_ITERATION_REACHED;
  res->iterations = r.iter;
  res->sub_iterations = r.total_solves;
  res->n_log = std::min(r.n_log, h->log_cap);
  res->final_chi2 = r.chi2;
  res->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (log && log_cap > 0) {
    const int n = std::min(res->n_log, log_cap);
    std::vector<DevLog> dl(n);
    if (n > 0) HIP_TRY(hipMemcpy(dl.data(), h->log, sizeof(DevLog) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { log[i].it = dl[i].it; log[i].subit = dl[i].subit; log[i].chi2 = dl[i].chi2; log[i].mu = dl[i].mu; }
  }
  h->linearized = true;
  h->stepped = true;
  return NSN_OK;
}

extern "C" int nsn_gn_solve(nsn_problem* h, const nsn_gn_settings* s, nsn_lm_result* res, nsn_log_entry* log,
                            int log_cap) {
  TraceRange trace_range("nsn_gn_solve");
  if (h && h->blocks_only) return fail(NSN_ESTATE, "%s: a block problem has no edges or state (import its Hessian)", __func__);
  if (!h || !s || !res) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  return h->prec == NSN_MIXED ? run_gn<float>(h, s, res, log, log_cap) : run_gn<double>(h, s, res, log, log_cap);
}
