// Device data views and kernel launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "chol_tasks.h"

#include "internal.h"

namespace nsn {

// Kernel gates: each conditional kernel first evaluates its gate against the
// device control block and returns at once when it is closed, so a fixed
// kernel sequence (and a captured graph of it) runs the LM loop's branches.
enum Gate : int {
  GATE_ALWAYS = 0,
  GATE_NOT_DONE = 1,        // !done
  GATE_NEED_LIN = 2,        // !done && need_lin
  GATE_SOLVE_OK = 3,        // !done && solve_ok
  GATE_ROLLBACK = 4,        // rollback_pending
  GATE_IR = 5,              // !done && ir_more (a refinement step goes on)
  GATE_IR_ANY = 6,          // ir_more
};

// Read-only structure and state pointers of one BA problem on the device.
struct BAView {
  int C, P, E, NU, NV, nW;
  int w_dup;                // some W block has more than one edge (repeated observations)
  double f, cx, cy, huber;
  const int* cam_var;       // [C]  user camera -> variable index or -1
  const int* pt_var;        // [P]  user point  -> variable index or -1
  const int* var_cam_user;  // [NU]
  const int* var_pt_user;   // [NV]
  // edges in point-major order (user point, then camera)
  const int* pt_ptr;        // [P+1]
  const int* pe_cam;        // [E]
  const double* pe_obs;     // [2E]
  const double* pe_w;       // [E]
  const int* pe_wblk;       // [E] W block or -1
  const int* pe_edge;       // [E] user index of the edge
  // edges of variable cameras in camera-major order
  const int* cam_ptr;       // [NU+1]
  const int* ce_pt;         // [Ec]
  const double* ce_obs;     // [2Ec]
  const double* ce_w;       // [Ec]
  // W blocks: point-major (blocks of variable point v are wptr[v]..wptr[v+1])
  const int* wptr;          // [NV+1]
  const int* w_cam;         // [nW] variable camera of the block
  const int* w_pt;          // [nW] variable point of the block
  const int* cw_ptr;        // [NU+1] camera-major list of W blocks
  const int* cw_idx;        // [nW]
  const int* cw_pt;         // [nW] variable point of the block cw_idx[i]
  double* cams;             // [12C] R (row-major) and t per camera
  double* pts;              // [3P]
  const double* intr;       // [3C] (f, k1, k2) per camera of the BAL model, or null: shared pinhole
  // unary priors (nsn_set_priors; null pointers: none): per user camera /
  // point the index of its prior or -1; camera priors omega0, t0 and 6
  // weights, point priors X0 and 3 weights
  const int* cam_prior;       // [C]
  const double* cam_prior_p;  // [6 ncp]
  const double* cam_prior_w;  // [6 ncp]
  const int* pt_prior;        // [P]
  const double* pt_prior_p;   // [3 npp]
  const double* pt_prior_w;   // [3 npp]
};

// Tiled lower storage of the reduced camera system S (plus the augmented
// right-hand-side row at index n) over an envelope: tile row i stores the
// tiles (i, j) for first[i] <= j <= i, each TB x TB values, column-major
// inside, packed row by row (slot of (i, j) = row_off[i] + j - first[i]).
// A Cholesky factor has no fill outside the envelope of its matrix, so the
// factor takes the same storage.  A dense S has first[i] = 0.  A tile row
// with no tile left of the diagonal (first[i] == i) starts an independent
// diagonal chain of the factorisation; every other tile row keeps the two
// sub-diagonal tiles of its chain (first[i] <= max(i - 2, start of its
// chain)).  The last tile row (the right-hand side) is dense.
struct TileView {
  int n;        // order of S (6 NU)
  int TB;       // tile size
  int NT;       // tiles per dimension (covers rows 0..n inclusive)
  int ntiles;   // stored tiles
  const int* first;    // [NT] first stored tile column of each tile row (device)
  const int* row_off;  // [NT] slot of tile (i, first[i]) (device)
  int nchain;          // chain workgroups of the factorisation
  const int* chain_start;  // [nchain + 1] first diagonal step of each (device); chain_start[nchain] = NT
  __device__ __forceinline__ long long slot(int i, int j) const { return (long long)row_off[i] + (j - first[i]); }
  __device__ __forceinline__ bool has(int i, int j) const { return j >= first[i]; }  // for i >= j
};

// the largest number of tile rows (the triangular solves keep the envelope
// of every tile row in LDS)
constexpr int kMaxTileRows = 4096;
// the most diagonal chain groups of the factorisation and the triangular
// solves (workgroups of their launches that run chains)
constexpr int kCholMaxChains = 16;

// the envelope of a tiled S on the host (the device copy of first / row_off
// is what TileView points to)
struct TileProfile {
  std::vector<int> first, row_off;
  std::vector<int> chain_start;  // chain groups of the factorisation (chol_chain_groups)
  int ntiles = 0;
  long long slot(int i, int j) const { return (long long)row_off[i] + (j - first[i]); }
  bool has(int i, int j) const { return j >= first[i]; }
};

// reduction scratch layout (doubles), one slot per block of the producing kernel
// lanes per point in the point-major linearisation pass (divides 64)
constexpr int kLinLanes = 4;

struct Scratch {
  double* part_chi2;    // [nblk_pts]
  double* part_xv;      // [2 * nblk_pts] (nblk_pts >= nblk_xv)  (incsq, den) of the points
  double* part_xu;      // [2 * nblk_cams] (incsq, den) of the cameras
  unsigned long long* maxes;  // [4] max|bU|, max|bV|, max|diag U|, max|diag V| as bit patterns
  // blocks of the point pass (kLinLanes P), xU stats (6 NU), back-substitution (kLinLanes NV)
  int nblk_pts, nblk_cams, nblk_xv;
};

// ---- BA kernels (ba_kernels.hip), T = double (FP64) or float (mixed) ----
template <typename T>
void launch_lin_points(const BAView& v, bool hess, T* V, T* bV, T* W, const Scratch& s,
                       DevCtl* ctl, int gate, hipStream_t st);
template <typename T>
void launch_lin_cams(const BAView& v, T* U, T* bU, const Scratch& s, DevCtl* ctl, int gate, hipStream_t st);
void launch_reset_lin(const Scratch& s, DevCtl* ctl, int gate, hipStream_t st);
void launch_finalize_lin(const Scratch& s, DevCtl* ctl, int gate, int full, hipStream_t st);
// Vinv = damped V^-1, Z = W L^-T (V = L L^T, so W_a V^-1 W_b^T = Z_a Z_b^T)
// and t = W V^-1 (-bV)
template <typename T>
void launch_vinv_y(const BAView& v, const T* V, const T* bV, const T* W, T* Vinv, T* Z, double* yV, DevCtl* ctl,
                   int gate, hipStream_t st);
template <typename T>
void launch_s_diag(const BAView& v, const T* U, const T* bU, const T* Z, const double* yV, T* Sd, T* bS,
                   DevCtl* ctl, int gate, hipStream_t st);
// one workgroup per part of a tile (part_tile, part_rect = local rows lo, hi,
// cols lo, hi, part_ptr / part_blk = the camera blocks inside)
template <typename T>
void launch_s_tiles(const TileView& tv, int nparts, const int* tile_order, const int* part_tile, const int* part_rect,
                    const int* part_ptr, const int* part_blk, const int* sb_i, const int* sb_j, const int* sb_ptr,
                    const int* sb_pairs, const T* Sd, const T* bS, const T* Z, T* tiles, DevCtl* ctl, int gate,
                    hipStream_t st);
// xV = V^-1 (rhsV - W^T xU) with rhsV = -bV when rhs_v is null; with `add`
// the result is added to xV; the increment statistics use bden (FP64
// gradient) when given, else bV; accU (nacc entries), when given, gets xU
// added (the refinement's camera update)
template <typename T>
void launch_backsub(const BAView& v, const T* W, const T* Vinv, const T* bV, const double* xU,
                    double* xV, const Scratch& s, DevCtl* ctl, int gate, hipStream_t st,
                    const double* rhs_v = nullptr, int add = 0, const double* bden = nullptr,
                    double* accU = nullptr, int nacc = 0);
// FP64 residual of the damped normal equations at the increment (xU, xV):
// rU/rV = -b - (H + D) x from the edges, and the FP64 gradient bU64/bV64
void launch_ir_residual(const BAView& v, const double* xU, const double* xV, double* rU, double* rV,
                        double* bU64, double* bV64, DevCtl* ctl, int gate, hipStream_t st);
// rS = rU - W V^-1 rV (Schur right-hand side of a correction); sets
// unset[0..3][0..nunset) to kUnsetBits (the solutions and destination sums
// of the two triangular solves that follow; nunset <= 256 NU)
template <typename T>
void launch_ir_reduce(const BAView& v, const T* W, const T* Vinv, const double* rU, const double* rV, double* rS,
                      double* const unset[4], int nunset, DevCtl* ctl, int gate, hipStream_t st);
// x += z
void launch_add(double* x, const double* z, long long n, DevCtl* ctl, int gate, hipStream_t st);
// out = a x + b y (out may alias x or y)
void launch_axpby(long long n, double a, const double* x, double b, const double* y, double* out, hipStream_t st);
// exact-Hessian refinement of the adjoint: eps = 1e-6 / max |lambda| (0 for a
// zero or non-finite lambda), t = sgn eps lambda, bp = g - (bp - bm) / (2 eps)
void launch_adj_eps(const double* lu, long long nu, const double* lv, long long nv, double* eps, hipStream_t st);
void launch_adj_offset(const double* lu, long long nu, const double* lv, long long nv, const double* eps, double sgn,
                       double* tu, double* tv, hipStream_t st);
void launch_adj_resid(const double* gu, const double* bmu, double* bpu, long long nu, const double* gv,
                      const double* bmv, double* bpv, long long nv, const double* eps, hipStream_t st);
// per-edge adjoint outputs at the current state from the multipliers
// (lamU, lamV): dL/d(obs) [2E] and dL/d(weight) [E], user edge order
void launch_adjoint_edges(const BAView& v, const double* lamU, const double* lamV, double* dobs, double* dw,
                          hipStream_t st);
template <typename T>
void launch_xu_stats(int NU, const double* xU, const T* bU, const Scratch& s, DevCtl* ctl, int gate,
                     hipStream_t st);
void launch_step_finish(const Scratch& s, DevCtl* ctl, int gate, int lm_checks, hipStream_t st);
void launch_oplus(const BAView& v, const double* xU, const double* xV, DevCtl* ctl, int gate,
                  hipStream_t st);
void launch_copy_state(const BAView& v, double* cams_dst, double* pts_dst, const double* cams_src,
                       const double* pts_src, DevCtl* ctl, int gate, hipStream_t st);
void launch_lm_after_lin(DevCtl* ctl, DevLog* log, hipStream_t st);
void launch_lm_control(DevCtl* ctl, DevLog* log, hipStream_t st);
void launch_gn_after_step(DevCtl* ctl, DevLog* log, hipStream_t st);
void launch_gn_after_lin(DevCtl* ctl, DevLog* log, hipStream_t st);
void launch_set_damping(DevCtl* ctl, double rel, double abs_, hipStream_t st);
void launch_begin_step(DevCtl* ctl, int gate, hipStream_t st);
// after a refinement residual: the step goes on (ctl->ir_more) while
// max |r| > tol * max |b| (tol <= 0: always); clears the maxima
void launch_ir_decide(DevCtl* ctl, int gate, double tol, hipStream_t st);
// adds the off-diagonal U blocks (i < j, 6x6 row-major, nb of them) to the
// assembled S tiles
template <typename T>
void launch_add_uoff(const TileView& tv, int nb, const int* ui, const int* uj, const T* Uoff, T* tiles, DevCtl* ctl,
                     int gate, hipStream_t st);
template <typename T>
void launch_zero(T* p, long long n, hipStream_t st);

// ---- dense tile Cholesky and triangular solve (chol_kernels.hip) --------
// bit pattern of a not yet published entry of a triangular solve's solution
// and destination sums: a signalling NaN, which arithmetic never produces
// (its NaN results are quiet)
constexpr unsigned long long kUnsetBits = 0x7FF4A5A5A5A5A5A5ull;
// zeroes the factorisation's flags and counters and sets unset0/unset1[0..nunset)
// to kUnsetBits (the solution and destination sums of the solve that follows)
template <typename T>
void launch_chol_reset(int* flags, int nflags, int* counters, double* unset0, double* unset1, int nunset,
                       DevCtl* ctl, int gate, hipStream_t st);
// factor in place; inv[j] receives L_jj^-1 (TB x TB, column-major); presub:
// NT scratch tiles; part: the CT_PART tasks' scratch tiles.  tasks: the
// worker task queue (kCholTaskInts per task, chol_queue); ready: ntiles
// flags, pre_sub / pre_diag: NT flags each, part_ready: one per part tile;
// counter zeroed
template <typename T>
void launch_chol(const TileView& tv, T* tiles, T* inv, T* presub, T* part, const int* tasks, int ntasks,
                 int* ready, int* pre_sub, int* pre_diag, int* part_ready, int* counter, DevCtl* ctl, int gate,
                 int max_wg, hipStream_t st);
// L^T sol = rhs (backward) or L sol = rhs (forward) over the first n rows;
// rhs == nullptr: row n of the factor.  far: n doubles of scratch; sol[0..n)
// and far[0..n) hold kUnsetBits before the launch (launch_chol_reset,
// launch_ir_reduce); rhs must not alias sol.
template <typename T>
void launch_trsv(const TileView& tv, const T* tiles, const T* inv, bool backward, const double* rhs, double* sol,
                 double* far, DevCtl* ctl, int gate, int max_wg, hipStream_t st);

}  // namespace nsn
