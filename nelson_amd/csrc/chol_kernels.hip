// Dense Cholesky of the reduced camera system on gfx950, as a persistent
// tile DAG in one launch, plus the backward triangular solve.
//
// The matrix is [S, .; bS^T, 1] padded with identity to NT*TB rows and
// stored as packed lower tiles (TileView).  Factorising it gives L with
// L y = bS in row n, i.e. the forward solve of SolverCholeskySchur.hpp:140
// comes out of the factorisation for free.  The reference factorises S with
// Eigen's pivoted LDL^T (SolverCholeskyDense.hpp:15-18); for the damped,
// positive definite S of an LM step the two give the same solution up to
// rounding.  A semidefinite S whose zero pivots have zero rows (a variable
// camera that nothing observes, undamped) is accepted by both with a zero
// solution entry there (potrf_panel); an indefinite S, which Eigen's LDL^T
// factorises and this LL^T cannot, is reported as a failure.
//
// Scheduling: workgroups take tiles from an atomic counter in column-major
// lower order (j, then i).  Tile (i, j) needs (i, k), (j, k) for k < j and
// (j, j); all of them come earlier in that order, and a workgroup finishes a
// tile before taking another, so every awaited tile is owned by a running
// workgroup: the DAG cannot deadlock whatever the residency.  Each wait is
// bounded; on expiry it raises ctl->wait_timeout and every later wait gives
// up at once, so the kernel always terminates.
// Cross-workgroup visibility uses agent-scope release/acquire fences around
// per-tile ready flags (L2 write-back / invalidate on gfx950).
#include <hip/hip_runtime.h>

#include "chol_tasks.h"
#include "kernels.h"
#include "tile_device.h"

namespace nsn {

__device__ __forceinline__ bool chol_gate(const DevCtl* c, int gate) {
  switch (gate) {
    case GATE_ALWAYS: return true;
    case GATE_NOT_DONE: return !c->done;
    case GATE_NEED_LIN: return !c->done && c->need_lin;
    case GATE_SOLVE_OK: return !c->done && c->solve_ok;
    case GATE_ROLLBACK: return c->rollback_pending != 0;
    case GATE_IR: return !c->done && c->ir_more;
    case GATE_IR_ANY: return c->ir_more != 0;
  }
  return true;
}

constexpr int kWaitLimit = 1 << 22;

// called by one thread; returns after the flag is set or the wait gave up
__device__ __forceinline__ void wait_flag(int* flag, DevCtl* ctl) {
  int it = 0;
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    __builtin_amdgcn_s_sleep(2);
    if (((++it) & 255) == 0) {
      if (__hip_atomic_load(&ctl->wait_timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
      if (it > kWaitLimit) {
        atomicAdd(&ctl->wait_timeout, 1);
        return;
      }
    }
  }
}

__device__ __forceinline__ void publish_flag(int* flag) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Tile DAG factorisation, one persistent launch.
//
// Workgroup 0 runs the diagonal chain, the critical path of the DAG: for
// column j it takes the pre-updated diagonal tile C_jj, subtracts
// L_{j,j-1} L_{j,j-1}^T with L_{j,j-1} still in its LDS from the previous
// step, factorises it in FP64 (POTRF) and inverts the factor; then it takes
// the pre-updated sub-diagonal tile C_{j+1,j} and forms L_{j+1,j} =
// C_{j+1,j} L_jj^-T as an MFMA product, keeping it for step j+1.  Nothing on
// that path crosses workgroups except the pre-updated tiles, which are
// ready long before they are needed.
//
// The other workgroups take tasks from an atomic counter, grouped in slots
// j = 0, 1, ...:
//   slot j: pre-update of C_{j+1,j} with every k < j into the scratch tile
//           presub[j], and of C_{j+1,j+1} (j >= 1) in place, each flagged;
//           L_{j+1,j} = C_{j+1,j} L_jj^-T into the tile storage for the
//           other tasks (the chain keeps its own copy and never publishes
//           it); then the tiles (i, j), i >= j+2: C_ij -= sum_{k<j} L_ik
//           L_jk^T, L_ij = C_ij L_jj^-T with the inverse published by
//           workgroup 0.
// A task of slot j only needs tiles of slots < j, the diagonal chain up to
// step j, and (pre-updates) sub-diagonal tiles of steps < j; step j of the
// chain needs only slots < j and the pre-update of C_{j+1,j} in slot j,
// which comes before every task that waits for step j.  A workgroup
// finishes a task before taking the next and tasks are taken in order, so
// every awaited task belongs to a running workgroup: the DAG cannot
// deadlock.  Every wait is bounded.
//
// MFMA tiling: each of the four waves owns a 32x32 quadrant of the 64x64
// tile in accumulator registers; update operands come straight from global
// memory into the operand registers, dependency waits are batched (lane l of
// wave 0 checks step k + l, one acquire fence per run of published steps).
// LDS: 64 KiB = two TB x TB FP64 regions.
// ---------------------------------------------------------------------------
// acc -= sum_{kbeg <= k < kend} L_ik L_jk^T over published tiles (batched
// waits); kbeg >= max(first[i], first[j]), the first k with both tiles
// inside the envelope.
// Operands go from global memory straight into the MFMA operand registers
// (a register double buffer across products measured slower: the diagonal
// chain shares this kernel's register budget).
template <typename T, int TB>
__device__ __forceinline__ void acc_update(CholAcc<T>& acc, const T* __restrict__ tiles, int* ready, DevCtl* ctl,
                                           int* shv, int i, int j, int kbeg, int kend, const TileView& tv, int lane,
                                           int wr, int wc) {
  const size_t T2 = (size_t)TB * TB;
  int k = kbeg;
  while (k < kend) {
    if (threadIdx.x < 64) {
      if (lane == 0) {
        wait_flag(ready + tv.slot(i, k), ctl);
        if (i != j) wait_flag(ready + tv.slot(j, k), ctl);
      }
      bool ok = false;
      if (k + lane < kend)
        ok = __hip_atomic_load(ready + tv.slot(i, k + lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
             (i == j || __hip_atomic_load(ready + tv.slot(j, k + lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      const unsigned long long m = __ballot(ok);
      const int run = ~m == 0ull ? 64 : __builtin_ctzll(~m);
      if (lane == 0) shv[1] = k + (run > 0 ? run : 1);
    }
    __syncthreads();
    const int ke = shv[1];
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (; k < ke; k++) {
      const T* Ag = tiles + (size_t)tv.slot(i, k) * T2 + wr * 32 + (lane & 15);
      const T* Bg = tiles + (size_t)tv.slot(j, k) * T2 + wc * 32 + (lane & 15);
      const int kq = lane >> 4;
#pragma unroll 8
      for (int m = 0; m < TB; m += 4) {
        const size_t off = (size_t)(m + kq) * TB;
        const T a0 = -Ag[off], a1 = -Ag[off + 16];
        const T b0 = Bg[off], b1 = Bg[off + 16];
        acc.a[0][0] = Mfma<T>::fma(a0, b0, acc.a[0][0]);
        acc.a[0][1] = Mfma<T>::fma(a0, b1, acc.a[0][1]);
        acc.a[1][0] = Mfma<T>::fma(a1, b0, acc.a[1][0]);
        acc.a[1][1] = Mfma<T>::fma(a1, b1, acc.a[1][1]);
      }
    }
  }
}

// a sum task's chunks (CT_PART scratch tiles p0 .. p0+np-1, each the
// negated sum of its chunk) added to acc in order; returns the first k of
// its own last chunk
template <typename T, int TB>
__device__ __forceinline__ int acc_add_parts(CholAcc<T>& acc, const T* __restrict__ part, int* part_ready,
                                             DevCtl* ctl, int p0, int np, int k0, int kend, int lane, int wr,
                                             int wc) {
  if (np <= 0) return k0;
  const size_t T2 = (size_t)TB * TB;
  for (int p = p0; p < p0 + np; p++) {
    if (threadIdx.x == 0) wait_flag(part_ready + p, ctl);
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    acc_add<T, TB>(acc, part + (size_t)p * T2, lane, wr, wc);
  }
  return max(k0, (kend - 1) / kCholPartLen * kCholPartLen);
}

template <typename T, int TB>
__global__ __launch_bounds__(256) void k_chol(TileView tv, T* __restrict__ tiles, T* __restrict__ inv,
                                              T* __restrict__ presub, T* __restrict__ part,
                                              const int* __restrict__ tasks, int ntasks, int* ready, int* pre_sub,
                                              int* pre_diag, int* part_ready, int* counter, DevCtl* ctl,
                                              int gate) {
  static_assert(TB == 64, "the MFMA tiling assumes 64x64 tiles and 4 waves");
  if (!chol_gate(ctl, gate)) return;
  __shared__ double lds[2 * TB * TB];
  double* R0d = lds;                 // FP64 diagonal tile (POTRF)
  double* R1d = lds + TB * TB;       // its inverse
  T* R0 = (T*)lds;                   // C / X tiles of the off-diagonal solves
  int* shv = (int*)(lds + TB * TB);  // scalars, in region 1 while it is unused
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv & 1, wc = wv >> 1;
  const int NT = tv.NT;
  const size_t T2 = (size_t)TB * TB;
  if (blockIdx.x < tv.nchain) {
    // ---- a diagonal chain: the steps chain_start[c] .. chain_start[c+1]-1 ----
    const int j0 = tv.chain_start[blockIdx.x], j1 = tv.chain_start[blockIdx.x + 1];
    CholAcc<T> acc;  // the next diagonal tile, loaded one step ahead
    acc_load<T, TB>(acc, tiles + (size_t)tv.slot(j0, j0) * T2, lane, wr, wc);
    bool have_sub = false;  // L_{j,j-1} of the previous step in R0
    for (int j = j0; j < j1; j++) {
      T* Cg = tiles + (size_t)tv.slot(j, j) * T2;
      if (have_sub) acc_gemm_nt<T, T, TB>(acc, R0, R0, true, lane, wr, wc);  // - L_{j,j-1} L_{j,j-1}^T
      __syncthreads();
      acc_store<T, double, TB>(acc, R0d, lane, wr, wc);
      __syncthreads();
      potrf_lds<TB>(R0d, R1d, j * TB, tv.n, ctl);  // R1d is free until the inverse
      tri_inverse<TB>(R0d, R1d);
      T* Ig = inv + (size_t)j * T2;
#pragma unroll 4
      for (int e = threadIdx.x; e < TB * TB; e += 256) Ig[e] = (T)R1d[e];
      publish_flag(ready + tv.slot(j, j));
      // the factor tile itself is read only after the launch (solves): its
      // stores go out after the flag, so the release above waits only for
      // the inverse (R0d is overwritten only after the next barrier)
#pragma unroll 4
      for (int e = threadIdx.x; e < TB * TB; e += 256) Cg[e] = (T)R0d[e];
      have_sub = false;
      if (j + 1 < j1) {
        const int f1 = tv.first[j + 1];
        if (f1 <= j) {
          // the pre-updated C_{j+1,j+1} (in flight during the product below);
          // its pre-update exists when row j+1 reaches a column before j
          if (threadIdx.x == 0) {
            if (f1 < j) wait_flag(pre_diag + j + 1, ctl);
            wait_flag(pre_sub + j, ctl);
          }
          __syncthreads();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          acc_load<T, TB>(acc, tiles + (size_t)tv.slot(j + 1, j + 1) * T2, lane, wr, wc);
          // L_{j+1,j} for the next step, from the pre-updated C_{j+1,j}; a
          // worker forms the same product for the tile storage (CT_SUB)
          global_to_lds<T, TB>(R0, presub + (size_t)j * T2);
          __syncthreads();
          TriProd<T> x;
          tri_gemm_nt<T, double, TB>(x, R0, R1d, lane, wv);  // C L_jj^-T
          __syncthreads();
          tri_store<T, TB>(x, R0, lane, wv);  // kept for step j + 1
          __syncthreads();
          have_sub = true;
        } else {
          // step j+1 starts another chain of this group: no tile left of its
          // diagonal, so its diagonal tile is final as assembled
          acc_load<T, TB>(acc, tiles + (size_t)tv.slot(j + 1, j + 1) * T2, lane, wr, wc);
        }
      }
    }
    return;
  }
  // ---- workers ----
  for (int round = 0; round < ntasks; round++) {  // at most ntasks tasks per workgroup
    if (threadIdx.x == 0) shv[0] = atomicAdd(counter, 1);
    __syncthreads();
    const int t = shv[0];
    __syncthreads();
    if (t >= ntasks) break;
    const int* rec = tasks + (size_t)kCholTaskInts * t;
    const int type = rec[0], i = rec[1], j = rec[2], pa = rec[3], pb = rec[4];
    if (type == CT_PART) {
      // the chunk pb of the sum over the rows (i, j): -sum L_ik L_jk^T into
      // the scratch tile pa
      const int kb = max(max(tv.first[i], tv.first[j]), pb * kCholPartLen);
      CholAcc<T> acc;
      acc_clear<T>(acc);
      acc_update<T, TB>(acc, tiles, ready, ctl, shv, i, j, kb, (pb + 1) * kCholPartLen, tv, lane, wr, wc);
      acc_store<T, T, TB>(acc, part + (size_t)pa * T2, lane, wr, wc);
      publish_flag(part_ready + pa);
    } else if (type == CT_TILE) {
      T* Cg = tiles + (size_t)tv.slot(i, j) * T2;
      CholAcc<T> acc;
      acc_load<T, TB>(acc, Cg, lane, wr, wc);
      const int kb = acc_add_parts<T, TB>(acc, part, part_ready, ctl, pa, pb, max(tv.first[i], tv.first[j]), j,
                                          lane, wr, wc);
      acc_update<T, TB>(acc, tiles, ready, ctl, shv, i, j, kb, j, tv, lane, wr, wc);
      acc_store<T, T, TB>(acc, R0, lane, wr, wc);
      if (threadIdx.x == 0) wait_flag(ready + tv.slot(j, j), ctl);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // X = C L_jj^-T with the inverse published by the diagonal chain
      TriProd<T> x;
      tri_gemm_nt<T, T, TB>(x, R0, inv + (size_t)j * T2, lane, wv);
      __syncthreads();
      tri_store<T, TB>(x, R0, lane, wv);
      __syncthreads();
      lds_to_global<T, TB>(Cg, R0);
      publish_flag(ready + tv.slot(i, j));
    } else if (type == CT_SUB) {
      // L_{j+1,j} = C_{j+1,j} L_jj^-T into the tile storage (the diagonal
      // chain keeps its own copy)
      if (threadIdx.x == 0) {
        wait_flag(pre_sub + j, ctl);
        wait_flag(ready + tv.slot(j, j), ctl);
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      global_to_lds<T, TB>(R0, presub + (size_t)j * T2);
      __syncthreads();
      TriProd<T> x;
      tri_gemm_nt<T, T, TB>(x, R0, inv + (size_t)j * T2, lane, wv);
      __syncthreads();
      tri_store<T, TB>(x, R0, lane, wv);
      __syncthreads();
      lds_to_global<T, TB>(tiles + (size_t)tv.slot(i, j) * T2, R0);
      publish_flag(ready + tv.slot(i, j));
    } else {
      // pre-update with every k < j of C_{j+1,j} (CT_PRE_SUB, into the
      // scratch tile presub[j]) or of C_{j+1,j+1} (CT_PRE_DIAG, in place)
      const int jc = type == CT_PRE_SUB ? j : j + 1;
      CholAcc<T> acc;
      acc_load<T, TB>(acc, tiles + (size_t)tv.slot(i, jc) * T2, lane, wr, wc);
      const int kb = acc_add_parts<T, TB>(acc, part, part_ready, ctl, pa, pb, max(tv.first[i], tv.first[jc]), j,
                                          lane, wr, wc);
      acc_update<T, TB>(acc, tiles, ready, ctl, shv, i, jc, kb, j, tv, lane, wr, wc);
      acc_store<T, T, TB>(acc, R0, lane, wr, wc);
      __syncthreads();
      lds_to_global<T, TB>(type == CT_PRE_SUB ? presub + (size_t)j * T2 : tiles + (size_t)tv.slot(i, jc) * T2,
                           R0);
      publish_flag(type == CT_PRE_SUB ? pre_sub + j : pre_diag + j + 1);
    }
  }
}

// Triangular solve with the tile factor and the inverses of its diagonal
// tiles (inv), over the NR = ceil(n / TB) tile rows of S, in solve order
// (backward, L^T x = y: blocks NR-1 .. 0; forward, L z = r: blocks 0 .. NR-1).
// The chain workgroups are the factorisation's chain groups: group c owns
// the blocks chain_start[c] .. chain_start[c+1]-1, a contiguous run of solve
// indices (independent parts of a nested-dissection order solve at the same
// time: forward the parts before their separators, backward the separators
// before their parts).  For the block k of each of its solve steps a chain
// forms
//   v = rhs_k - (far_k + sum_{d=1..kNear} N_{k,d} x_{k-d})
//   x_k = Linv_kk^(T) v
// with its kNear previous solutions kept in its LDS (N_{k,d} = L_{k,k-d}
// forward, L_{k+d,k}^T backward; only sources of its own run: near_lo) and
// the tiles, rhs_k and far_k of the next step loaded into registers during
// the current one (far_k speculatively: an entry still unset is polled again
// when it is needed).  Every other workgroup owns destination blocks (the
// solve indices t >= tf with a possible far source, round robin): it takes
// the solutions in solve order and accumulates, in its LDS and in source
// order (deterministic), far_t = sum of N x_s over the sources s <
// near_lo(t), and publishes far_t after its last far source.  Deadlock
// freedom: the unsolved block of least solve index has every earlier block
// solved, so its far sum gets published, and its chain has solved every
// earlier block of its run.
//
// Publication carries no flags or fences: sol and far hold kUnsetBits (a
// signalling NaN, which arithmetic never produces) before the launch, every
// entry is written once with an agent-scope atomic store, and a reader polls
// the entries it needs with agent-scope atomic loads until none is unset.
// Each entry is its own ready flag, so a writer never waits for its stores
// to drain before signalling, and a reader gets the data with the poll.
// rhs == nullptr takes the right-hand side from row n of the factor (the
// forward solve of bS that the factorisation leaves there).
constexpr int kNear = 2;
constexpr int kMaxOwned = 16;  // destination blocks per workgroup
constexpr int kLdsFirst = kMaxTileRows;  // tile rows of the envelope kept in LDS

__device__ __forceinline__ bool is_unset(double v) { return __builtin_bit_cast(unsigned long long, v) == kUnsetBits; }

__device__ __forceinline__ double load_pub(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_pub(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// polls one published entry (per lane) until it is set or the wait gave up
// (then 0 and ctl->wait_timeout raised); v: a value already loaded from p
__device__ __forceinline__ double wait_pub(const double* p, double v, DevCtl* ctl) {
  int it = 0;
  while (is_unset(v)) {
    __builtin_amdgcn_s_sleep(1);
    if (((++it) & 255) == 0) {
      if (__hip_atomic_load(&ctl->wait_timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return 0.0;
      if (it > kWaitLimit) {
        atomicAdd(&ctl->wait_timeout, 1);
        return 0.0;
      }
    }
    v = load_pub(p);
  }
  return v;
}

template <typename T, int TB, bool BWD>
__global__ __launch_bounds__(256) void k_trsv(TileView tv, const T* __restrict__ tiles, const T* __restrict__ inv,
                                              const double* rhs, double* sol, double* far, DevCtl* ctl, int gate) {
  static_assert(TB == 64, "64-row tiles");
  if (!chol_gate(ctl, gate)) return;
  typedef TileGemv<T, TB, BWD> Gemv;
  __shared__ __align__(16) double xs[2][TB];
  __shared__ __align__(16) double xr[kNear][TB];
  __shared__ double acc[kMaxOwned][TB];
  // the envelope (first tile column and slot offset per tile row) in LDS:
  // every tile address and the destination loop's test of every (source,
  // destination) pair read it, and a global read would put a memory round
  // trip before each tile load
  __shared__ int sfirst[kLdsFirst], soff[kLdsFirst];
  __shared__ int scs[kCholMaxChains + 1];
  const int n = tv.n, NT = tv.NT;
  for (int i = threadIdx.x; i < NT; i += blockDim.x) {
    sfirst[i] = tv.first[i];
    soff[i] = tv.row_off[i];
  }
  const int nch = tv.nchain;
  for (int i = threadIdx.x; i <= nch; i += blockDim.x) scs[i] = tv.chain_start[i];
  __syncthreads();
  auto first_of = [&](int i) { return sfirst[i]; };
  auto slot_of = [&](int i, int j) { return (long long)soff[i] + (j - sfirst[i]); };
  const int NR = (n + TB - 1) / TB;
  const size_t T2 = (size_t)TB * TB;
  const int step = BWD ? -1 : 1;  // solve order
  auto blk = [&](int s) { return BWD ? NR - 1 - s : s; };
  // the tile applied to x_k for destination i (L_ik forward, L_ki^T backward)
  auto ntile = [&](int i, int k) { return tiles + (size_t)(BWD ? slot_of(k, i) : slot_of(i, k)) * T2; };
  // the tile applied to x_k for destination i is inside the envelope
  auto ntile_has = [&](int i, int k) { return BWD ? i >= first_of(k) : k >= first_of(i); };
  const bool own = Gemv::owner();
  const int i0 = Gemv::index(0), i1 = Gemv::index(1);
  // the solve-index run [run_lo(c), run_hi(c)) of chain group c
  auto run_lo = [&](int c) { return BWD ? NR - min(scs[c + 1], NR) : min(scs[c], NR); };
  auto run_hi = [&](int c) { return BWD ? NR - min(scs[c], NR) : min(scs[c + 1], NR); };
  // the first near source of solve index t: kNear back, not before its run
  auto near_lo = [&](int t) {
    const int k = blk(t);
    int c = 0;
    while (c + 1 < nch && scs[c + 1] <= k) c++;
    return max(t - kNear, run_lo(c));
  };
  // the first solve index with a possible far source (near_lo >= 1); every
  // later one has one too (near_lo does not decrease)
  int tf = 1;
  while (tf < NR && near_lo(tf) < 1) tf++;
  if (blockIdx.x < nch) {
    // ---- a chain: the solve indices [sb, se) ----
    const int sb = run_lo(blockIdx.x), se = run_hi(blockIdx.x);
    if (sb >= se) return;
    // the loads of a step's tiles are issued as soon as the registers they
    // replace are consumed (issuing 96 KB takes about a microsecond on one
    // CU: spread over the step it overlaps the rest of the step's work)
    Gemv gd, gn[kNear];
    double r0 = 0.0, r1 = 0.0, f0 = 0.0, f1 = 0.0;
    auto rhs_at = [&](int row) {
      return rhs ? rhs[row] : (double)tiles[(size_t)slot_of(n / TB, row / TB) * T2 + (row % TB) * TB + (n % TB)];
    };
    // the near tile d of solve step s exists (inside the matrix and the
    // envelope: a tile row that starts a chain has no tile left of its
    // diagonal, and the one after it none before that chain)
    auto near_has = [&](int s, int d) {
      const int k = blk(s);
      return s - d >= sb && ntile_has(k, k - d * step);
    };
    auto load_near = [&](int s) {
      const int k = blk(s);
#pragma unroll
      for (int d = 1; d <= kNear; d++)
        if (near_has(s, d)) gn[d - 1].load(ntile(k, k - d * step));
    };
    auto load_vec = [&](int s) {
      const int k = blk(s);
      if (own) {
        const int a = k * TB + i0, b = k * TB + i1;
        f0 = (s >= tf && a < n) ? load_pub(far + a) : 0.0;
        f1 = (s >= tf && b < n) ? load_pub(far + b) : 0.0;
        r0 = a < n ? rhs_at(a) : 0.0;
        r1 = b < n ? rhs_at(b) : 0.0;
      }
    };
    load_vec(sb);
    gd.load(inv + (size_t)blk(sb) * T2);
    load_near(sb);
    for (int s = sb; s < se; s++) {
      const int k = blk(s);
      // the near blocks' contributions, one reduction for all of them
      double h0 = 0.0, h1 = 0.0;
#pragma unroll
      for (int d = 1; d <= kNear; d++)
        if (near_has(s, d)) gn[d - 1].partial(xr[(s - d) % kNear], h0, h1);
      if (s + 1 < se) load_near(s + 1);
      Gemv::reduce(h0, h1);
      if (own) {
        const int a = k * TB + i0, b = k * TB + i1;
        if (a < n) f0 = wait_pub(far + a, f0, ctl);
        if (b < n) f1 = wait_pub(far + b, f1, ctl);
        xs[0][i0] = a < n ? r0 - (f0 + h0) : 0.0;
        xs[0][i1] = b < n ? r1 - (f1 + h1) : 0.0;
      }
      __syncthreads();
      double x0 = 0.0, x1 = 0.0;
      gd.partial(xs[0], x0, x1);
      if (s + 1 < se) gd.load(inv + (size_t)blk(s + 1) * T2);
      Gemv::reduce(x0, x1);
      if (own) {
        const int a = k * TB + i0, b = k * TB + i1;
        if (a < n) store_pub(sol + a, x0);
        if (b < n) store_pub(sol + b, x1);
        xr[s % kNear][i0] = a < n ? x0 : 0.0;
        xr[s % kNear][i1] = b < n ? x1 : 0.0;
      }
      // x_k complete in the LDS before the next step reads it (and xs is
      // read by every wave before it is written again)
      __syncthreads();
      if (s + 1 < se) load_vec(s + 1);
    }
    return;
  }
  // ---- destination sums ----
  const int G = gridDim.x - nch, w = blockIdx.x - nch;
  const int first = tf + w;  // solve index of the first owned destination
  if (first >= NR) return;
  const int nown = (NR - 1 - first) / G + 1;  // <= kMaxOwned (launch)
  if (own)
    for (int m = 0; m < nown; m++) acc[m][i0] = acc[m][i1] = 0.0;
  // the products (source s, destination m) in source order, destinations in
  // order within a source: destination m takes the far sources s <
  // near_lo(its solve index) whose tile is inside the envelope; the tile of
  // the next product is loaded while the current one runs
  __shared__ int slast[kMaxOwned];
  if (threadIdx.x < nown) slast[threadIdx.x] = near_lo(first + G * threadIdx.x) - 1;
  __syncthreads();
  auto s_last = [&](int m) { return slast[m]; };
  const int s_end = s_last(nown - 1);  // no product after the last destination's last source
  auto find_next = [&](int& ps, int& pm) {  // the first product at or after (ps, pm)
    for (; ps <= s_end; ps++, pm = 0) {
      for (; pm < nown; pm++)
        if (ps <= s_last(pm) && ntile_has(blk(first + G * pm), blk(ps))) return true;
    }
    return false;
  };
  Gemv gg;
  int cs = 0, cm = 0;
  bool have = find_next(cs, cm);
  if (have) gg.load(ntile(blk(first + G * cm), blk(cs)));
  int m0 = 0;  // first owned destination still open
  int pc = 0;  // sources polled so far (the parity of the LDS buffer)
  // publishes the destinations whose last far source is s (a destination
  // with no tile of its row inside the envelope publishes 0)
  auto publish_done = [&](int s) {
    while (m0 < nown && s_last(m0) <= s) {
      const int d = blk(first + G * m0);
      if (own) {
        if (d * TB + i0 < n) store_pub(far + d * TB + i0, acc[m0][i0]);
        if (d * TB + i1 < n) store_pub(far + d * TB + i1, acc[m0][i1]);
      }
      m0++;
    }
  };
  for (int s = 0; s < NR && m0 < nown; s++) {  // source steps
    if (have && cs == s) {
      const int k = blk(s);
      // wave 0 polls the source's solution into the LDS (double-buffered by
      // the count of polled sources: the barrier of the previous one orders
      // every read of this buffer, two polled sources back, before the write)
      double* xb = xs[pc & 1];
      if (threadIdx.x < TB) {
        const int row = k * TB + threadIdx.x;
        xb[threadIdx.x] = row < n ? wait_pub(sol + row, load_pub(sol + row), ctl) : 0.0;
      }
      __syncthreads();
      pc++;
      for (;;) {
        double v0 = 0.0, v1 = 0.0;
        gg.partial(xb, v0, v1);
        Gemv::reduce(v0, v1);
        if (own) {
          acc[cm][i0] += v0;
          acc[cm][i1] += v1;
        }
        // the next destination with a product at this source
        int nm = cm + 1;
        while (nm < nown && !(s <= s_last(nm) && ntile_has(blk(first + G * nm), k))) nm++;
        if (nm == nown) break;
        cm = nm;
        gg.load(ntile(blk(first + G * cm), k));
      }
      // the sums complete at this source go out before the scan for the
      // next product, which may pass many sources
      publish_done(s);
      cs = s + 1;
      cm = 0;
      have = find_next(cs, cm);
      if (have) gg.load(ntile(blk(first + G * cm), blk(cs)));
    } else {
      publish_done(s);
    }
  }
}

template <typename T>
__global__ void k_chol_reset(int* flags, int nflags, int* counters, double* unset0, double* unset1, int nunset,
                             DevCtl* ctl, int gate) {
  if (!chol_gate(ctl, gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nflags) flags[i] = 0;
  if (i < 4) counters[i] = 0;
  if (i < nunset) {
    const double u = __builtin_bit_cast(double, kUnsetBits);
    unset0[i] = u;
    unset1[i] = u;
  }
}

template <typename T>
void launch_chol_reset(int* flags, int nflags, int* counters, double* unset0, double* unset1, int nunset,
                       DevCtl* ctl, int gate, hipStream_t st) {
  int n = nflags > 4 ? nflags : 4;
  if (nunset > n) n = nunset;
  hipLaunchKernelGGL((k_chol_reset<T>), dim3((n + 255) / 256), dim3(256), 0, st, flags, nflags, counters, unset0,
                     unset1, nunset, ctl, gate);
}

template <typename T>
void launch_chol(const TileView& tv, T* tiles, T* inv, T* presub, T* part, const int* tasks, int ntasks,
                 int* ready, int* pre_sub, int* pre_diag, int* part_ready, int* counter, DevCtl* ctl, int gate,
                 int max_wg, hipStream_t st) {
  // the chain workgroups, then the workers (the structure stage keeps
  // nchain well below max_wg)
  const int nc = tv.nchain;
  int g = nc + (ntasks < max_wg - nc ? ntasks : max_wg - nc);
  hipLaunchKernelGGL((k_chol<T, 64>), dim3(g), dim3(256), 0, st, tv, tiles, inv, presub, part, tasks, ntasks,
                     ready, pre_sub, pre_diag, part_ready, counter, ctl, gate);
}

template <typename T>
void launch_trsv(const TileView& tv, const T* tiles, const T* inv, bool backward, const double* rhs, double* sol,
                 double* far, DevCtl* ctl, int gate, int max_wg, hipStream_t st) {
  const int NR = (tv.n + 63) / 64;
  if (NR == 0 || tv.NT > kLdsFirst || tv.nchain > kCholMaxChains) return;  // (the structure stage refuses these)
  // the chain groups plus G workgroups owning the destinations with far
  // sources (at most NR - 1 of them: solve index 0 has none)
  const int nc = tv.nchain;
  const int nd = NR - 1;
  int G = nd < max_wg - nc ? nd : max_wg - nc;
  if (G > 0 && (nd + G - 1) / G > kMaxOwned) G = (nd + kMaxOwned - 1) / kMaxOwned;
  const int g = nc + G;
  if (backward)
    hipLaunchKernelGGL((k_trsv<T, 64, true>), dim3(g), dim3(256), 0, st, tv, tiles, inv, rhs, sol, far, ctl, gate);
  else
    hipLaunchKernelGGL((k_trsv<T, 64, false>), dim3(g), dim3(256), 0, st, tv, tiles, inv, rhs, sol, far, ctl, gate);
}

template void launch_chol_reset<double>(int*, int, int*, double*, double*, int, DevCtl*, int, hipStream_t);
template void launch_chol_reset<float>(int*, int, int*, double*, double*, int, DevCtl*, int, hipStream_t);
template void launch_chol<double>(const TileView&, double*, double*, double*, double*, const int*, int, int*,
                                  int*, int*, int*, int*, DevCtl*, int, int, hipStream_t);
template void launch_chol<float>(const TileView&, float*, float*, float*, float*, const int*, int, int*, int*,
                                 int*, int*, int*, DevCtl*, int, int, hipStream_t);
template void launch_trsv<double>(const TileView&, const double*, const double*, bool, const double*, double*,
                                  double*, DevCtl*, int, int, hipStream_t);
template void launch_trsv<float>(const TileView&, const float*, const float*, bool, const double*, double*, double*,
                                 DevCtl*, int, int, hipStream_t);

}  // namespace nsn
