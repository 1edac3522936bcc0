// Dense Cholesky of the reduced camera system on gfx950, as a persistent
// tile DAG in one launch, plus the backward triangular solve.
//
// The matrix is [S, .; bS^T, 1] padded with identity to NT*TB rows and
// stored as packed lower tiles (TileView).  Factorising it gives L with
// L y = bS in row n, i.e. the forward solve of SolverCholeskySchur.hpp:140
// comes out of the factorisation for free.  The reference factorises S with
// Eigen's pivoted LDL^T (SolverCholeskyDense.hpp:15-18); for the damped,
// positive definite S of an LM step the two give the same solution up to
// rounding, and both report failure when S is not positive definite.
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

#include "kernels.h"

namespace nsn {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// 16x16x4 MFMA on one wave: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// output register v of lane l holds D[row(v, l)][l%16] where the f32 form
// keeps four consecutive rows per lane group (row = 4*(l/16) + v) and the
// f64 form interleaves them (row = 4*v + l/16).
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t fma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t fma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ bool chol_gate(const DevCtl* c, int gate) {
  switch (gate) {
    case GATE_ALWAYS: return true;
    case GATE_NOT_DONE: return !c->done;
    case GATE_NEED_LIN: return !c->done && c->need_lin;
    case GATE_SOLVE_OK: return !c->done && c->solve_ok;
    case GATE_ROLLBACK: return c->rollback_pending != 0;
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

// orders this wave's LDS accesses (cross-lane exchange through LDS inside one
// wave without a workgroup barrier)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Diagonal-tile factorisation on one 64x64 FP64 tile in LDS, in four
// 16-column panels.  A panel is processed by one wave with one row per lane
// and the panel row in registers: each step takes its pivot from the owning
// lane and broadcasts the pivot column by readlane; between panels the
// trailing update runs on the MFMA units of the four waves.  The factored
// tile is then inverted, so that the solves against it (the tiles below it
// and the backward substitution) become matrix products.
// ---------------------------------------------------------------------------
constexpr int kPW = 16;  // panel width

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The reciprocal of the factor's diagonal entry c is kept in the strictly
// upper triangle of the diagonal tile, which nothing else uses: (0, c+1) for
// c < TB-1 and (1, TB-1) for the last one.  Rows 0 and 1 are never written by
// the panel updates (rows >= 16) nor by the panels below the diagonal.
template <int TB>
__device__ __forceinline__ int rdiag_slot(int c) { return c < TB - 1 ? (c + 1) * TB : (TB - 1) * TB + 1; }

// 1/sqrt(d) from the hardware estimate refined by two Newton steps
__device__ __forceinline__ double rsqrt_f64(double d) {
  double r = __builtin_amdgcn_rsq(d);
  r = r * (1.5 - 0.5 * d * r * r);
  r = r * (1.5 - 0.5 * d * r * r);
  return r;
}

// right-looking Cholesky of columns C0 .. C0+15 of the 64-row FP64 tile A
// (rows >= C0), wave 0 only.  Rows with global index >= n (the bS row and the
// identity padding) take pivot 1 and never count as failures; a failed pivot
// is replaced by 1 and counted in `bad`.  The step is branch-free: every
// lane computes the same pivot, the reciprocal square root of the selected
// value, and the updates run on every lane (entries above the diagonal in the
// registers are garbage and never written back).  The pivot reciprocals go to
// their slots in the upper triangle after the panel.
template <int TB, int C0>
__device__ __forceinline__ void potrf_panel(double* A, int g0, int n, int& bad) {
  const int lane = threadIdx.x;
  double a[kPW];
#pragma unroll
  for (int c = 0; c < kPW; c++) a[c] = A[(C0 + c) * TB + lane];
  double myr = 1.0;
#pragma unroll
  for (int k = 0; k < kPW; k++) {
    const double d = readlane_f64(a[k], C0 + k);
    const bool pad = g0 + C0 + k >= n;
    const bool good = d > 0.0 && d < __builtin_inf();
    bad |= (!pad && !good) ? 1 : 0;
    const double r = (pad || !good) ? 1.0 : rsqrt_f64(d);
    const double piv = (pad || !good) ? 1.0 : d * r;
    a[k] = (lane == C0 + k) ? piv : a[k] * r;
    myr = (lane == C0 + k) ? r : myr;
#pragma unroll
    for (int c = k + 1; c < kPW; c++) a[c] -= a[k] * readlane_f64(a[k], C0 + c);
  }
#pragma unroll
  for (int c = 0; c < kPW; c++)
    if (lane >= C0 + c) A[(C0 + c) * TB + lane] = a[c];
  if (lane >= C0 && lane < C0 + kPW) A[rdiag_slot<TB>(lane)] = myr;
}

// D_(rb, cb) -= A_(rb, p) A_(cb, p)^T for one 16x16 block of the FP64 tile
// A in LDS (column-major), with the 16 columns of panel p; one wave, four
// 16x16x4 MFMAs
template <int TB>
__device__ __forceinline__ void block_update(double* A, int rb, int cb, int k0) {
  typedef Mfma<double>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  acc_t acc;
#pragma unroll
  for (int vv = 0; vv < 4; vv++) acc[vv] = A[(16 * cb + (lane & 15)) * TB + 16 * rb + Mfma<double>::row(vv, lane)];
#pragma unroll
  for (int m = 0; m < kPW; m += 4) {
    const int mm = k0 + m + (lane >> 4);
    acc = Mfma<double>::fma(-A[mm * TB + 16 * rb + (lane & 15)], A[mm * TB + 16 * cb + (lane & 15)], acc);
  }
#pragma unroll
  for (int vv = 0; vv < 4; vv++) A[(16 * cb + (lane & 15)) * TB + 16 * rb + Mfma<double>::row(vv, lane)] = acc[vv];
}

// Cholesky of a 64x64 FP64 tile in LDS (lower triangle written, the
// reciprocals of its diagonal in the upper triangle), four 16-column panels
// with look-ahead: after panel p the four waves first update block column
// p+1 (what panel p+1 needs), then wave 0 factorises panel p+1 while waves
// 1-3 apply panel p to the block columns beyond it.
template <int TB, int P>
__device__ __forceinline__ void potrf_next(double* A, int g0, int n, int& bad) {
  const int wv = threadIdx.x >> 6;
  if (wv < 3 - P) block_update<TB>(A, P + 1 + wv, P + 1, kPW * P);
  __syncthreads();
  if (wv == 0) {
    potrf_panel<TB, kPW * (P + 1)>(A, g0, n, bad);
  } else if (P == 0) {
    // blocks (2,2), (3,2), (3,3) with panel 0
    const int rb = wv == 3 ? 3 : wv + 1, cb = wv == 1 ? 2 : (wv == 2 ? 2 : 3);
    block_update<TB>(A, rb, cb, 0);
  } else if (P == 1 && wv == 1) {
    block_update<TB>(A, 3, 3, kPW);
  }
  __syncthreads();
}

template <int TB>
__device__ __forceinline__ void potrf_lds(double* A, int g0, int n, DevCtl* ctl) {
  static_assert(TB == 4 * kPW, "four panels");
  int bad = 0;
  if (threadIdx.x < 64) potrf_panel<TB, 0>(A, g0, n, bad);
  __syncthreads();
  potrf_next<TB, 0>(A, g0, n, bad);
  potrf_next<TB, 1>(A, g0, n, bad);
  potrf_next<TB, 2>(A, g0, n, bad);
  if (threadIdx.x == 0 && bad) atomicAdd(&ctl->fail_s, 1);
}

// Li = L^-1 for the lower-triangular 64x64 FP64 tile L in LDS (diagonal
// reciprocals in their upper-triangle slots); Li is written lower with
// zeros above the diagonal.  By 16x16 blocks: the four diagonal blocks are
// inverted by forward substitution (wave b, lane j < 16 solves
// L_bb x = e_j), then the off-diagonal blocks level by level,
//   Li_ib,jb = -Li_ib,ib * sum_{kb = jb}^{ib-1} L_ib,kb Li_kb,jb,
// one wave per block on the MFMA units.  The inner product is staged in the
// strictly upper block (jb, ib) of Li, which is cleared at the end.
template <int TB>
__device__ __forceinline__ void tri_inverse(const double* L, double* Li) {
  static_assert(TB == 64, "four 16x16 blocks");
  typedef Mfma<double>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane < 16) {
    const int b0 = wv * 16, j = lane;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; k++) s -= L[(b0 + k) * TB + b0 + i] * x[k];
      x[i] = (i >= j) ? s * L[rdiag_slot<TB>(b0 + i)] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) Li[(b0 + j) * TB + b0 + i] = x[i];
  }
  __syncthreads();
#pragma unroll
  for (int lev = 1; lev < 4; lev++) {
    if (wv < 4 - lev) {
      const int jb = wv, ib = wv + lev;
      acc_t t = {0.0, 0.0, 0.0, 0.0};
      for (int kb = jb; kb < ib; kb++) {
#pragma unroll
        for (int m = 0; m < 16; m += 4) {
          const int mm = 16 * kb + m + (lane >> 4);
          t = Mfma<double>::fma(L[mm * TB + 16 * ib + (lane & 15)], Li[(16 * jb + (lane & 15)) * TB + mm], t);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) Li[(16 * ib + Mfma<double>::row(v, lane)) * TB + 16 * jb + (lane & 15)] = t[v];
      wave_lds_sync();
      acc_t o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 0; m < 16; m += 4) {
        const int mm = 16 * ib + m + (lane >> 4);
        o = Mfma<double>::fma(-Li[mm * TB + 16 * ib + (lane & 15)], Li[mm * TB + 16 * jb + (lane & 15)], o);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) Li[(16 * jb + (lane & 15)) * TB + 16 * ib + Mfma<double>::row(v, lane)] = o[v];
    }
    __syncthreads();
  }
  // clear the staging blocks (jb, ib), jb < ib: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
  for (int e = threadIdx.x; e < 6 * 256; e += 256) {
    const int blk = e >> 8, w = e & 255;
    const int jb = blk < 3 ? 0 : (blk < 5 ? 1 : 2);
    const int ib = blk < 3 ? blk + 1 : (blk < 5 ? blk - 1 : 3);
    Li[(16 * ib + (w >> 4)) * TB + 16 * jb + (w & 15)] = 0.0;
  }
  __syncthreads();
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
enum CholTask : int { CT_TILE = 0, CT_PRE_SUB = 1, CT_PRE_DIAG = 2, CT_SUB = 3 };

template <typename T>
struct CholAcc {
  typename Mfma<T>::acc_t a[2][2];
};

// acc <- tile (global, column-major TB x TB)
template <typename T, int TB>
__device__ __forceinline__ void acc_load(CholAcc<T>& acc, const T* Cg, int lane, int wr, int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int vv = 0; vv < 4; vv++) {
        const int r = wr * 32 + bi * 16 + Mfma<T>::row(vv, lane);
        const int c = wc * 32 + bj * 16 + (lane & 15);
        acc.a[bi][bj][vv] = Cg[c * TB + r];
      }
}

// acc -> LDS tile (column-major), element type U
template <typename T, typename U, int TB>
__device__ __forceinline__ void acc_store(const CholAcc<T>& acc, U* D, int lane, int wr, int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int vv = 0; vv < 4; vv++) {
        const int r = wr * 32 + bi * 16 + Mfma<T>::row(vv, lane);
        const int c = wc * 32 + bj * 16 + (lane & 15);
        D[c * TB + r] = (U)acc.a[bi][bj][vv];
      }
}

// acc -= sum_{k < kend} L_ik L_jk^T over published tiles (batched waits)
template <typename T, int TB>
__device__ __forceinline__ void acc_update(CholAcc<T>& acc, const T* __restrict__ tiles, int* ready, DevCtl* ctl,
                                           int* shv, int i, int j, int kend, int NT, int lane, int wr, int wc) {
  const size_t T2 = (size_t)TB * TB;
  int k = 0;
  while (k < kend) {
    if (threadIdx.x < 64) {
      if (lane == 0) {
        wait_flag(ready + tile_index(i, k, NT), ctl);
        if (i != j) wait_flag(ready + tile_index(j, k, NT), ctl);
      }
      bool ok = false;
      if (k + lane < kend)
        ok = __hip_atomic_load(ready + tile_index(i, k + lane, NT), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
             (i == j ||
              __hip_atomic_load(ready + tile_index(j, k + lane, NT), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      const unsigned long long m = __ballot(ok);
      const int run = ~m == 0ull ? 64 : __builtin_ctzll(~m);
      if (lane == 0) shv[1] = k + (run > 0 ? run : 1);
    }
    __syncthreads();
    const int ke = shv[1];
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (; k < ke; k++) {
      const T* Ag = tiles + (size_t)tile_index(i, k, NT) * T2 + wr * 32 + (lane & 15);
      const T* Bg = tiles + (size_t)tile_index(j, k, NT) * T2 + wc * 32 + (lane & 15);
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

// acc (+)= A B^T for TB x TB operands: A column-major in LDS (element type
// T), B column-major (element type U, LDS or global); `sub` subtracts
template <typename T, typename U, int TB>
__device__ __forceinline__ void acc_gemm_nt(CholAcc<T>& acc, const T* A, const U* B, bool sub, int lane, int wr,
                                            int wc) {
  const T* As = A + wr * 32 + (lane & 15);
  const U* Bs = B + wc * 32 + (lane & 15);
  const int kq = lane >> 4;
#pragma unroll 8
  for (int m = 0; m < TB; m += 4) {
    const int off = (m + kq) * TB;
    const T a0 = sub ? -As[off] : As[off], a1 = sub ? -As[off + 16] : As[off + 16];
    const T b0 = (T)Bs[off], b1 = (T)Bs[off + 16];
    acc.a[0][0] = Mfma<T>::fma(a0, b0, acc.a[0][0]);
    acc.a[0][1] = Mfma<T>::fma(a0, b1, acc.a[0][1]);
    acc.a[1][0] = Mfma<T>::fma(a1, b0, acc.a[1][0]);
    acc.a[1][1] = Mfma<T>::fma(a1, b1, acc.a[1][1]);
  }
}

// tile copies between global memory and LDS: every load of a thread is
// issued before the first use (the copies sit on the diagonal chain)
template <typename T, int TB>
__device__ __forceinline__ void lds_to_global(T* Cg, const T* R) {
  constexpr int NQ = TB * TB / 256;
  T v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) v[q] = R[threadIdx.x + 256 * q];
#pragma unroll
  for (int q = 0; q < NQ; q++) Cg[threadIdx.x + 256 * q] = v[q];
}

template <typename T, int TB>
__device__ __forceinline__ void global_to_lds(T* R, const T* Cg) {
  constexpr int NQ = TB * TB / 256;
  T v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) v[q] = Cg[threadIdx.x + 256 * q];
#pragma unroll
  for (int q = 0; q < NQ; q++) R[threadIdx.x + 256 * q] = v[q];
}

template <typename T, int TB>
__global__ __launch_bounds__(256) void k_chol(TileView tv, T* __restrict__ tiles, T* __restrict__ inv,
                                              T* __restrict__ presub, const int* __restrict__ tasks, int ntasks,
                                              int* ready, int* pre_sub, int* pre_diag, int* counter, DevCtl* ctl,
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
  if (blockIdx.x == 0) {
    // ---- the diagonal chain ----
    CholAcc<T> acc;  // the next diagonal tile, loaded one step ahead
    acc_load<T, TB>(acc, tiles, lane, wr, wc);
    for (int j = 0; j < NT; j++) {
      T* Cg = tiles + (size_t)tile_index(j, j, NT) * T2;
      if (j >= 1) acc_gemm_nt<T, T, TB>(acc, R0, R0, true, lane, wr, wc);  // - L_{j,j-1} L_{j,j-1}^T
      __syncthreads();
      acc_store<T, double, TB>(acc, R0d, lane, wr, wc);
      __syncthreads();
      potrf_lds<TB>(R0d, j * TB, tv.n, ctl);
      tri_inverse<TB>(R0d, R1d);
      T* Ig = inv + (size_t)j * T2;
#pragma unroll 4
      for (int e = threadIdx.x; e < TB * TB; e += 256) {
        Cg[e] = (T)R0d[e];
        Ig[e] = (T)R1d[e];
      }
      publish_flag(ready + tile_index(j, j, NT));
      if (j + 1 < NT) {
        // the pre-updated C_{j+1,j+1} (in flight during the product below)
        if (threadIdx.x == 0) {
          if (j + 1 >= 2) wait_flag(pre_diag + j + 1, ctl);
          wait_flag(pre_sub + j, ctl);
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        acc_load<T, TB>(acc, tiles + (size_t)tile_index(j + 1, j + 1, NT) * T2, lane, wr, wc);
        // L_{j+1,j} for the next step, from the pre-updated C_{j+1,j}; a
        // worker forms the same product for the tile storage (CT_SUB)
        global_to_lds<T, TB>(R0, presub + (size_t)j * T2);
        __syncthreads();
        CholAcc<T> x;
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
          for (int bj = 0; bj < 2; bj++) x.a[bi][bj] = typename Mfma<T>::acc_t{};
        acc_gemm_nt<T, double, TB>(x, R0, R1d, false, lane, wr, wc);  // C L_jj^-T
        __syncthreads();
        acc_store<T, T, TB>(x, R0, lane, wr, wc);  // kept for step j + 1
        __syncthreads();
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
    const int type = tasks[3 * t], i = tasks[3 * t + 1], j = tasks[3 * t + 2];
    if (type == CT_TILE) {
      T* Cg = tiles + (size_t)tile_index(i, j, NT) * T2;
      CholAcc<T> acc;
      acc_load<T, TB>(acc, Cg, lane, wr, wc);
      acc_update<T, TB>(acc, tiles, ready, ctl, shv, i, j, j, NT, lane, wr, wc);
      acc_store<T, T, TB>(acc, R0, lane, wr, wc);
      if (threadIdx.x == 0) wait_flag(ready + tile_index(j, j, NT), ctl);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // X = C L_jj^-T with the inverse published by the diagonal chain
      CholAcc<T> x;
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) x.a[bi][bj] = typename Mfma<T>::acc_t{};
      acc_gemm_nt<T, T, TB>(x, R0, inv + (size_t)j * T2, false, lane, wr, wc);
      __syncthreads();
      acc_store<T, T, TB>(x, R0, lane, wr, wc);
      __syncthreads();
      lds_to_global<T, TB>(Cg, R0);
      publish_flag(ready + tile_index(i, j, NT));
    } else if (type == CT_SUB) {
      // L_{j+1,j} = C_{j+1,j} L_jj^-T into the tile storage (the diagonal
      // chain keeps its own copy)
      if (threadIdx.x == 0) {
        wait_flag(pre_sub + j, ctl);
        wait_flag(ready + tile_index(j, j, NT), ctl);
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      global_to_lds<T, TB>(R0, presub + (size_t)j * T2);
      __syncthreads();
      CholAcc<T> x;
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) x.a[bi][bj] = typename Mfma<T>::acc_t{};
      acc_gemm_nt<T, T, TB>(x, R0, inv + (size_t)j * T2, false, lane, wr, wc);
      __syncthreads();
      acc_store<T, T, TB>(x, R0, lane, wr, wc);
      __syncthreads();
      lds_to_global<T, TB>(tiles + (size_t)tile_index(i, j, NT) * T2, R0);
      publish_flag(ready + tile_index(i, j, NT));
    } else {
      // pre-update with every k < j of C_{j+1,j} (CT_PRE_SUB, into the
      // scratch tile presub[j]) or of C_{j+1,j+1} (CT_PRE_DIAG, in place)
      const int jc = type == CT_PRE_SUB ? j : j + 1;
      CholAcc<T> acc;
      acc_load<T, TB>(acc, tiles + (size_t)tile_index(i, jc, NT) * T2, lane, wr, wc);
      acc_update<T, TB>(acc, tiles, ready, ctl, shv, i, jc, j, NT, lane, wr, wc);
      acc_store<T, T, TB>(acc, R0, lane, wr, wc);
      __syncthreads();
      lds_to_global<T, TB>(type == CT_PRE_SUB ? presub + (size_t)j * T2 : tiles + (size_t)tile_index(i, jc, NT) * T2,
                           R0);
      publish_flag(type == CT_PRE_SUB ? pre_sub + j : pre_diag + j + 1);
    }
  }
}

// 64 x 64 tile times a 64-vector in LDS, FP64 accumulation, with the tile
// prefetched into registers (16 entries per thread) before the vector is
// known.  TRANS: out[c] = sum_r A[r][c] x[r], thread (c = tid/4, part
// tid%4) holds rows 16*part .. +15 of column c (contiguous); combined by two
// lane exchanges.  Otherwise: out[r] = sum_c A[r][c] x[c], thread (r =
// tid%64, wave w) holds columns 16*w .. +15 of row r (coalesced loads);
// combined across the four waves through LDS.  Both sums run in a fixed
// order.
template <typename T, int TB, bool TRANS>
struct TileGemv {
  T a[16];
  __device__ __forceinline__ void load(const T* A) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 16; q++)
      a[q] = TRANS ? A[(tid >> 2) * TB + (tid & 3) * 16 + q] : A[((tid >> 6) * 16 + q) * TB + (tid & 63)];
  }
  // returns the full sum in every thread whose index() is valid (TRANS: all
  // four threads of a column; otherwise: wave 0); red: 4 x TB doubles of LDS
  __device__ __forceinline__ double apply(const double* x, double* red) const {
    const int tid = threadIdx.x;
    double s = 0.0;
    const int base = TRANS ? (tid & 3) * 16 : (tid >> 6) * 16;
#pragma unroll
    for (int q = 0; q < 16; q++) s += (double)a[q] * x[base + q];
    if (TRANS) {
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      return s;
    }
    red[(tid >> 6) * TB + (tid & 63)] = s;
    __syncthreads();
    s = 0.0;
    if (tid < TB) s = ((red[tid] + red[TB + tid]) + red[2 * TB + tid]) + red[3 * TB + tid];
    __syncthreads();
    return s;
  }
  static __device__ __forceinline__ int index() { return TRANS ? (threadIdx.x >> 2) : (int)threadIdx.x; }
  static __device__ __forceinline__ bool owner() { return TRANS ? (threadIdx.x & 3) == 0 : threadIdx.x < TB; }
};

// Triangular solve with the tile factor and the inverses of its diagonal
// tiles (inv), as a task DAG over the NR = ceil(n / TB) tile rows of S.
//   backward (L^T x = y): step s handles block k = NR-1-s: task D_k, then
//                         G_{k,i} for i = k-2 down to 0;
//   forward (L z = r):    step s handles block k = s: task D_k, then G_{k,i}
//                         for i = k+2 up to NR-1.
// G_{k,i} adds block k's contribution to block i (L_ki^T x_k backward,
// L_ik z_k forward) to the running sum of the blocks solved before k, so
// part(i, k) = part(i, k_prev) + contribution: one chain per destination i,
// in solve order (deterministic).
// D_k:  v = rhs_k - (part(k, last) + N sol_nb),  sol_k = Linv_kk^(T) v, with
//       nb the neighbouring block (k+1 backward, k-1 forward), N its tile
//       (L_{k+1,k}^T or L_{k,k-1}) and part(k, last) the chain of every other
//       solved block.
// rhs == nullptr takes the right-hand side from row n of the factor (the
// forward solve of bS that the factorisation leaves there).  Every task waits
// for at most two tasks, both earlier in the queue, so the DAG cannot
// deadlock.
template <typename T, int TB, bool BWD>
__global__ __launch_bounds__(256) void k_trsv(TileView tv, const T* __restrict__ tiles, const T* __restrict__ inv,
                                              const double* rhs, double* sol, double* __restrict__ parts,
                                              int* xready, int* gready, int* counter, DevCtl* ctl, int gate) {
  static_assert(TB == 64, "64-row tiles");
  if (!chol_gate(ctl, gate)) return;
  typedef TileGemv<T, TB, BWD> Gemv;
  __shared__ double red[4 * TB];
  __shared__ double xs[TB];
  __shared__ double hs[TB];
  __shared__ int sh_t;
  const int n = tv.n, NT = tv.NT;
  const int NR = (n + TB - 1) / TB;
  const int ntask = NR + (NR >= 2 ? (NR - 1) * (NR - 2) / 2 : 0);
  const size_t T2 = (size_t)TB * TB;
  const int tid = threadIdx.x;
  const int step = BWD ? -1 : 1;  // solve order
  for (int round = 0; round < ntask; round++) {  // bounded
    if (tid == 0) sh_t = atomicAdd(counter, 1);
    __syncthreads();
    int t = sh_t;
    __syncthreads();
    if (t >= ntask) break;
    int s = 0;
    while (t >= 1 + (NR - 2 - s > 0 ? NR - 2 - s : 0)) {
      t -= 1 + (NR - 2 - s > 0 ? NR - 2 - s : 0);
      s++;
    }
    const int k = BWD ? NR - 1 - s : s;
    if (t == 0) {
      // ---- D_k ----
      const int nb = k - step;       // solved just before k
      const int pl = k - 2 * step;   // last block of k's partial-sum chain
      const bool has_nb = nb >= 0 && nb < NR, has_part = pl >= 0 && pl < NR;
      Gemv gd, gn;
      gd.load(inv + (size_t)k * T2);
      if (has_nb) gn.load(tiles + (size_t)(BWD ? tile_index(nb, k, NT) : tile_index(k, nb, NT)) * T2);
      double r0 = 0.0;
      if (tid < TB && k * TB + tid < n)
        r0 = rhs ? rhs[k * TB + tid] : (double)tiles[(size_t)tile_index(n / TB, k, NT) * T2 + tid * TB + (n % TB)];
      if (tid == 0) {
        if (has_nb) wait_flag(xready + nb, ctl);
        if (has_part) wait_flag(gready + (size_t)k * NR + pl, ctl);
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      double pk = 0.0;
      if (tid < TB) {
        xs[tid] = (has_nb && nb * TB + tid < n) ? sol[nb * TB + tid] : 0.0;
        if (has_part) pk = parts[((size_t)k * NR + pl) * TB + tid];
      }
      __syncthreads();
      double h = 0.0;
      if (has_nb) h = gn.apply(xs, red);
      if (Gemv::owner()) hs[Gemv::index()] = h;
      __syncthreads();
      if (tid < TB) xs[tid] = (k * TB + tid < n) ? r0 - (pk + hs[tid]) : 0.0;
      __syncthreads();
      const double x = gd.apply(xs, red);
      if (Gemv::owner() && k * TB + Gemv::index() < n) sol[k * TB + Gemv::index()] = x;
      publish_flag(xready + k);
    } else {
      // ---- G_{k,i}: part(i, k) = part(i, k - step) + contribution of block k ----
      const int i = BWD ? k - 1 - t : k + 1 + t;
      const int kp = k - step;  // previous link of i's chain, if it is a source of i
      const bool has_prev = BWD ? kp < NR : kp >= 0;
      Gemv gg;
      gg.load(tiles + (size_t)(BWD ? tile_index(k, i, NT) : tile_index(i, k, NT)) * T2);
      if (tid == 0) {
        wait_flag(xready + k, ctl);
        if (has_prev) wait_flag(gready + (size_t)i * NR + kp, ctl);
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (tid < TB) xs[tid] = (k * TB + tid < n) ? sol[k * TB + tid] : 0.0;
      __syncthreads();
      const double v = gg.apply(xs, red);
      if (Gemv::owner()) {
        const int e = Gemv::index();
        parts[((size_t)i * NR + k) * TB + e] = (has_prev ? parts[((size_t)i * NR + kp) * TB + e] : 0.0) + v;
      }
      publish_flag(gready + (size_t)i * NR + k);
    }
  }
}

template <typename T>
__global__ void k_chol_reset(int* flags, int nflags, int* counters, DevCtl* ctl, int gate) {
  if (!chol_gate(ctl, gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nflags) flags[i] = 0;
  if (i < 4) counters[i] = 0;
}

template <typename T>
void launch_chol_reset(int* flags, int nflags, int* counters, DevCtl* ctl, int gate, hipStream_t st) {
  const int n = nflags > 4 ? nflags : 4;
  hipLaunchKernelGGL((k_chol_reset<T>), dim3((n + 255) / 256), dim3(256), 0, st, flags, nflags, counters, ctl, gate);
}

template <typename T>
void launch_chol(const TileView& tv, T* tiles, T* inv, T* presub, const int* tasks, int ntasks, int* ready,
                 int* pre_sub, int* pre_diag, int* counter, DevCtl* ctl, int gate, int max_wg, hipStream_t st) {
  int g = 1 + (ntasks < max_wg - 1 ? ntasks : max_wg - 1);
  hipLaunchKernelGGL((k_chol<T, 64>), dim3(g), dim3(256), 0, st, tv, tiles, inv, presub, tasks, ntasks, ready,
                     pre_sub, pre_diag, counter, ctl, gate);
}

template <typename T>
void launch_trsv(const TileView& tv, const T* tiles, const T* inv, bool backward, const double* rhs, double* sol,
                 double* parts, int* xready, int* gready, int* counter, DevCtl* ctl, int gate, int max_wg,
                 hipStream_t st) {
  const int NR = (tv.n + 63) / 64;
  const int ntask = NR + (NR >= 2 ? (NR - 1) * (NR - 2) / 2 : 0);
  if (ntask == 0) return;
  int g = ntask < max_wg ? ntask : max_wg;
  if (backward)
    hipLaunchKernelGGL((k_trsv<T, 64, true>), dim3(g), dim3(256), 0, st, tv, tiles, inv, rhs, sol, parts, xready,
                       gready, counter, ctl, gate);
  else
    hipLaunchKernelGGL((k_trsv<T, 64, false>), dim3(g), dim3(256), 0, st, tv, tiles, inv, rhs, sol, parts, xready,
                       gready, counter, ctl, gate);
}

template void launch_chol_reset<double>(int*, int, int*, DevCtl*, int, hipStream_t);
template void launch_chol_reset<float>(int*, int, int*, DevCtl*, int, hipStream_t);
template void launch_chol<double>(const TileView&, double*, double*, double*, const int*, int, int*, int*, int*,
                                  int*, DevCtl*, int, int, hipStream_t);
template void launch_chol<float>(const TileView&, float*, float*, float*, const int*, int, int*, int*, int*, int*,
                                 DevCtl*, int, int, hipStream_t);
template void launch_trsv<double>(const TileView&, const double*, const double*, bool, const double*, double*,
                                  double*, int*, int*, int*, DevCtl*, int, int, hipStream_t);
template void launch_trsv<float>(const TileView&, const float*, const float*, bool, const double*, double*, double*,
                                 int*, int*, int*, DevCtl*, int, int, hipStream_t);

}  // namespace nsn
