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

// unblocked right-looking Cholesky of a TB x TB column-major tile in LDS
// (FP64); g0 = global index of the tile's first row.  Rows >= n (the bS row
// and the identity padding) take pivot 1 and never count as failures.
template <int TB>
__device__ void potrf_lds(double* A, int g0, int n, DevCtl* ctl) {
  const int tid = threadIdx.x;
  constexpr int NS = 256 / TB;
  for (int k = 0; k < TB; k++) {
    if (tid == 0) {
      const double akk = A[k * TB + k];
      double piv;
      if (g0 + k >= n) piv = 1.0;
      else if (!(akk > 0.0) || !isfinite(akk)) {
        atomicAdd(&ctl->fail_s, 1);
        piv = 1.0;
      } else piv = sqrt(akk);
      A[k * TB + k] = piv;
    }
    __syncthreads();
    const double inv = 1.0 / A[k * TB + k];
    if (tid > k && tid < TB) A[k * TB + tid] *= inv;
    __syncthreads();
    const int i = tid % TB, js = tid / TB;
    if (i > k) {
      const double lik = A[k * TB + i];
      for (int jj = k + 1 + js; jj <= i; jj += NS) A[jj * TB + i] -= lik * A[k * TB + jj];
    }
    __syncthreads();
  }
}

// X L^T = C in place (C in Cs, L lower in Ls), both TB x TB column-major
template <typename T, int TB>
__device__ void trsm_lds(T* Cs, const T* Ls) {
  const int tid = threadIdx.x;
  constexpr int NS = 256 / TB;
  const int r = tid % TB, cs = tid / TB;
  for (int c = 0; c < TB; c++) {
    if (tid < TB) Cs[c * TB + tid] = Cs[c * TB + tid] / Ls[c * TB + c];
    __syncthreads();
    const T xr = Cs[c * TB + r];
    for (int c2 = c + 1 + cs; c2 < TB; c2 += NS) Cs[c2 * TB + r] -= xr * Ls[c * TB + c2];
    __syncthreads();
  }
}

// LDS budget: exactly 64 KiB per workgroup (tile buffer + staging buffer,
// plus a double tile for the FP64 pivot block of the FP32 path, which then
// has half-size buffers).  The dequeued tile index is passed through the
// first word of the tile buffer, which is dead between tiles.
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_chol(TileView tv, T* __restrict__ tiles, const int* __restrict__ tile_order,
                                              int* ready, int* counter, DevCtl* ctl, int gate) {
  static_assert(TB == 64, "the MFMA tiling below assumes 64x64 tiles and 4 waves");
  constexpr int KC = TB / 2;  // staged K columns per GEMM chunk
  if (!chol_gate(ctl, gate)) return;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Ct[TB * TB];        // C tile during potrf / trsm
  __shared__ T stage[2 * TB * KC];  // A | B chunks during GEMM, L_jj during trsm
  __shared__ double Dd[(sizeof(T) == 8) ? 1 : TB * TB];
  static_assert(sizeof(T) * TB * TB * 2 + ((sizeof(T) == 8) ? 8 : 8 * TB * TB) <= 65536 + 8,
                "LDS budget");
  T* As = stage;
  T* Bs = stage + TB * KC;
  int* sh_t = (int*)Ct;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv & 1, wc = wv >> 1;
  const int NT = tv.NT;
  const size_t T2 = (size_t)TB * TB;
  // a workgroup takes at most ntiles tiles: the loop is bounded
  for (int round = 0; round < tv.ntiles; round++) {
    __syncthreads();
    if (threadIdx.x == 0) *sh_t = atomicAdd(counter, 1);
    __syncthreads();
    const int t = *sh_t;
    __syncthreads();
    if (t >= tv.ntiles) break;
    const int i = tile_order[2 * t], j = tile_order[2 * t + 1];
    T* Cg = tiles + (size_t)t * T2;
    acc_t acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int vv = 0; vv < 4; vv++) {
          const int r = wr * 32 + bi * 16 + Mfma<T>::row(vv, lane);
          const int c = wc * 32 + bj * 16 + (lane & 15);
          acc[bi][bj][vv] = Cg[c * TB + r];
        }
    for (int k = 0; k < j; k++) {
      const long long tik = tile_index(i, k, NT), tjk = tile_index(j, k, NT);
      if (threadIdx.x == 0) {
        wait_flag(ready + tik, ctl);
        if (i != j) wait_flag(ready + tjk, ctl);
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const T* Ag = tiles + (size_t)tik * T2;
      const T* Bg = tiles + (size_t)tjk * T2;
      for (int kk = 0; kk < TB; kk += KC) {
        for (int e = threadIdx.x; e < TB * KC; e += 256) {
          As[e] = Ag[(size_t)kk * TB + e];
          Bs[e] = Bg[(size_t)kk * TB + e];
        }
        __syncthreads();
        // C[r][c] -= sum_m A[r][kk+m] B[c][kk+m]
#pragma unroll 4
        for (int m = 0; m < KC; m += 4) {
          const int mm = m + (lane >> 4);
          T a[2], b[2];
#pragma unroll
          for (int bi = 0; bi < 2; bi++) a[bi] = -As[mm * TB + wr * 32 + bi * 16 + (lane & 15)];
#pragma unroll
          for (int bj = 0; bj < 2; bj++) b[bj] = Bs[mm * TB + wc * 32 + bj * 16 + (lane & 15)];
#pragma unroll
          for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < 2; bj++) acc[bi][bj] = Mfma<T>::fma(a[bi], b[bj], acc[bi][bj]);
        }
        __syncthreads();
      }
    }
    // accumulators -> Ct (column-major)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int vv = 0; vv < 4; vv++) {
          const int r = wr * 32 + bi * 16 + Mfma<T>::row(vv, lane);
          const int c = wc * 32 + bj * 16 + (lane & 15);
          Ct[c * TB + r] = acc[bi][bj][vv];
        }
    __syncthreads();
    if (i == j) {
      if constexpr (sizeof(T) == 8) {
        potrf_lds<TB>((double*)Ct, j * TB, tv.n, ctl);
      } else {
        for (int e = threadIdx.x; e < TB * TB; e += 256) Dd[e] = (double)Ct[e];
        __syncthreads();
        potrf_lds<TB>(Dd, j * TB, tv.n, ctl);
        for (int e = threadIdx.x; e < TB * TB; e += 256) Ct[e] = (T)Dd[e];
        __syncthreads();
      }
    } else {
      const long long tjj = tile_index(j, j, NT);
      if (threadIdx.x == 0) wait_flag(ready + tjj, ctl);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const T* Lg = tiles + (size_t)tjj * T2;
      for (int e = threadIdx.x; e < TB * TB; e += 256) stage[e] = Lg[e];
      __syncthreads();
      trsm_lds<T, TB>(Ct, stage);
    }
    for (int e = threadIdx.x; e < TB * TB; e += 256) Cg[e] = Ct[e];
    publish_flag(ready + t);
  }
}

// Backward solve L^T x = y, y = row n of the factor (the forward solve's
// result), persistent over tile rows in reverse order.
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_trsv_bwd(TileView tv, const T* __restrict__ tiles, double* xU, int* ready,
                                                  int* counter, DevCtl* ctl, int gate) {
  if (!chol_gate(ctl, gate)) return;
  __shared__ T Ls[TB * TB];
  __shared__ double acc[TB], xk[TB], part[256 / TB][TB];
  __shared__ int sh_t;
  constexpr int NS = 256 / TB;
  const int n = tv.n, NT = tv.NT;
  const int NTr = (n + TB - 1) / TB;
  const size_t T2 = (size_t)TB * TB;
  const int tid = threadIdx.x;
  // a workgroup takes at most NTr tile rows: the loop is bounded
  for (int round = 0; round < NTr; round++) {
    if (tid == 0) sh_t = atomicAdd(counter, 1);
    __syncthreads();
    const int t = sh_t;
    __syncthreads();
    if (t >= NTr) break;
    const int i = NTr - 1 - t;
    const T* ytile = tiles + (size_t)tile_index(n / TB, i, NT) * T2;
    if (tid < TB) {
      const int R = i * TB + tid;
      acc[tid] = (R < n) ? (double)ytile[tid * TB + (n % TB)] : 0.0;
    }
    for (int k = NTr - 1; k > i; k--) {
      if (tid == 0) wait_flag(ready + k, ctl);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (tid < TB) xk[tid] = (k * TB + tid < n) ? xU[k * TB + tid] : 0.0;
      const T* Lg = tiles + (size_t)tile_index(k, i, NT) * T2;
      for (int e = tid; e < TB * TB; e += 256) Ls[e] = Lg[e];
      __syncthreads();
      {
        const int c = tid % TB, rs = tid / TB;
        double s = 0.0;
        for (int r = rs * (TB / NS); r < (rs + 1) * (TB / NS); r++) s += (double)Ls[c * TB + r] * xk[r];
        part[rs][c] = s;
      }
      __syncthreads();
      if (tid < TB) {
        double s = 0.0;
        for (int q = 0; q < NS; q++) s += part[q][tid];
        acc[tid] -= s;
      }
      __syncthreads();
    }
    const T* Dg = tiles + (size_t)tile_index(i, i, NT) * T2;
    for (int e = tid; e < TB * TB; e += 256) Ls[e] = Dg[e];
    __syncthreads();
    // one wave: lane c holds acc[c]; columns from the last to the first
    if (tid < 64) {
      double a = (tid < TB) ? acc[tid] : 0.0;
      for (int c = TB - 1; c >= 0; c--) {
        const int R = i * TB + c;
        double xc = 0.0;
        if (R < n) xc = __shfl(a, c) / (double)Ls[c * TB + c];
        if (tid == c) a = xc;
        if (tid < c) a -= (double)Ls[tid * TB + c] * xc;
      }
      if (tid < TB) acc[tid] = a;
    }
    __syncthreads();
    if (tid < TB) {
      const int R = i * TB + tid;
      if (R < n) xU[R] = acc[tid];
    }
    publish_flag(ready + i);
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
void launch_chol(const TileView& tv, T* tiles, const int* tile_order, int* ready, int* counter, DevCtl* ctl, int gate,
                 int max_wg, hipStream_t st) {
  int g = tv.ntiles < max_wg ? tv.ntiles : max_wg;
  hipLaunchKernelGGL((k_chol<T, 64>), dim3(g), dim3(256), 0, st, tv, tiles, tile_order, ready, counter, ctl, gate);
}

template <typename T>
void launch_trsv_bwd(const TileView& tv, const T* tiles, double* xU, int* ready, int* counter, DevCtl* ctl, int gate,
                     int max_wg, hipStream_t st) {
  const int NTr = (tv.n + 63) / 64;
  int g = NTr < max_wg ? NTr : max_wg;
  hipLaunchKernelGGL((k_trsv_bwd<T, 64>), dim3(g), dim3(256), 0, st, tv, tiles, xU, ready, counter, ctl, gate);
}

template void launch_chol_reset<double>(int*, int, int*, DevCtl*, int, hipStream_t);
template void launch_chol_reset<float>(int*, int, int*, DevCtl*, int, hipStream_t);
template void launch_chol<double>(const TileView&, double*, const int*, int*, int*, DevCtl*, int, int, hipStream_t);
template void launch_chol<float>(const TileView&, float*, const int*, int*, int*, DevCtl*, int, int, hipStream_t);
template void launch_trsv_bwd<double>(const TileView&, const double*, double*, int*, int*, DevCtl*, int, int,
                                      hipStream_t);
template void launch_trsv_bwd<float>(const TileView&, const float*, double*, int*, int*, DevCtl*, int, int,
                                     hipStream_t);

}  // namespace nsn
