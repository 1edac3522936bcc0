// Building blocks of the 64x64 tile Cholesky on gfx950, shared by the
// single-problem factorisation (chol_kernels.hip) and the batched-window
// solver (batch_kernels.hip): the 16x16x4 MFMA operand / result layouts, the
// FP64 diagonal-tile POTRF and inverse in LDS, MFMA tile products and the
// tile GEMV of the triangular solves.
#pragma once

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

// orders a wave's LDS write before its later reads of other lanes' entries
// (wavefront-scope fences: no instructions, since one wave's LDS operations
// execute in order, but the compiler keeps the order)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
// lane and broadcasts the pivot column (lane reads and LDS); between panels the
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

// 1/sqrt(d) from the hardware estimate y refined by one third-order step,
// y (1 + e/2 + 3 e^2 / 8) with e = 1 - d y^2 (four dependent operations; two
// Newton steps take seven)
__device__ __forceinline__ double rsqrt_f64(double d) {
  const double y = __builtin_amdgcn_rsq(d);
  const double e = __builtin_fma(-(d * y), y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// right-looking Cholesky of columns C0 .. C0+15 of the 64-row FP64 tile A
// (rows >= C0), wave 0 only.  Rows with global index >= n (the bS row and the
// identity padding) take the reciprocal 1 and never count as failures; a
// failed pivot (not positive and finite) also takes 1 and sets `bad`.  Each
// column is scaled by its reciprocal square root r on every lane, so the
// diagonal lane keeps d r = sqrt(d) for a good pivot and d itself for a
// padded or failed one (nothing reads the diagonal of a factor tile except
// potrf_lds's check of failed pivots; the solves use the reciprocals, stored
// in the upper triangle).  The step is branch-free: every
// lane computes the same pivot, the reciprocal square root of the selected
// value, and the updates run on every lane (entries above the diagonal in the
// registers are garbage and never written back).  After scaling, the entry
// of column k that the next pivot needs comes by a lane read; the rest of the
// column goes through the LDS buffer col (two 64-double halves, alternating
// by step) and every lane reads rows C0+k+2 .. C0+15 back as a broadcast, so
// those updates overlap the next pivot and their operands land in vector
// registers.  The pivot reciprocals go to their slots in the upper triangle
// after the panel.
template <int TB, int C0>
__device__ __forceinline__ void potrf_panel(double* A, double* col, int g0, int n, int& bad) {
  // the lane index is opaque to the optimiser, so that the lane tests below
  // are formed where they are used: hoisted out of the enclosing loops they
  // become scalar masks that the register allocator spills to vector lanes
  // and reads back on the pivot path (measured 3.06 -> 2.92 ms for the
  // 5994-row factorisation)
  int lane = threadIdx.x;
  asm volatile("" : "+v"(lane));
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
    a[k] *= r;
    myr = (lane == C0 + k) ? r : myr;
    if (k + 1 < kPW) {
      // the next pivot's column first, by a lane read (the critical path)
      a[k + 1] -= a[k] * readlane_f64(a[k], C0 + k + 1);
      if (k + 2 < kPW) {
        // the rest through LDS: the write and the broadcast reads are LDS
        // instructions of this one wave, which the LDS executes in order
        double* cb = col + (k & 1) * 64;
        cb[lane] = a[k];
        wave_sync();
#pragma unroll
        for (int c = k + 2; c < kPW; c++) a[c] -= a[k] * cb[C0 + c];
      }
    }
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
__device__ __forceinline__ void potrf_next(double* A, double* col, int g0, int n, int& bad) {
  const int wv = threadIdx.x >> 6;
  if (wv < 3 - P) block_update<TB>(A, P + 1 + wv, P + 1, kPW * P);
  __syncthreads();
  if (wv == 0) {
    potrf_panel<TB, kPW * (P + 1)>(A, col, g0, n, bad);
    // the last panel leaves the tile's failure flag for every thread
    if (P == 2 && threadIdx.x == 0) ((volatile int*)(col + 128))[0] = bad;
  } else if (P == 0) {
    // blocks (2,2), (3,2), (3,3) with panel 0
    const int rb = wv == 3 ? 3 : wv + 1, cb = wv == 1 ? 2 : (wv == 2 ? 2 : 3);
    block_update<TB>(A, rb, cb, 0);
  } else if (P == 1 && wv == 1) {
    block_update<TB>(A, 3, 3, kPW);
  }
  __syncthreads();
}

// col: 129 doubles of LDS that nothing else uses during the call (the
// broadcast buffer and the failure flag).  When a pivot failed, the tile is
// checked column by column (all threads, rare): a pivot that is exactly
// zero (its diagonal entry was left at d = 0) with an exactly zero column
// below it is the semidefinite case that the reference's pivoted LDL^T
// accepts (SolverCholeskyDense.hpp:15-18) and solves with the pseudo-inverse
// of D: its stored reciprocal becomes 0, so the inverse of the factor tile
// has that row and column zero and the solution entry comes out 0, as in the
// reference (a row and column of S that no term touches: a variable camera
// without observations, undamped).  Any other failed pivot, or a zero pivot
// over a non-zero column, is a failure, as there.
template <int TB>
__device__ __forceinline__ void potrf_lds(double* A, double* col, int g0, int n, DevCtl* ctl) {
  static_assert(TB == 4 * kPW, "four panels");
  int bad = 0;
  if (threadIdx.x < 64) potrf_panel<TB, 0>(A, col, g0, n, bad);
  __syncthreads();
  potrf_next<TB, 0>(A, col, g0, n, bad);
  potrf_next<TB, 1>(A, col, g0, n, bad);
  potrf_next<TB, 2>(A, col, g0, n, bad);
  if (((volatile int*)(col + 128))[0] != 0) {  // uniform
    volatile int* fails = (volatile int*)(col + 128) + 1;
    if (threadIdx.x == 0) *fails = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
      const int c = e / TB, r = e % TB;
      const double v = A[c * TB + c];
      if (g0 + c >= n || (v > 0.0 && v < __builtin_inf())) continue;  // padded or good
      if (v != 0.0 || (r > c && A[c * TB + r] != 0.0)) *fails = 1;
    }
    __syncthreads();
    if (threadIdx.x < TB && g0 + threadIdx.x < n && A[threadIdx.x * TB + threadIdx.x] == 0.0)
      A[rdiag_slot<TB>(threadIdx.x)] = 0.0;
    bad = *fails;
    __syncthreads();
  }
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
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // lane tests formed in place (see potrf_panel)
  const int lane = tid & 63, wv = threadIdx.x >> 6;
  if (lane < 16) {
    // branch-free: for i < j every term is zero, so x_i = 0 without a test,
    // and the loads of L (the same for every lane) can all be issued early
    const int b0 = wv * 16, j = lane;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; k++) s -= L[(b0 + k) * TB + b0 + i] * x[k];
      x[i] = s * L[rdiag_slot<TB>(b0 + i)];
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

// acc += tile (global, column-major TB x TB)
template <typename T, int TB>
__device__ __forceinline__ void acc_add(CholAcc<T>& acc, const T* Pg, int lane, int wr, int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int vv = 0; vv < 4; vv++) {
        const int r = wr * 32 + bi * 16 + Mfma<T>::row(vv, lane);
        const int c = wc * 32 + bj * 16 + (lane & 15);
        acc.a[bi][bj][vv] += Pg[c * TB + r];
      }
}

template <typename T>
__device__ __forceinline__ void acc_clear(CholAcc<T>& acc) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int vv = 0; vv < 4; vv++) acc.a[bi][bj][vv] = 0;
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

// X = A B^T for TB x TB column-major operands with B lower triangular
// (B[c][m] = 0 for m > c, as the inverse of a factor tile): wave w forms the
// 16-row block w of X for the four 16-column blocks, and column block cb
// accumulates only m < 16 (cb + 1), 40 MFMAs per wave instead of 64.  The
// terms left out are exact zeros, so X is the same as the full product's.
// A in LDS (element type T), B in LDS or global (element type U).
template <typename T>
struct TriProd {
  typename Mfma<T>::acc_t a[4];
};

template <typename T, typename U, int TB>
__device__ __forceinline__ void tri_gemm_nt(TriProd<T>& x, const T* A, const U* B, int lane, int wv) {
  static_assert(TB == 64, "four 16-column blocks");
  const T* As = A + wv * 16 + (lane & 15);
  const U* Bs = B + (lane & 15);
  const int kq = lane >> 4;
#pragma unroll
  for (int cb = 0; cb < 4; cb++) x.a[cb] = typename Mfma<T>::acc_t{};
#pragma unroll
  for (int m = 0; m < TB; m += 4) {
    const int off = (m + kq) * TB;
    const T a = As[off];
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
      if (m < 16 * (cb + 1)) x.a[cb] = Mfma<T>::fma(a, (T)Bs[off + 16 * cb], x.a[cb]);
  }
}

// TriProd -> LDS tile (column-major)
template <typename T, int TB>
__device__ __forceinline__ void tri_store(const TriProd<T>& x, T* D, int lane, int wv) {
#pragma unroll
  for (int cb = 0; cb < 4; cb++)
#pragma unroll
    for (int vv = 0; vv < 4; vv++) D[(16 * cb + (lane & 15)) * TB + 16 * wv + Mfma<T>::row(vv, lane)] = x.a[cb][vv];
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

// 64 x 64 column-major tile (entry (r, c) at A[c * TB + r]) times a
// 64-vector in LDS with FP64 accumulation; the tile is loaded into registers
// (16 entries per thread, as pairs of adjacent entries of a column) before
// the vector is known.  Wave w owns 16 outputs and reduces them within the
// wave (three lane exchanges), so no barrier separates the partial sums from
// the result:
//   TRANS:  out[c] = sum_r A(r, c) x[r] for c = 16w + l/8 and c + 8; lane l
//           holds the rows 2(l%8) + 16q, +1, q = 0..3, of both columns;
//   !TRANS: out[r] = sum_c A(r, c) x[c] for r in 16w .. 16w+15; lane l holds
//           the row pair 16w + 2(l%8), +1 in the columns l/8 + 8q, q = 0..7.
// In both, eight lanes read 16 adjacent entries of a column per load.
// reduce() leaves the two sums of a lane group in every lane of the group;
// the owner lanes (TRANS: l%8 == 0; otherwise l < 8) hand out outputs
// index(0) and index(1).  All sums run in a fixed order.
template <typename T, int TB, bool TRANS>
struct TileGemv {
  static_assert(TB == 64, "64-row tiles");
  typedef T T2 __attribute__((ext_vector_type(2)));
  T a[16];
  static __device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63); }
  static __device__ __forceinline__ int wave() { return (int)(threadIdx.x >> 6); }
  __device__ __forceinline__ void load(const T* A) {
    const int l = lane(), w = wave();
    if (TRANS) {
      const T* p = A + (size_t)(16 * w + (l >> 3)) * TB + 2 * (l & 7);
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const T2 v = *(const T2*)(p + (size_t)8 * h * TB + 16 * q);
          a[8 * h + 2 * q] = v.x;
          a[8 * h + 2 * q + 1] = v.y;
        }
    } else {
      const T* p = A + (size_t)(l >> 3) * TB + 16 * w + 2 * (l & 7);
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const T2 v = *(const T2*)(p + (size_t)8 * q * TB);
        a[2 * q] = v.x;
        a[2 * q + 1] = v.y;
      }
    }
  }
  // adds this lane's partial sums of its two outputs with x (LDS, 16-byte
  // aligned) to s0, s1
  __device__ __forceinline__ void partial(const double* x, double& s0, double& s1) const {
    const int l = lane();
    if (TRANS) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double2 xv = *(const double2*)(x + 2 * (l & 7) + 16 * q);
        s0 += (double)a[2 * q] * xv.x + (double)a[2 * q + 1] * xv.y;
        s1 += (double)a[8 + 2 * q] * xv.x + (double)a[8 + 2 * q + 1] * xv.y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const double xv = x[(l >> 3) + 8 * q];
        s0 += (double)a[2 * q] * xv;
        s1 += (double)a[2 * q + 1] * xv;
      }
    }
  }
  static __device__ __forceinline__ void reduce(double& s0, double& s1) {
#pragma unroll
    for (int m = TRANS ? 1 : 8; m <= (TRANS ? 4 : 32); m *= 2) {
      s0 += __shfl_xor(s0, m);
      s1 += __shfl_xor(s1, m);
    }
  }
  static __device__ __forceinline__ bool owner() { return TRANS ? (lane() & 7) == 0 : lane() < 8; }
  static __device__ __forceinline__ int index(int e) {
    return TRANS ? 16 * wave() + (lane() >> 3) + 8 * e : 16 * wave() + 2 * lane() + e;
  }
};

}  // namespace nsn
