// Batched independent BA windows: one persistent launch in which each
// workgroup takes windows (largest first) from an atomic counter and runs the complete
// Levenberg-Marquardt solve of a window (nelson/LevenbergMarquardt.hpp:35-146
// with the Schur step of nelson/SolverCholeskySchur.hpp:91-157) on its own:
// linearisation, V^-1 / W V^-1, S assembly from the (W, W) pair lists, a
// dense Cholesky of S (n <= 256, padded to a multiple of 64) with FP64 pivot
// blocks, the two triangular solves, back-substitution, oplus, cost and the
// accept / reject schedule.  Two factorisation variants, picked per batch by
// the autotuner: 64x64 tiles (the POTRF / inverse / MFMA tile routines of the
// large-problem factorisation, solves as products with the diagonal-tile
// inverses) or NB-column panels factorised in LDS with column sweeps.
// Windows never wait for each other, so the launch needs no inter-workgroup
// synchronisation; every loop is bounded.
// The per-edge, per-point and LM-control arithmetic is shared with the
// single-problem kernels (ba_device.h).
#include <hip/hip_runtime.h>

#include "ba_device.h"
#include "batch.h"
#include "tile_device.h"

namespace nsn {

namespace {

constexpr int kBT = 256;      // threads per window
constexpr int kMaxS = 256;    // largest S order of a window

// deterministic block reductions of per-thread values (4 waves, fixed order)
__device__ __forceinline__ double bsum(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__device__ __forceinline__ double bmax(double v, double* sh) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

template <typename T>
__device__ __forceinline__ T* tp(void* p) {
  return (T*)p;
}

// update(true): V, bV, W (point-major), U, bU (one wave per camera), chi2,
// max |b| and max |diag H|
template <typename T>
__device__ __forceinline__ void win_linearize(const BWin& w, DevCtl* c, double* sh) {
  const BAView v = w.v;  // a copy: the stores below cannot alias it
  T* V = tp<T>(w.V);
  T* bV = tp<T>(w.bV);
  T* W = tp<T>(w.W);
  T* U = tp<T>(w.U);
  T* bU = tp<T>(w.bU);
  double chi2 = 0.0, mb = 0.0, md = 0.0;
  for (int p = threadIdx.x; p < v.P; p += kBT) lin_point<T, true>(v, p, V, bV, W, chi2, mb, md);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int u = wv; u < v.NU; u += 4) {
    const int cam = v.var_cam_user[u];
    double R[kCamRec];
    load_cam(v, cam, R);
    T acc[27];
#pragma unroll
    for (int i = 0; i < 27; i++) acc[i] = (T)0;
    const int e1 = v.cam_ptr[u + 1];
    for (int e = v.cam_ptr[u] + lane; e < e1; e += 64) {
      const int q = v.ce_pt[e];
      const Proj pr = project(v, R, v.pts[3 * q], v.pts[3 * q + 1], v.pts[3 * q + 2], v.ce_obs[2 * e],
                              v.ce_obs[2 * e + 1], v.ce_w[e]);
      T j0[6], j1[6];
      jac_cam<T>(v, pr, j0, j1);
      const T om = (T)pr.om, r0 = (T)pr.r0, r1 = (T)pr.r1;
#pragma unroll
      for (int i = 0; i < 6; i++) {
#pragma unroll
        for (int j = i; j < 6; j++) acc[up6(i, j)] += om * (j0[i] * j0[j] + j1[i] * j1[j]);
        acc[21 + i] += om * (j0[i] * r0 + j1[i] * r1);
      }
    }
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const double s = wave_sum((double)acc[i]);
      if (lane == 0) {
        if (i < 21) U[21 * (size_t)u + i] = (T)s;
        else bU[6 * (size_t)u + (i - 21)] = (T)s;
      }
      const double sr = (double)(T)s;
      if (i >= 21) mb = fmax(mb, fabs(sr));
      else if (i == 0 || i == 6 || i == 11 || i == 15 || i == 18 || i == 20) md = fmax(md, fabs(sr));
    }
  }
  const double tot = bsum(chi2, sh);
  const double mbt = bmax(mb, sh);
  const double mdt = bmax(md, sh);
  if (threadIdx.x == 0) {
    c->chi2 = tot;
    c->chi2_lin = tot;
    c->maxb = mbt;
    c->maxdiag = mdt;
  }
}

// update(false): chi2 only
template <typename T>
__device__ __forceinline__ void win_cost(const BWin& w, DevCtl* c, double* sh) {
  double chi2 = 0.0, mb = 0.0, md = 0.0;
  const BAView v = w.v;
  for (int p = threadIdx.x; p < v.P; p += kBT)
    lin_point<T, false>(v, p, nullptr, nullptr, nullptr, chi2, mb, md);
  const double tot = bsum(chi2, sh);
  if (threadIdx.x == 0) c->chi2 = tot;
}

__device__ __forceinline__ void win_copy(const BAView& v, double* cd, double* pd, const double* cs, const double* ps) {
  for (int i = threadIdx.x; i < 12 * v.C; i += kBT) cd[i] = cs[i];
  for (int i = threadIdx.x; i < 3 * v.P; i += kBT) pd[i] = ps[i];
}

// panel variant: Cholesky of the n x n lower S (column-major, leading
// dimension ld) in place; panels of NB columns are factorised in FP64 in LDS
// (pan: NB columns of kMaxS rows), the trailing matrix is updated from the
// panel; false if a pivot is not positive
template <typename T, int NB>
__device__ __forceinline__ bool win_cholesky_panels(T* S, int n, int ld, double* pan) {
  const int tid = threadIdx.x;
  bool ok = true;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = min(NB, n - k0), m = n - k0;
    for (int e = tid; e < nb * m; e += kBT) {
      const int cc = e / m, r = e % m;
      pan[cc * kMaxS + r] = (double)S[(size_t)(k0 + cc) * ld + k0 + r];
    }
    __syncthreads();
    for (int cc = 0; cc < nb; cc++) {
      const double d = pan[cc * kMaxS + cc];
      const bool good = d > 0.0 && isfinite(d);
      ok = ok && good;
      const double piv = good ? sqrt(d) : 1.0;
      const double rp = 1.0 / piv;
      __syncthreads();
      for (int r = cc + tid; r < m; r += kBT) pan[cc * kMaxS + r] = (r == cc) ? piv : pan[cc * kMaxS + r] * rp;
      __syncthreads();
      const int rem = nb - cc - 1;
      for (int e = tid; e < rem * m; e += kBT) {
        const int j = cc + 1 + e / m, r = e % m;
        if (r >= j) pan[j * kMaxS + r] -= pan[cc * kMaxS + r] * pan[cc * kMaxS + j];
      }
      __syncthreads();
    }
    for (int e = tid; e < nb * m; e += kBT) {
      const int cc = e / m, r = e % m;
      if (r >= cc) S[(size_t)(k0 + cc) * ld + k0 + r] = (T)pan[cc * kMaxS + r];
    }
    const int mt = m - nb;
    for (int e = tid; e < mt * mt; e += kBT) {
      const int jj = e / mt, ii = e % mt;
      if (ii >= jj) {
        double s = 0.0;
#pragma unroll 4
        for (int cc = 0; cc < nb; cc++) s += pan[cc * kMaxS + nb + ii] * pan[cc * kMaxS + nb + jj];
        T* a = S + (size_t)(k0 + nb + jj) * ld + k0 + nb + ii;
        *a = (T)((double)*a - s);
      }
    }
    __syncthreads();
  }
  return ok;
}

// column sweeps L y = b, then L^T x = y, on the panel variant's factor; b in
// yv (LDS), x to xU
template <typename T>
__device__ __forceinline__ void win_solve_sweeps(const T* S, int n, int ld, double* yv, double* ys, double* xU) {
  const int tid = threadIdx.x;
  for (int cc = 0; cc < n; cc++) {
    const double yc = yv[cc] / (double)S[(size_t)cc * ld + cc];
    if (tid == 0) ys[cc] = yc;
    if (tid > cc && tid < n) yv[tid] -= (double)S[(size_t)cc * ld + tid] * yc;
    __syncthreads();
  }
  for (int cc = n - 1; cc >= 0; cc--) {
    const double xc = ys[cc] / (double)S[(size_t)cc * ld + cc];
    if (tid == 0) xU[cc] = xc;
    if (tid < cc) ys[tid] -= (double)S[(size_t)tid * ld + cc] * xc;
    __syncthreads();
  }
}

// ---- tile variant -----------------------------------------------------------
// 64x64 tile (ti, tj) of the padded S (column-major, leading dimension ld)
template <typename T>
__device__ __forceinline__ T* tile_at(T* S, int ld, int ti, int tj) {
  return S + (size_t)(64 * tj) * ld + 64 * ti;
}

template <typename T, typename U>
__device__ __forceinline__ void tile_to_lds(U* R, const T* A, int ld) {
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = threadIdx.x + 256 * q;
    v[q] = A[(size_t)(e >> 6) * ld + (e & 63)];
  }
#pragma unroll
  for (int q = 0; q < 16; q++) R[threadIdx.x + 256 * q] = (U)v[q];
}

template <typename T, typename U>
__device__ __forceinline__ void lds_to_tile(T* A, int ld, const U* R) {
  U v[16];
#pragma unroll
  for (int q = 0; q < 16; q++) v[q] = R[threadIdx.x + 256 * q];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = threadIdx.x + 256 * q;
    A[(size_t)(e >> 6) * ld + (e & 63)] = (T)v[q];
  }
}

// acc (+)= A B^T for 64x64 operands in global memory (leading dimension ld)
template <typename T>
__device__ __forceinline__ void acc_gemm_nt_g(CholAcc<T>& acc, const T* A, const T* B, int ld, bool sub, int lane,
                                              int wr, int wc) {
  const T* As = A + wr * 32 + (lane & 15);
  const T* Bs = B + wc * 32 + (lane & 15);
  const int kq = lane >> 4;
#pragma unroll 8
  for (int m = 0; m < 64; m += 4) {
    const size_t off = (size_t)(m + kq) * ld;
    const T a0 = sub ? -As[off] : As[off], a1 = sub ? -As[off + 16] : As[off + 16];
    const T b0 = Bs[off], b1 = Bs[off + 16];
    acc.a[0][0] = Mfma<T>::fma(a0, b0, acc.a[0][0]);
    acc.a[0][1] = Mfma<T>::fma(a0, b1, acc.a[0][1]);
    acc.a[1][0] = Mfma<T>::fma(a1, b0, acc.a[1][0]);
    acc.a[1][1] = Mfma<T>::fma(a1, b1, acc.a[1][1]);
  }
}

template <typename T>
__device__ __forceinline__ void acc_zero(CholAcc<T>& acc) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) acc.a[bi][bj] = typename Mfma<T>::acc_t{};
}

// right-looking tile Cholesky of the padded S (nP x nP, NT = nP / 64 <= 4
// tiles per side): per tile column j the diagonal tile is factorised and
// inverted in FP64 in LDS (potrf_lds / tri_inverse), its inverse kept in
// Linv[j], the tiles below become C L_jj^-T and the trailing tiles are
// updated, all products on MFMA.  Pivot failures are counted in c->fail_s.
template <typename T>
__device__ __forceinline__ void win_cholesky_tiles(const BWin& w, DevCtl* c, double* lds) {
  double* R0d = lds;
  double* R1d = lds + 4096;
  T* R0 = (T*)lds;
  T* S = tp<T>(w.S);
  T* Li = tp<T>(w.Linv);
  const int ld = w.nP, NTw = w.nP / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv & 1, wc = wv >> 1;
  for (int j = 0; j < NTw; j++) {
    tile_to_lds<T, double>(R0d, tile_at(S, ld, j, j), ld);
    __syncthreads();
    potrf_lds<64>(R0d, R1d, 64 * j, w.n, c);  // R1d is free until the inverse
    tri_inverse<64>(R0d, R1d);
    lds_to_tile<T, double>(tile_at(S, ld, j, j), ld, R0d);
    {
      T* Ig = Li + (size_t)j * 4096;
#pragma unroll 4
      for (int e = threadIdx.x; e < 4096; e += 256) Ig[e] = (T)R1d[e];
    }
    __syncthreads();
    for (int i = j + 1; i < NTw; i++) {
      tile_to_lds<T, T>(R0, tile_at(S, ld, i, j), ld);
      __syncthreads();
      CholAcc<T> x;
      acc_zero(x);
      acc_gemm_nt<T, double, 64>(x, R0, R1d, false, lane, wr, wc);
      __syncthreads();
      acc_store<T, T, 64>(x, R0, lane, wr, wc);
      __syncthreads();
      lds_to_tile<T, T>(tile_at(S, ld, i, j), ld, R0);
    }
    __syncthreads();
    for (int k = j + 1; k < NTw; k++)
      for (int i = k; i < NTw; i++) {
        CholAcc<T> acc;
        T* Cg = tile_at(S, ld, i, k);
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
          for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int vv = 0; vv < 4; vv++)
              acc.a[bi][bj][vv] = Cg[(size_t)(wc * 32 + bj * 16 + (lane & 15)) * ld + wr * 32 + bi * 16 +
                                     Mfma<T>::row(vv, lane)];
        acc_gemm_nt_g<T>(acc, tile_at(S, ld, i, j), tile_at(S, ld, k, j), ld, true, lane, wr, wc);
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
          for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int vv = 0; vv < 4; vv++)
              Cg[(size_t)(wc * 32 + bj * 16 + (lane & 15)) * ld + wr * 32 + bi * 16 + Mfma<T>::row(vv, lane)] =
                  acc.a[bi][bj][vv];
      }
    __syncthreads();
  }
}

// out[0..63] = M v (TRANS = false) or M^T v (TRANS = true) for a 64x64
// column-major tile M (leading dimension ld) and v in LDS; four partial sums
// per entry combined in a fixed order through red (4 x 64 doubles); the
// result is valid in threads 0..63 (returned) after the call
template <typename T, bool TRANS>
__device__ __forceinline__ double tile_gemv(const T* M, int ld, const double* v, double* red) {
  const int tid = threadIdx.x, e = tid & 63, part = tid >> 6;
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int k = part * 16 + q;
    s += (double)(TRANS ? M[(size_t)e * ld + k] : M[(size_t)k * ld + e]) * v[k];
  }
  red[part * 64 + e] = s;
  __syncthreads();
  double r = 0.0;
  if (tid < 64) r = ((red[e] + red[64 + e]) + red[128 + e]) + red[192 + e];
  __syncthreads();
  return r;
}

// L y = b then L^T x = y by 64-row blocks with the diagonal-tile inverses:
// y_k = Linv_kk (b_k - sum_{i<k} L_ki y_i), x_k = Linv_kk^T (y_k - sum_{i>k}
// L_ik^T x_i).  b (global, nP entries, zero beyond n) and x (n entries
// written) may be the same array.  ws: LDS workspace (nP + 64 + 256 doubles)
template <typename T>
__device__ __forceinline__ void win_solve_tiles(const BWin& w, double* ws, const double* b, double* x) {
  const T* S = tp<T>(w.S);
  const T* Li = tp<T>(w.Linv);
  const int ld = w.nP, NTw = w.nP / 64, tid = threadIdx.x;
  double* yv = ws;                 // nP: y, then x
  double* tv = ws + w.nP;          // 64
  double* red = tv + 64;           // 256
  for (int e = tid; e < w.nP; e += 256) yv[e] = b[e];
  __syncthreads();
  for (int k = 0; k < NTw; k++) {
    double acc = 0.0;
    for (int i = 0; i < k; i++) acc += tile_gemv<T, false>(tile_at(S, ld, k, i), ld, yv + 64 * i, red);
    if (tid < 64) tv[tid] = yv[64 * k + tid] - acc;
    __syncthreads();
    const double y = tile_gemv<T, false>(Li + (size_t)k * 4096, 64, tv, red);
    if (tid < 64) yv[64 * k + tid] = y;
    __syncthreads();
  }
  for (int k = NTw - 1; k >= 0; k--) {
    double acc = 0.0;
    for (int i = k + 1; i < NTw; i++) acc += tile_gemv<T, true>(tile_at(S, ld, i, k), ld, yv + 64 * i, red);
    if (tid < 64) tv[tid] = yv[64 * k + tid] - acc;
    __syncthreads();
    const double x = tile_gemv<T, true>(Li + (size_t)k * 4096, 64, tv, red);
    if (tid < 64) yv[64 * k + tid] = x;
    __syncthreads();
  }
  for (int e = tid; e < w.n; e += 256) x[e] = yv[e];
  __syncthreads();
}

// one Schur step with damping (rel, abs) = (c->rel, c->abs_); writes xU, xV
// and the step statistics (the small-increment test of
// LevenbergMarquardt.hpp:73-75 included).  ALG: 8 / 16 = panel variant of
// that width, 64 = tile variant.  lds: the workgroup's 64 KiB; sh: 4 doubles.
// the 18 values of a 6x3 block (Y or W) of edge k; blocks start at a
// multiple of 18 elements of a 256-byte aligned array, so 8-byte (float) or
// 16-byte (double) vector loads are aligned
__device__ __forceinline__ void load18(const float* p, float* o) {
  const float2* q = (const float2*)p;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const float2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}

__device__ __forceinline__ void load18(const double* p, double* o) {
  const double2* q = (const double2*)p;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const double2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}

// phase 1 of the S assembly: a lane sums the terms Z_a Z_b^T (Z = W L^-T,
// V = L L^T, so Z_a Z_b^T = W_a V^-1 W_b^T, and one array serves both
// operands: half the bytes of Y and W, a diagonal term reads one block) of one chunk
// (at most kSChunk terms of one block) into 36 partial sums, and for a
// diagonal block also the (W V^-1 bV)_a of the bS row of its camera into 6
// more, stored under the chunk's number; the terms are taken U at a time (all
// their loads issued before the products), the last group padded with repeats
// of its first term that are not added
template <typename T>
__device__ __forceinline__ void win_s_chunks(const BWin& w) {
  constexpr int U = sizeof(T) == 4 ? 4 : 2;
  const T* Y = tp<T>(w.Y);  // Z
  const T* tW = tp<T>(w.tW);
  T* part = tp<T>(w.part);
  const int nch = w.nch;
  const int4* chs = (const int4*)w.ch;
  const int2* terms = (const int2*)w.terms;
  for (int c = threadIdx.x; c < nch; c += kBT) {
    const int4 cr = chs[c];
    T acc[36];
#pragma unroll
    for (int e = 0; e < 36; e++) acc[e] = (T)0;
    for (int t = cr.x; t < cr.y; t += U) {
      int2 k[U];
#pragma unroll
      for (int u = 0; u < U; u++) k[u] = terms[t + u < cr.y ? t + u : t];
      T y[U][18], ww[U][18];
#pragma unroll
      for (int u = 0; u < U; u++) {
        load18(Y + 18 * (size_t)k[u].x, y[u]);
        load18(Y + 18 * (size_t)k[u].y, ww[u]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (t + u >= cr.y) break;
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
          for (int bb = 0; bb < 6; bb++)
            acc[6 * a + bb] += y[u][3 * a] * ww[u][3 * bb] + y[u][3 * a + 1] * ww[u][3 * bb + 1] +
                               y[u][3 * a + 2] * ww[u][3 * bb + 2];
      }
    }
    T* o = part + kPart * (size_t)cr.z;
#pragma unroll
    for (int e = 0; e < 36; e++) o[e] = acc[e];
    if (cr.w) {
      T bs[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
      for (int t = cr.x; t < cr.y; t++) {
        const T* q = tW + 6 * (size_t)terms[t].x;
#pragma unroll
        for (int i = 0; i < 6; i++) bs[i] += q[i];
      }
#pragma unroll
      for (int i = 0; i < 6; i++) o[36 + i] = bs[i];
    }
  }
}

// the FP64 residual rows of variable camera u of a window (the camera part
// of ir_res_cam) by one wave, lanes striding over the camera's edges and the
// 18 sums closed by a butterfly: r_u = -b_u - (H x)_u - (rel diag(H)_u +
// abs) x_u with b_u = sum_e J_e^T om_e r_e, (H x)_u = sum_e J_e^T om_e (J_e
// x), J_e = [camera block, point block] of edge e.  Windows carry no unary
// priors.  Lane 0 writes rU and bU64 and folds max |r|, max |b| into mr, mb.
__device__ __forceinline__ void ir_res_cam_wave(const BAView& v, const double* __restrict__ xU,
                                                const double* __restrict__ xV, double* __restrict__ rU,
                                                double* __restrict__ bU64, const DevCtl* ctl, int u, int lane,
                                                double& mr, double& mb) {
  const int c = v.var_cam_user[u];
  double R[kCamRec];
  load_cam(v, c, R);
  double xc[6];
#pragma unroll
  for (int i = 0; i < 6; i++) xc[i] = xU[6 * (size_t)u + i];
  double acc[18];
#pragma unroll
  for (int i = 0; i < 18; i++) acc[i] = 0.0;
  const int e1 = v.cam_ptr[u + 1];
  for (int e = v.cam_ptr[u] + lane; e < e1; e += 64) {
    const int q = v.ce_pt[e];
    const Proj pr = project(v, R, v.pts[3 * q], v.pts[3 * q + 1], v.pts[3 * q + 2], v.ce_obs[2 * e],
                            v.ce_obs[2 * e + 1], v.ce_w[e]);
    double j0[6], j1[6];
    jac_cam<double>(v, pr, j0, j1);
    double q0 = 0.0, q1 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      q0 += j0[i] * xc[i];
      q1 += j1[i] * xc[i];
    }
    const int pv = v.pt_var[q];
    if (pv >= 0) {
      double g0[3], g1[3];
      jac_pt<double>(v, pr, R, g0, g1);
      const double* xp = xV + 3 * (size_t)pv;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        q0 += g0[k] * xp[k];
        q1 += g1[k] * xp[k];
      }
    }
    const double om = pr.om;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      acc[i] += om * (j0[i] * pr.r0 + j1[i] * pr.r1);
      acc[6 + i] += om * (j0[i] * q0 + j1[i] * q1);
      acc[12 + i] += om * (j0[i] * j0[i] + j1[i] * j1[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < 18; i++) acc[i] = wave_sum(acc[i]);
  if (lane == 0) {
    const double rel = ctl->rel, ab = ctl->abs_;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const double r = -acc[i] - acc[6 + i] - (rel * acc[12 + i] + ab) * xc[i];
      rU[6 * (size_t)u + i] = r;
      bU64[6 * (size_t)u + i] = acc[i];
      mr = fmax(mr, fabs(r));
      mb = fmax(mb, fabs(acc[i]));
    }
  }
}

// Iterative refinement of a window's mixed-precision increment (as for a
// single problem): each step evaluates the FP64 residual r = -b - (H + D) dx
// of the damped normal equations from the edges (ir_res_point per point,
// ir_res_cam_wave per camera, a wave each), stops once max |r| <= ir_tol *
// max |b|, reduces r to the camera system (rS = rU - W V^-1 rV), solves it
// with the window's FP32 factor and back-substitutes, adding the correction
// to both parts of the increment.  Leaves the FP64 gradient in bU64 / bV64
// for the step statistics.  Every thread of the workgroup calls it; lds: the
// workgroup's 64 KiB (free after the solve, the control scalars at its end
// untouched).
template <typename T, int ALG>
__device__ __forceinline__ void win_refine(const BWin& w, DevCtl* c, double* lds, double* sh) {
  const BAView& v = w.v;
  const int tid = threadIdx.x, n = w.n, ld = w.nP;
  const double ab = c->abs_;
  const T* W = tp<T>(w.W);
  const T* Vinv = tp<T>(w.Vinv);
  const T* bV = tp<T>(w.bV);
  for (int it = 0; it < w.ir_steps; it++) {
    double mr = 0.0, mb = 0.0;
    for (int q = tid; q < v.P; q += kBT) ir_res_point(v, w.xU, w.xV, w.rV, w.bV64, c, q, mr, mb);
    for (int u = tid >> 6; u < v.NU; u += kBT / 64)
      ir_res_cam_wave(v, w.xU, w.xV, w.rU, w.bU64, c, u, tid & 63, mr, mb);
    mr = bmax(mr, sh);
    mb = bmax(mb, sh);
    if (w.ir_tol > 0.0 && mr <= w.ir_tol * mb) break;
    // rS_u = rU_u - sum_{W blocks k of u} W_k V^-1 rV_point(k)
    for (int e = tid; e < ld; e += kBT) {
      double val = 0.0;
      if (e < n) {
        const int u = e / 6, a = e - 6 * (e / 6);
        double sacc = w.rU[e];
        const int i1 = v.cw_ptr[u + 1];
        for (int ii = v.cw_ptr[u]; ii < i1; ii++) {
          const int k = v.cw_idx[ii], q = v.w_pt[k];
          const double* r = w.rV + 3 * (size_t)q;
          const T* m = Vinv + 6 * (size_t)q;
          const double r0 = r[0], r1 = r[1], r2 = r[2];
          const double g0 = (double)m[0] * r0 + (double)m[1] * r1 + (double)m[2] * r2;
          const double g1 = (double)m[1] * r0 + (double)m[3] * r1 + (double)m[4] * r2;
          const double g2 = (double)m[2] * r0 + (double)m[4] * r1 + (double)m[5] * r2;
          const T* wk = W + 18 * (size_t)k + 3 * a;
          sacc -= (double)wk[0] * g0 + (double)wk[1] * g1 + (double)wk[2] * g2;
        }
        val = sacc;
      }
      w.rS[e] = val;
    }
    __syncthreads();
    // z = S^-1 rS with the FP32 factor
    if constexpr (ALG == 64) {
      win_solve_tiles<T>(w, lds, w.rS, w.zU);
    } else {
      double* yv = lds + 4096;
      if (tid < ld) yv[tid] = w.rS[tid];
      __syncthreads();
      win_solve_sweeps<T>(tp<T>(w.S), n, ld, yv, yv + 256, w.zU);
    }
    __syncthreads();
    // xU += zU; xV += V^-1 (rV - W^T zU)
    for (int e = tid; e < n; e += kBT) w.xU[e] += w.zU[e];
    for (int q = tid; q < v.NV; q += kBT) {
      double i1, d1;
      backsub_point<T>(v, q, W, Vinv, bV, w.zU, w.xV, ab, w.rV, 1, nullptr, i1, d1);
    }
    __syncthreads();
    if (tid == 0) c->ir_count++;
  }
}

// the step's statistics (the small-increment test of
// LevenbergMarquardt.hpp:73-75 included) from per-thread sums of |x|^2 and
// x^T (mu x - b) over the camera part (incu, denu) and the point part
__device__ __forceinline__ void win_step_stats(DevCtl* c, double* sh, double incu, double denu, double inc, double den) {
  const double a1 = bsum(incu, sh), b1 = bsum(denu, sh), c1 = bsum(inc, sh), d1 = bsum(den, sh);
  if (threadIdx.x == 0) {
    const int fails = __hip_atomic_load(&c->fail_v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                      __hip_atomic_load(&c->fail_s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int ok = fails == 0 ? 1 : 0;
    c->solve_ok = ok;
    c->incsq = a1 + c1;
    c->den = b1 + d1;
    c->total_solves++;
    if (ok && c->iter >= c->minNumIt && c->incsq <= c->epsIncSquare) {
      c->done = 1;
      c->reason = 1;  // SuccessSmallIncrement
    }
  }
}

// mixed precision with refinement: refine the window's increment, then take
// the step statistics with the FP64 gradient the refinement evaluated
template <typename T, int ALG>
__device__ __forceinline__ void win_refine_step(const BWin& w, DevCtl* c, double* lds, double* sh) {
  win_refine<T, ALG>(w, c, lds, sh);
  const int tid = threadIdx.x;
  const double ab = c->abs_;
  double inc = 0.0, den = 0.0;
  for (int q = tid; q < w.v.NV; q += kBT)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double x = w.xV[3 * (size_t)q + k];
      inc += x * x;
      den += x * (ab * x - w.bV64[3 * (size_t)q + k]);
    }
  double incu = 0.0, denu = 0.0;
  if (tid < w.n) {
    const double x = w.xU[tid];
    incu = x * x;
    denu = x * (ab * x - w.bU64[tid]);
  }
  win_step_stats(c, sh, incu, denu, inc, den);
}

// IR: the mixed-precision instance with refinement, which leaves the step
// statistics to win_refine_step
template <typename T, int ALG, bool IR>
__device__ __forceinline__ void win_schur(const BWin& w, DevCtl* c, double* lds, double* sh) {
  const BAView v = w.v;  // a copy: the stores below cannot alias it
  const int tid = threadIdx.x, n = w.n, ld = w.nP;
  const double rel = c->rel, ab = c->abs_;
  const T* V = tp<T>(w.V);
  const T* bV = tp<T>(w.bV);
  const T* W = tp<T>(w.W);
  const T* U = tp<T>(w.U);
  const T* bU = tp<T>(w.bU);
  T* Vinv = tp<T>(w.Vinv);
  T* Y = tp<T>(w.Y);
  T* tW = tp<T>(w.tW);
  T* S = tp<T>(w.S);
  if (tid == 0) {
    c->fail_v = 0;
    c->fail_s = 0;
  }
  __syncthreads();
  int fv = 0;
  for (int q = tid; q < v.NV; q += kBT)
    if (!vinv_y_point<T, true>(v, q, V, bV, W, Vinv, Y, tW, rel, ab)) fv = 1;
  if (fv) atomicAdd(&c->fail_v, 1);
  // the padded S: zeros, identity on the padding diagonal
  for (int e = tid; e < ld * ld; e += kBT) {
    const int r = e % ld, cc = e / ld;
    S[e] = (r == cc && r >= n) ? (T)1 : (T)0;
  }
  __syncthreads();
  // S blocks: S_ii = U_i - sum_k Z_k Z_k^T (+ damping), S_ij = -sum_pairs Z_a Z_b^T
  // (SolverCholeskySchur.hpp:114-124, MatrixWWtMultiplier.hpp:124-153).
  // Phase 1: one lane per chunk of a block's terms, the whole 6x6 product per
  // term; phase 2: one thread per block entry sums its chunks in term order
  // (deterministic)
  win_s_chunks<T>(w);
  __syncthreads();
  {
    const T* part = tp<T>(w.part);
    for (int e = tid; e < 36 * w.nsb; e += kBT) {
      const int sb = e / 36, r = e - 36 * sb, a = r / 6, bb = r - 6 * (r / 6);
      const int i = w.sb_i[sb], j = w.sb_j[sb];
      if (i == j && bb < a) continue;
      T sum = (T)0;
      const int c1 = w.blk_ch_ptr[sb + 1];
#pragma unroll 8
      for (int cc = w.blk_ch_ptr[sb]; cc < c1; cc++) sum += part[kPart * (size_t)cc + r];
      if (i == j) {
        const double uu = (double)U[21 * (size_t)i + up6(a, bb)];
        double val = uu - (double)sum;
        if (a == bb && (rel != 0.0 || ab != 0.0)) {
          val += rel * uu;
          val += ab;
        }
        S[(size_t)(6 * i + a) * ld + 6 * i + bb] = (T)val;
      } else {
        S[(size_t)(6 * i + a) * ld + 6 * j + bb] = -sum;
      }
    }
  }
  // bS = -bU - W V^-1 (bV-part): one thread per entry (the tile variant keeps
  // it in xU, padded with zeros; the panel variant in LDS)
  double* yv = lds + 4096;
  if (tid < ld) {
    double val = 0.0;
    if (tid < n) {
      const int u = tid / 6, aa = tid % 6;
      const T* part = tp<T>(w.part);
      const int sb = w.sb_diag[u], c1 = w.blk_ch_ptr[sb + 1];
      T s = (T)0;
#pragma unroll 8
      for (int cc = w.blk_ch_ptr[sb]; cc < c1; cc++) s += part[kPart * (size_t)cc + 36 + aa];
      val = -(double)bU[6 * (size_t)u + aa] - (double)s;
    }
    if constexpr (ALG == 64) w.xU[tid] = val;
    else yv[tid] = val;
  }
  __syncthreads();
  if constexpr (ALG == 64) {
    win_cholesky_tiles<T>(w, c, lds);
    win_solve_tiles<T>(w, lds, w.xU, w.xU);
  } else {
    const bool ok = win_cholesky_panels<T, ALG>(S, n, ld, lds);
    if (!ok && tid == 0) atomicAdd(&c->fail_s, 1);
    win_solve_sweeps<T>(S, n, ld, yv, yv + 256, w.xU);
  }
  double inc = 0.0, den = 0.0;
  for (int q = tid; q < v.NV; q += kBT) {
    double i1, d1;
    backsub_point<T>(v, q, W, Vinv, bV, w.xU, w.xV, ab, nullptr, 0, nullptr, i1, d1);
    inc += i1;
    den += d1;
  }
  if constexpr (IR) return;
  double incu = 0.0, denu = 0.0;
  if (tid < n) {
    const double x = w.xU[tid];
    incu = x * x;
    denu = x * (ab * x - (double)bU[tid]);
  }
  win_step_stats(c, sh, incu, denu, inc, den);
}

// the LM loop of one window (LevenbergMarquardt.hpp:35-146); the control
// scalars live in c and are broadcast through LDS after every update
template <typename T, int ALG, bool IR>
__device__ __forceinline__ void win_lm(const BWin& w, DevCtl* c, double* lds, double* sh, int* shi) {
  const int tid = threadIdx.x;
  const long long maxsub = (long long)max(1, c->maxNumIt) * max(1, c->maxNumSubIt) + 4;
  for (long long it = 0; it < maxsub; it++) {
    if (tid == 0) {
      shi[1] = c->done;
      shi[2] = c->need_lin;
    }
    __syncthreads();
    const int done = shi[1], need_lin = shi[2];
    __syncthreads();
    if (done) break;
    if (need_lin) {
      win_linearize<T>(w, c, sh);
      win_copy(w.v, w.cams_bak, w.pts_bak, w.v.cams, w.v.pts);
      __syncthreads();
      if (tid == 0) {
        lm_after_lin(c, nullptr);
        shi[1] = c->done;
      }
      __syncthreads();
      if (shi[1]) break;
    }
    win_schur<T, ALG, IR>(w, c, lds, sh);
    __syncthreads();
    if constexpr (IR) {
      win_refine_step<T, ALG>(w, c, lds, sh);
      __syncthreads();
    }
    if (tid == 0) shi[3] = !c->done && c->solve_ok;
    __syncthreads();
    if (shi[3]) {
      {
        const BAView v = w.v;
        const double *xU = w.xU, *xV = w.xV;
        for (int idx = tid; idx < v.NU + v.NV; idx += kBT) {
          if (idx < v.NU) oplus_cam(v, idx, xU);
          else oplus_pt(v, idx - v.NU, xV);
        }
      }
      __syncthreads();
      win_cost<T>(w, c, sh);
    }
    __syncthreads();
    if (tid == 0) {
      lm_control(c, nullptr);
      shi[4] = c->rollback_pending;
    }
    __syncthreads();
    if (shi[4]) {
      win_copy(w.v, w.v.cams, w.v.pts, w.cams_bak, w.pts_bak);
      __syncthreads();
      if (tid == 0) c->rollback_pending = 0;
    }
  }
}

// LDS: one 64 KiB buffer; the tile variant's two 64x64 FP64 tiles or the
// panel variant's panel and vectors, with the reduction / control scalars in
// its last 96 bytes, which only the tile factorisation (inside a Schur step)
// overwrites
template <typename T, int ALG, bool IR>
__global__ __launch_bounds__(256) void k_batch_lm(const BWin* __restrict__ wins, int nwin, DevCtl* ctls,
                                                  int* counter) {
  __shared__ double lds[8192];
  double* sh = lds + 8184;
  int* shi = (int*)(lds + 8176);
  for (int round = 0; round < nwin; round++) {  // at most nwin windows per workgroup
    if (threadIdx.x == 0) {
      const int k = atomicAdd(counter, 1);
      shi[7] = k < nwin ? counter[1 + k] : nwin;
    }
    __syncthreads();
    const int wi = shi[7];
    __syncthreads();
    if (wi >= nwin) break;
    if constexpr (ALG == 64) {
      // the tile variant reads the descriptor's fields where it uses them
      // (its factorisation needs the registers)
      const BWin& w = wins[wi];
      win_lm<T, ALG, IR>(w, ctls + wi, lds, sh, shi);
    } else {
      const BWin w = wins[wi];
      win_lm<T, ALG, IR>(w, ctls + wi, lds, sh, shi);
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

}  // namespace

void launch_batch_reset(int* counter, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, counter);
}

template <typename T, bool IR>
static int launch_lm(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int g, hipStream_t st) {
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_batch_lm<T, 8, IR>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    case 16: hipLaunchKernelGGL((k_batch_lm<T, 16, IR>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    case 64: hipLaunchKernelGGL((k_batch_lm<T, 64, IR>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    default: return -1;
  }
  return 0;
}

template <>
int launch_batch_lm<double>(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int grid, bool refine,
                            hipStream_t st) {
  const int g = grid < nwin ? grid : nwin;
  if (g <= 0) return 0;
  if (refine) return -2;  // FP64 windows are not refined
  return launch_lm<double, false>(wins, nwin, ctls, counter, nb, g, st);
}

template <>
int launch_batch_lm<float>(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int grid, bool refine,
                           hipStream_t st) {
  const int g = grid < nwin ? grid : nwin;
  if (g <= 0) return 0;
  return refine ? launch_lm<float, true>(wins, nwin, ctls, counter, nb, g, st)
                : launch_lm<float, false>(wins, nwin, ctls, counter, nb, g, st);
}

}  // namespace nsn
