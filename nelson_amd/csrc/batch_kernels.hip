// Batched independent BA windows: one persistent launch in which each
// workgroup takes windows from an atomic counter and runs the complete
// Levenberg-Marquardt solve of a window (nelson/LevenbergMarquardt.hpp:35-146
// with the Schur step of nelson/SolverCholeskySchur.hpp:91-157) on its own:
// linearisation, V^-1 / W V^-1, S assembly from the (W, W) pair lists, a
// dense blocked Cholesky of S (n <= 256) whose panels are factorised in FP64
// in LDS, the two triangular solves, back-substitution, oplus, cost and the
// accept / reject schedule.  Windows never wait for each other, so the launch
// needs no inter-workgroup synchronisation; every loop is bounded.
// The per-edge, per-point and LM-control arithmetic is shared with the
// single-problem kernels (ba_device.h).
#include <hip/hip_runtime.h>

#include "ba_device.h"
#include "batch.h"

namespace nsn {

namespace {

constexpr int kBT = 256;      // threads per window
constexpr int kMaxS = 256;    // largest S order of a window
constexpr int kTeams = 7;     // 36-lane teams of the S assembly

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
__device__ void win_linearize(const BWin& w, DevCtl* c, double* sh) {
  const BAView& v = w.v;
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
    double R[12];
#pragma unroll
    for (int i = 0; i < 12; i++) R[i] = v.cams[12 * cam + i];
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
__device__ void win_cost(const BWin& w, DevCtl* c, double* sh) {
  double chi2 = 0.0, mb = 0.0, md = 0.0;
  for (int p = threadIdx.x; p < w.v.P; p += kBT)
    lin_point<T, false>(w.v, p, nullptr, nullptr, nullptr, chi2, mb, md);
  const double tot = bsum(chi2, sh);
  if (threadIdx.x == 0) c->chi2 = tot;
}

__device__ void win_copy(const BAView& v, double* cd, double* pd, const double* cs, const double* ps) {
  for (int i = threadIdx.x; i < 12 * v.C; i += kBT) cd[i] = cs[i];
  for (int i = threadIdx.x; i < 3 * v.P; i += kBT) pd[i] = ps[i];
}

// dense Cholesky of the n x n lower S (column-major) in place: panels of NB
// columns are factorised in FP64 in LDS (pan: NB columns of kMaxS rows), the
// trailing matrix is updated from the panel; false if a pivot is not positive
template <typename T, int NB>
__device__ bool win_cholesky(T* S, int n, double* pan) {
  const int tid = threadIdx.x;
  bool ok = true;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = min(NB, n - k0), m = n - k0;
    for (int e = tid; e < nb * m; e += kBT) {
      const int cc = e / m, r = e % m;
      pan[cc * kMaxS + r] = (double)S[(size_t)(k0 + cc) * n + k0 + r];
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
      if (r >= cc) S[(size_t)(k0 + cc) * n + k0 + r] = (T)pan[cc * kMaxS + r];
    }
    const int mt = m - nb;
    for (int e = tid; e < mt * mt; e += kBT) {
      const int jj = e / mt, ii = e % mt;
      if (ii >= jj) {
        double s = 0.0;
#pragma unroll 4
        for (int cc = 0; cc < nb; cc++) s += pan[cc * kMaxS + nb + ii] * pan[cc * kMaxS + nb + jj];
        T* a = S + (size_t)(k0 + nb + jj) * n + k0 + nb + ii;
        *a = (T)((double)*a - s);
      }
    }
    __syncthreads();
  }
  return ok;
}

// one Schur step with damping (rel, abs) = (c->rel, c->abs_); writes xU, xV
// and the step statistics (the small-increment test of
// LevenbergMarquardt.hpp:73-75 included)
template <typename T, int NB>
__device__ void win_schur(const BWin& w, DevCtl* c, double* pan, double* yv, double* ys, double* sh, int* shi) {
  const BAView& v = w.v;
  const int tid = threadIdx.x, n = w.n;
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
  if (tid == 0) shi[0] = 0;
  __syncthreads();
  int fv = 0;
  for (int q = tid; q < v.NV; q += kBT)
    if (!vinv_y_point<T>(v, q, V, bV, W, Vinv, Y, tW, rel, ab)) fv = 1;
  if (fv) atomicOr(shi, 1);
  for (int e = tid; e < n * n; e += kBT) S[e] = (T)0;
  __syncthreads();
  // S blocks: S_ii = U_i - sum_k Y_k W_k^T (+ damping), S_ij = -sum_pairs Y_a W_b^T
  // (SolverCholeskySchur.hpp:114-124, MatrixWWtMultiplier.hpp:124-153);
  // a 36-lane team per block, lane (a, bb) owns one entry
  const int team = tid / 36, l36 = tid % 36, a = l36 / 6, bb = l36 % 6;
  if (team < kTeams) {
    for (int sb = team; sb < w.nsb; sb += kTeams) {
      const int i = w.sb_i[sb], j = w.sb_j[sb];
      T s = (T)0;
      if (i == j) {
        if (bb < a) continue;
        const int k1 = v.cw_ptr[i + 1];
        for (int kk = v.cw_ptr[i]; kk < k1; kk++) {
          const int k = v.cw_idx[kk];
          const T* y = Y + 18 * (size_t)k + 3 * a;
          const T* ww = W + 18 * (size_t)k + 3 * bb;
          s += y[0] * ww[0] + y[1] * ww[1] + y[2] * ww[2];
        }
        const double uu = (double)U[21 * (size_t)i + up6(a, bb)];
        double val = uu - (double)s;
        if (a == bb && (rel != 0.0 || ab != 0.0)) {
          val += rel * uu;
          val += ab;
        }
        S[(size_t)(6 * i + a) * n + 6 * i + bb] = (T)val;
      } else {
        const int p1 = w.sb_ptr[sb + 1];
        for (int pp = w.sb_ptr[sb]; pp < p1; pp++) {
          const T* y = Y + 18 * (size_t)w.sb_pairs[2 * pp] + 3 * a;
          const T* ww = W + 18 * (size_t)w.sb_pairs[2 * pp + 1] + 3 * bb;
          s += y[0] * ww[0] + y[1] * ww[1] + y[2] * ww[2];
        }
        S[(size_t)(6 * i + a) * n + 6 * j + bb] = -s;
      }
    }
  }
  // bS = -bU - W V^-1 (bV-part): one thread per entry
  if (tid < n) {
    const int u = tid / 6, aa = tid % 6;
    T s = (T)0;
    const int k1 = v.cw_ptr[u + 1];
    for (int kk = v.cw_ptr[u]; kk < k1; kk++) s += tW[6 * (size_t)v.cw_idx[kk] + aa];
    yv[tid] = -(double)bU[6 * (size_t)u + aa] - (double)s;
  }
  __syncthreads();
  const bool ok_s = win_cholesky<T, NB>(S, n, pan);
  // forward solve L y = bS, then backward L^T x = y (column sweeps)
  for (int cc = 0; cc < n; cc++) {
    const double yc = yv[cc] / (double)S[(size_t)cc * n + cc];
    if (tid == 0) ys[cc] = yc;
    if (tid > cc && tid < n) yv[tid] -= (double)S[(size_t)cc * n + tid] * yc;
    __syncthreads();
  }
  for (int cc = n - 1; cc >= 0; cc--) {
    const double xc = ys[cc] / (double)S[(size_t)cc * n + cc];
    if (tid == 0) w.xU[cc] = xc;
    if (tid < cc) ys[tid] -= (double)S[(size_t)tid * n + cc] * xc;
    __syncthreads();
  }
  double inc = 0.0, den = 0.0;
  for (int q = tid; q < v.NV; q += kBT) {
    double i1, d1;
    backsub_point<T>(v, q, W, Vinv, bV, w.xU, w.xV, ab, nullptr, 0, nullptr, i1, d1);
    inc += i1;
    den += d1;
  }
  double incu = 0.0, denu = 0.0;
  if (tid < n) {
    const double x = w.xU[tid];
    incu = x * x;
    denu = x * (ab * x - (double)bU[tid]);
  }
  const double a1 = bsum(incu, sh), b1 = bsum(denu, sh), c1 = bsum(inc, sh), d1 = bsum(den, sh);
  if (tid == 0) {
    const int ok = (shi[0] == 0 && ok_s) ? 1 : 0;
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

// the LM loop of one window (LevenbergMarquardt.hpp:35-146); the control
// scalars live in c and are broadcast through LDS after every update
template <typename T, int NB>
__device__ void win_lm(const BWin& w, DevCtl* c, double* pan, double* yv, double* ys, double* sh, int* shi) {
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
    win_schur<T, NB>(w, c, pan, yv, ys, sh, shi);
    __syncthreads();
    if (tid == 0) shi[3] = !c->done && c->solve_ok;
    __syncthreads();
    if (shi[3]) {
      for (int idx = tid; idx < w.v.NU + w.v.NV; idx += kBT) {
        if (idx < w.v.NU) oplus_cam(w.v, idx, w.xU);
        else oplus_pt(w.v, idx - w.v.NU, w.xV);
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

template <typename T, int NB>
__global__ __launch_bounds__(256) void k_batch_lm(const BWin* __restrict__ wins, int nwin, DevCtl* ctls,
                                                  int* counter) {
  __shared__ double pan[NB * kMaxS];
  __shared__ double yv[kMaxS], ys[kMaxS];
  __shared__ double sh[4];
  __shared__ int shi[8];
  for (int round = 0; round < nwin; round++) {  // at most nwin windows per workgroup
    if (threadIdx.x == 0) shi[7] = atomicAdd(counter, 1);
    __syncthreads();
    const int wi = shi[7];
    __syncthreads();
    if (wi >= nwin) break;
    const BWin w = wins[wi];
    win_lm<T, NB>(w, ctls + wi, pan, yv, ys, sh, shi);
    __syncthreads();
  }
}

}  // namespace

template <typename T>
int launch_batch_lm(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int grid, hipStream_t st) {
  const int g = grid < nwin ? grid : nwin;
  if (g <= 0) return 0;
  switch (nb) {
    case 4: hipLaunchKernelGGL((k_batch_lm<T, 4>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    case 8: hipLaunchKernelGGL((k_batch_lm<T, 8>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    case 16: hipLaunchKernelGGL((k_batch_lm<T, 16>), dim3(g), dim3(kBT), 0, st, wins, nwin, ctls, counter); break;
    default: return -1;
  }
  return 0;
}

template int launch_batch_lm<double>(const BWin*, int, DevCtl*, int*, int, int, hipStream_t);
template int launch_batch_lm<float>(const BWin*, int, DevCtl*, int*, int, int, hipStream_t);

}  // namespace nsn
