// Bundle-adjustment kernels of the Schur LM step on gfx950 (wave64).
//
// Data flow of one LM sub-iteration (nelson/SolverCholeskySchur.hpp:91-157,
// nelson/LevenbergMarquardt.hpp:67-130):
//   k_lin_points / k_lin_cams  : residuals, Jacobians, V, bV, W (point-major)
//                                and U, bU (camera-major)        [update(true)]
//   k_vinv_y                   : damped V^-1, Y = W V^-1, t = W V^-1 (-bV)
//   k_s_diag                   : diagonal S blocks and bS per camera
//   k_s_tiles                  : off-diagonal S blocks gathered per output tile
//   (chol_kernels.hip)         : tile Cholesky of [S; bS^T], backward solve
//   k_backsub                  : xV = V^-1 (-bV - W^T xU)
//   k_oplus, k_lin_points<cost>: candidate state and its chi2     [update(false)]
//   k_lm_control               : rho, accept/reject, damping schedule
// All reductions are deterministic: fixed per-thread order, fixed wave
// butterfly, fixed cross-wave order; maxima use integer atomics on the bit
// patterns of non-negative doubles, which are order independent.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nsn {

__device__ __forceinline__ bool gate_open(const DevCtl* c, int gate) {
  switch (gate) {
    case GATE_ALWAYS: return true;
    case GATE_NOT_DONE: return !c->done;
    case GATE_NEED_LIN: return !c->done && c->need_lin;
    case GATE_SOLVE_OK: return !c->done && c->solve_ok;
    case GATE_ROLLBACK: return c->rollback_pending != 0;
  }
  return true;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// deterministic block sum for blockDim.x == 256; result valid in thread 0
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* a, double v) {
  if (v > 0.0) atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// packed upper index of a 6x6 symmetric block, i <= j
__host__ __device__ constexpr int up6(int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); }

// ---------------------------------------------------------------------------
// the reprojection edge
// ---------------------------------------------------------------------------
struct Proj {
  double Xc0, Xc1, Xc2, iz, r0, r1, cost, om;
};

__device__ __forceinline__ Proj project(const BAView& v, const double* __restrict__ R, double X0,
                                        double X1, double X2, double o0, double o1, double we) {
  Proj p;
  p.Xc0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + R[9];
  p.Xc1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + R[10];
  p.Xc2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + R[11];
  p.iz = 1.0 / p.Xc2;
  p.r0 = v.f * p.Xc0 * p.iz + v.cx - o0;
  p.r1 = v.f * p.Xc1 * p.iz + v.cy - o1;
  double s = p.r0 * p.r0 + p.r1 * p.r1;
  double psi = 1.0;
  if (v.huber > 0.0 && s > v.huber * v.huber) {
    double nn = sqrt(s);
    p.cost = 2.0 * v.huber * nn - v.huber * v.huber;
    psi = v.huber / nn;
  } else {
    p.cost = s;
  }
  p.cost *= we;
  p.om = we * psi;
  return p;
}

// d(r)/d(omega, upsilon) for the left SE(3) perturbation: D [-[Xc]x | I]
template <typename T>
__device__ __forceinline__ void jac_cam(const BAView& v, const Proj& p, T* j0, T* j1) {
  const double a = v.f * p.iz;
  const double b0 = -a * p.Xc0 * p.iz;
  const double b1 = -a * p.Xc1 * p.iz;
  j0[0] = (T)(b0 * p.Xc1);
  j0[1] = (T)(a * p.Xc2 - b0 * p.Xc0);
  j0[2] = (T)(-a * p.Xc1);
  j0[3] = (T)a;
  j0[4] = (T)0;
  j0[5] = (T)b0;
  j1[0] = (T)(-a * p.Xc2 + b1 * p.Xc1);
  j1[1] = (T)(-b1 * p.Xc0);
  j1[2] = (T)(a * p.Xc0);
  j1[3] = (T)0;
  j1[4] = (T)a;
  j1[5] = (T)b1;
}

// d(r)/d(X) = D R
template <typename T>
__device__ __forceinline__ void jac_pt(const BAView& v, const Proj& p, const double* __restrict__ R,
                                       T* g0, T* g1) {
  const double a = v.f * p.iz;
  const double b0 = -a * p.Xc0 * p.iz;
  const double b1 = -a * p.Xc1 * p.iz;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    g0[k] = (T)(a * R[k] + b0 * R[6 + k]);
    g1[k] = (T)(a * R[3 + k] + b1 * R[6 + k]);
  }
}

// ---------------------------------------------------------------------------
// linearisation, point-major pass: V, bV, W and the chi2 of every edge
// (nelson/BaseSection.cpp:88-169 + updateH22Block/updateH12Block)
// ---------------------------------------------------------------------------
template <typename T, bool HESS>
__global__ __launch_bounds__(256) void k_lin_points(BAView v, T* __restrict__ V, T* __restrict__ bV,
                                                    T* __restrict__ W, double* __restrict__ part_chi2,
                                                    unsigned long long* __restrict__ maxes, DevCtl* ctl,
                                                    int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double chi2 = 0.0, mb = 0.0, md = 0.0;
  if (p < v.P) {
    const double X0 = v.pts[3 * p], X1 = v.pts[3 * p + 1], X2 = v.pts[3 * p + 2];
    const int pv = v.pt_var[p];
    T vv[6] = {0, 0, 0, 0, 0, 0};
    T bv[3] = {0, 0, 0};
    T w[18];
    int cur = -1;
    const int e0 = v.pt_ptr[p], e1 = v.pt_ptr[p + 1];
    for (int e = e0; e < e1; e++) {
      const int c = v.pe_cam[e];
      double R[12];
#pragma unroll
      for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
      const Proj pr = project(v, R, X0, X1, X2, v.pe_obs[2 * e], v.pe_obs[2 * e + 1], v.pe_w[e]);
      chi2 += pr.cost;
      if (HESS) {
        const int cv = v.cam_var[c];
        if (pv < 0) continue;
        T g0[3], g1[3];
        jac_pt<T>(v, pr, R, g0, g1);
        const T om = (T)pr.om, r0 = (T)pr.r0, r1 = (T)pr.r1;
        vv[0] += om * (g0[0] * g0[0] + g1[0] * g1[0]);
        vv[1] += om * (g0[0] * g0[1] + g1[0] * g1[1]);
        vv[2] += om * (g0[0] * g0[2] + g1[0] * g1[2]);
        vv[3] += om * (g0[1] * g0[1] + g1[1] * g1[1]);
        vv[4] += om * (g0[1] * g0[2] + g1[1] * g1[2]);
        vv[5] += om * (g0[2] * g0[2] + g1[2] * g1[2]);
#pragma unroll
        for (int k = 0; k < 3; k++) bv[k] += om * (g0[k] * r0 + g1[k] * r1);
        if (cv >= 0) {
          T j0[6], j1[6];
          jac_cam<T>(v, pr, j0, j1);
          const int blk = v.pe_wblk[e];
          if (blk != cur) {
            if (cur >= 0) {
#pragma unroll
              for (int i = 0; i < 18; i++) W[18 * (size_t)cur + i] = w[i];
            }
#pragma unroll
            for (int i = 0; i < 18; i++) w[i] = (T)0;
            cur = blk;
          }
#pragma unroll
          for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) w[i * 3 + j] += om * (j0[i] * g0[j] + j1[i] * g1[j]);
        }
      }
    }
    if (HESS) {
      if (cur >= 0) {
#pragma unroll
        for (int i = 0; i < 18; i++) W[18 * (size_t)cur + i] = w[i];
      }
      if (pv >= 0) {
#pragma unroll
        for (int i = 0; i < 6; i++) V[6 * (size_t)pv + i] = vv[i];
#pragma unroll
        for (int i = 0; i < 3; i++) bV[3 * (size_t)pv + i] = bv[i];
        mb = fmax(fabs((double)bv[0]), fmax(fabs((double)bv[1]), fabs((double)bv[2])));
        md = fmax(fabs((double)vv[0]), fmax(fabs((double)vv[3]), fabs((double)vv[5])));
      }
    }
  }
  double s = block_sum256(chi2, sh);
  if (threadIdx.x == 0) part_chi2[blockIdx.x] = s;
  if (HESS) {
    mb = wave_max(mb);
    md = wave_max(md);
    if ((threadIdx.x & 63) == 0) {
      atomic_max_nonneg(maxes + 1, mb);
      atomic_max_nonneg(maxes + 3, md);
    }
  }
}

// ---------------------------------------------------------------------------
// linearisation, camera-major pass: U (21 packed upper) and bU per variable
// camera (updateH11Block); one workgroup per camera
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_lin_cams(BAView v, T* __restrict__ U, T* __restrict__ bU,
                                                  unsigned long long* __restrict__ maxes, DevCtl* ctl,
                                                  int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][27];
  const int u = blockIdx.x;
  const int c = v.var_cam_user[u];
  double R[12];
#pragma unroll
  for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
  T acc[27];
#pragma unroll
  for (int i = 0; i < 27; i++) acc[i] = (T)0;
  const int e1 = v.cam_ptr[u + 1];
  for (int e = v.cam_ptr[u] + threadIdx.x; e < e1; e += 256) {
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
    double s = wave_sum((double)acc[i]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    const int i = threadIdx.x;
    const double s = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    if (i < 21) U[21 * (size_t)u + i] = (T)s;
    else bU[6 * (size_t)u + (i - 21)] = (T)s;
  }
  if (threadIdx.x == 0) {
    double mb = 0.0, md = 0.0;
    for (int i = 0; i < 6; i++) {
      mb = fmax(mb, fabs((double)(T)(((red[0][21 + i] + red[1][21 + i]) + red[2][21 + i]) + red[3][21 + i])));
      const int d = up6(i, i);
      md = fmax(md, fabs((double)(T)(((red[0][d] + red[1][d]) + red[2][d]) + red[3][d])));
    }
    atomic_max_nonneg(maxes + 0, mb);
    atomic_max_nonneg(maxes + 2, md);
  }
}

__global__ void k_reset_lin(unsigned long long* maxes, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  if (threadIdx.x < 4) maxes[threadIdx.x] = 0ull;
}

// chi2 = sum of the block partials in a fixed order; full linearisation also
// publishes max|b| and max|diag H|
__global__ __launch_bounds__(256) void k_finalize_lin(const double* __restrict__ part_chi2, int nblk,
                                                      const unsigned long long* __restrict__ maxes,
                                                      DevCtl* ctl, int gate, int full) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) s += part_chi2[i];
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) {
    ctl->chi2 = s;
    if (full) {
      ctl->chi2_lin = s;
      ctl->maxb = fmax(__longlong_as_double((long long)maxes[0]), __longlong_as_double((long long)maxes[1]));
      ctl->maxdiag = fmax(__longlong_as_double((long long)maxes[2]), __longlong_as_double((long long)maxes[3]));
    }
  }
}

// ---------------------------------------------------------------------------
// V^-1 (MatrixDiagInv::compute, nelson/MatrixDiagInv.hpp:25-140) fused with
// Y = W V^-1 (MatrixWVinvMultiplier::multiply, MatrixWVinvMultiplier.hpp:23-106)
// and t = W V^-1 (-bV) (the W * V^-1 bV term of bS, SolverCholeskySchur.hpp:114).
// The damped 3x3 block is inverted through its Cholesky factor in FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_vinv_y(BAView v, const T* __restrict__ V, const T* __restrict__ bV,
                                                const T* __restrict__ W, T* __restrict__ Vinv,
                                                T* __restrict__ Y, T* __restrict__ tW, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= v.NV) return;
  const double rel = ctl->rel, ab = ctl->abs_;
  double a00 = V[6 * (size_t)q + 0], a01 = V[6 * (size_t)q + 1], a02 = V[6 * (size_t)q + 2];
  double a11 = V[6 * (size_t)q + 3], a12 = V[6 * (size_t)q + 4], a22 = V[6 * (size_t)q + 5];
  if (rel != 0.0 || ab != 0.0) {
    a00 = a00 * (1.0 + rel) + ab;
    a11 = a11 * (1.0 + rel) + ab;
    a22 = a22 * (1.0 + rel) + ab;
  }
  double vi[6];
  bool ok = a00 > 0.0;
  const double l00 = sqrt(a00);
  const double l10 = a01 / l00, l20 = a02 / l00;
  const double d1 = a11 - l10 * l10;
  ok = ok && d1 > 0.0;
  const double l11 = sqrt(d1);
  const double l21 = (a12 - l20 * l10) / l11;
  const double d2 = a22 - l20 * l20 - l21 * l21;
  ok = ok && d2 > 0.0;
  const double l22 = sqrt(d2);
  const double m00 = 1.0 / l00, m11 = 1.0 / l11, m22 = 1.0 / l22;
  const double m10 = -l10 * m00 * m11;
  const double m21 = -l21 * m11 * m22;
  const double m20 = -(l20 * m00 + l21 * m10) * m22;
  vi[0] = m00 * m00 + m10 * m10 + m20 * m20;
  vi[1] = m10 * m11 + m20 * m21;
  vi[2] = m20 * m22;
  vi[3] = m11 * m11 + m21 * m21;
  vi[4] = m21 * m22;
  vi[5] = m22 * m22;
  ok = ok && isfinite(vi[0]) && isfinite(vi[3]) && isfinite(vi[5]);
  if (!ok) {
    atomicAdd(&ctl->fail_v, 1);
    for (int i = 0; i < 6; i++) vi[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 6; i++) Vinv[6 * (size_t)q + i] = (T)vi[i];
  const double b0 = bV[3 * (size_t)q], b1 = bV[3 * (size_t)q + 1], b2 = bV[3 * (size_t)q + 2];
  // x0 = V^-1 (-bV)
  const double x0 = -(vi[0] * b0 + vi[1] * b1 + vi[2] * b2);
  const double x1 = -(vi[1] * b0 + vi[3] * b1 + vi[4] * b2);
  const double x2 = -(vi[2] * b0 + vi[4] * b1 + vi[5] * b2);
  const T M[9] = {(T)vi[0], (T)vi[1], (T)vi[2], (T)vi[1], (T)vi[3], (T)vi[4], (T)vi[2], (T)vi[4], (T)vi[5]};
  const int k1 = v.wptr[q + 1];
  for (int k = v.wptr[q]; k < k1; k++) {
    const T* w = W + 18 * (size_t)k;
    T y[18];
    T t[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const T w0 = w[3 * i], w1 = w[3 * i + 1], w2 = w[3 * i + 2];
#pragma unroll
      for (int j = 0; j < 3; j++) y[3 * i + j] = w0 * M[j] + w1 * M[3 + j] + w2 * M[6 + j];
      t[i] = (T)((double)w0 * x0 + (double)w1 * x1 + (double)w2 * x2);
    }
#pragma unroll
    for (int i = 0; i < 18; i++) Y[18 * (size_t)k + i] = y[i];
#pragma unroll
    for (int i = 0; i < 6; i++) tW[6 * (size_t)k + i] = t[i];
  }
}

// ---------------------------------------------------------------------------
// diagonal S blocks and bS, one workgroup per variable camera:
//   S_uu = U_u - sum_k Y_k W_k^T, then S_ii += rel * U_ii + abs
//   bS_u = -bU_u - sum_k t_k
// (SolverCholeskySchur.hpp:114-124; MatrixWWtMultiplier.hpp:139-153 for the
// diagonal pairs)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_s_diag(BAView v, const T* __restrict__ U, const T* __restrict__ bU,
                                                const T* __restrict__ Y, const T* __restrict__ W,
                                                const T* __restrict__ tW, T* __restrict__ Sd,
                                                T* __restrict__ bS, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][27];
  const int u = blockIdx.x;
  T acc[27];
#pragma unroll
  for (int i = 0; i < 27; i++) acc[i] = (T)0;
  const int i1 = v.cw_ptr[u + 1];
  for (int ii = v.cw_ptr[u] + threadIdx.x; ii < i1; ii += 256) {
    const int k = v.cw_idx[ii];
    T y[18], w[18];
#pragma unroll
    for (int i = 0; i < 18; i++) { y[i] = Y[18 * (size_t)k + i]; w[i] = W[18 * (size_t)k + i]; }
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
      for (int b = a; b < 6; b++)
        acc[up6(a, b)] += y[3 * a] * w[3 * b] + y[3 * a + 1] * w[3 * b + 1] + y[3 * a + 2] * w[3 * b + 2];
#pragma unroll
    for (int a = 0; a < 6; a++) acc[21 + a] += tW[6 * (size_t)k + a];
  }
#pragma unroll
  for (int i = 0; i < 27; i++) {
    double s = wave_sum((double)acc[i]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    const int i = threadIdx.x;
    const double s = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    if (i < 21) {
      const double uu = (double)U[21 * (size_t)u + i];
      double val = uu - s;
      const bool diag = (i == 0 || i == 6 || i == 11 || i == 15 || i == 18 || i == 20);
      if (diag) {
        const double rel = ctl->rel, ab = ctl->abs_;
        if (rel != 0.0 || ab != 0.0) {
          val += rel * uu;
          val += ab;
        }
      }
      Sd[21 * (size_t)u + i] = (T)val;
    } else {
      bS[6 * (size_t)u + (i - 21)] = (T)(-(double)bU[6 * (size_t)u + (i - 21)] - s);
    }
  }
}

// ---------------------------------------------------------------------------
// S tiles: one workgroup per lower TB x TB tile of [S; bS^T]. The tile is
// built in LDS (zeros, identity padding, the bS row, diagonal blocks from Sd,
// off-diagonal blocks S_ij = -sum_p Y_ip W_jp^T gathered over their pair
// lists, MatrixWWtMultiplier.hpp:124-153) and written once, coalesced.
// An upper block (i, j), i <= j, is stored transposed at rows 6j.., cols 6i..
// ---------------------------------------------------------------------------
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_s_tiles(TileView tv, const int* __restrict__ tile_order,
                                                 const int* __restrict__ tile_blk_ptr,
                                                 const int* __restrict__ tile_blk, const int* __restrict__ sb_i,
                                                 const int* __restrict__ sb_j, const int* __restrict__ sb_ptr,
                                                 const int* __restrict__ sb_pairs, const T* __restrict__ Sd,
                                                 const T* __restrict__ bS, const T* __restrict__ Y,
                                                 const T* __restrict__ W, T* __restrict__ tiles, DevCtl* ctl,
                                                 int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ T lt[TB * TB];
  const int t = blockIdx.x;
  const int ti = tile_order[2 * t], tj = tile_order[2 * t + 1];
  const int r0 = ti * TB, c0 = tj * TB, n = tv.n;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int lr = e % TB, lc = e / TB;
    const int R = r0 + lr, Cc = c0 + lc;
    T val = (T)0;
    if (R == Cc && R >= n) val = (T)1;
    else if (R == n && Cc < n) val = bS[Cc];
    lt[e] = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int a = lane / 6, bb = lane % 6;
  const int b1 = tile_blk_ptr[t + 1];
  for (int b = tile_blk_ptr[t] + wv; b < b1; b += 4) {
    const int sb = tile_blk[b];
    const int i = sb_i[sb], j = sb_j[sb];
    const int R = 6 * j + bb, Cc = 6 * i + a;
    const bool inside = lane < 36 && R >= r0 && R < r0 + TB && Cc >= c0 && Cc < c0 + TB;
    if (i == j) {
      if (inside && bb >= a) lt[(Cc - c0) * TB + (R - r0)] = Sd[21 * (size_t)i + up6(a, bb)];
      continue;
    }
    // S_ij = -sum_p Y_ip W_jp^T: lane (a, bb) < 36 owns one entry; the pair
    // indices of up to 64 pairs are loaded at once and broadcast by readlane
    T s = (T)0;
    const int p0 = sb_ptr[sb], np = sb_ptr[sb + 1] - p0;
    for (int base = 0; base < np; base += 64) {
      const int cnt = min(64, np - base);
      int my_ka = 0, my_kb = 0;
      if (lane < cnt) {
        my_ka = sb_pairs[2 * (p0 + base + lane)];
        my_kb = sb_pairs[2 * (p0 + base + lane) + 1];
      }
#pragma unroll 4
      for (int q = 0; q < cnt; q++) {
        const int ka = __builtin_amdgcn_readlane(my_ka, q);
        const int kb = __builtin_amdgcn_readlane(my_kb, q);
        if (lane < 36) {
          const T* y = Y + 18 * (size_t)ka + 3 * a;
          const T* w = W + 18 * (size_t)kb + 3 * bb;
          s += y[0] * w[0] + y[1] * w[1] + y[2] * w[2];
        }
      }
    }
    if (inside) lt[(Cc - c0) * TB + (R - r0)] = -s;
  }
  __syncthreads();
  T* out = tiles + (size_t)t * TB * TB;
  for (int e = threadIdx.x; e < TB * TB; e += 256) out[e] = lt[e];
}

// ---------------------------------------------------------------------------
// back-substitution (SolverCholeskySchur.hpp:144-147; MatrixWtXMultiplier.hpp:49-129):
//   xV_p = V_p^-1 (-bV_p - sum_k W_k^T xU_c(k)), plus the increment statistics
//   ||xV||^2 and xV^T (mu xV - bV) for the rho denominator (DoubleSection.h:300-311)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_backsub(BAView v, const T* __restrict__ W, const T* __restrict__ Vinv,
                                                 const T* __restrict__ bV, const double* __restrict__ xU,
                                                 double* __restrict__ xV, double* __restrict__ part,
                                                 DevCtl* ctl, int gate, const double* __restrict__ rhs_v,
                                                 int add, const double* __restrict__ bden) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  double inc = 0.0, den = 0.0;
  if (q < v.NV) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const int k1 = v.wptr[q + 1];
    for (int k = v.wptr[q]; k < k1; k++) {
      const T* w = W + 18 * (size_t)k;
      const double* x = xU + 6 * (size_t)v.w_cam[k];
#pragma unroll
      for (int i = 0; i < 6; i++) {
        const double xi = x[i];
        s0 += (double)w[3 * i] * xi;
        s1 += (double)w[3 * i + 1] * xi;
        s2 += (double)w[3 * i + 2] * xi;
      }
    }
    double bv0, bv1, bv2;
    if (bden) {
      bv0 = bden[3 * (size_t)q]; bv1 = bden[3 * (size_t)q + 1]; bv2 = bden[3 * (size_t)q + 2];
    } else {
      bv0 = bV[3 * (size_t)q]; bv1 = bV[3 * (size_t)q + 1]; bv2 = bV[3 * (size_t)q + 2];
    }
    double t0, t1, t2;
    if (rhs_v) {
      t0 = rhs_v[3 * (size_t)q] - s0; t1 = rhs_v[3 * (size_t)q + 1] - s1; t2 = rhs_v[3 * (size_t)q + 2] - s2;
    } else {
      t0 = -(double)bV[3 * (size_t)q] - s0; t1 = -(double)bV[3 * (size_t)q + 1] - s1;
      t2 = -(double)bV[3 * (size_t)q + 2] - s2;
    }
    const T* m = Vinv + 6 * (size_t)q;
    const double m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5];
    double x0 = m0 * t0 + m1 * t1 + m2 * t2;
    double x1 = m1 * t0 + m3 * t1 + m4 * t2;
    double x2 = m2 * t0 + m4 * t1 + m5 * t2;
    if (add) {
      x0 += xV[3 * (size_t)q];
      x1 += xV[3 * (size_t)q + 1];
      x2 += xV[3 * (size_t)q + 2];
    }
    xV[3 * (size_t)q] = x0;
    xV[3 * (size_t)q + 1] = x1;
    xV[3 * (size_t)q + 2] = x2;
    const double mu = ctl->abs_;
    inc = x0 * x0 + x1 * x1 + x2 * x2;
    den = x0 * (mu * x0 - bv0) + x1 * (mu * x1 - bv1) + x2 * (mu * x2 - bv2);
  }
  const double a = block_sum256(inc, sh);
  const double b = block_sum256(den, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// ---------------------------------------------------------------------------
// Iterative refinement of the mixed-precision increment.  The FP32 Schur
// solve M approximates the damped normal equations (H + D) x = -b; each
// refinement step takes the FP64 residual
//   res = -b - (H + D) x,  H = sum_e J_e^T om_e J_e,  b = sum_e J_e^T om_e r_e,
//   D = diag(rel H_ii + abs) (SolverCholeskySchur.hpp:121-124,
//   MatrixDiagInv.hpp:36-40),
// evaluated matrix-free from the edges with J, r and om in FP64 at the
// current state (no FP32 block enters it), and adds M^-1 res to x.  The
// FP64 gradient b is written as well: the gain-ratio denominator
// x^T (mu x - b) uses it (DoubleSection.h:300-311).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ir_res_points(BAView v, const double* __restrict__ xU,
                                                       const double* __restrict__ xV, double* __restrict__ rV,
                                                       double* __restrict__ bV64, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.P) return;
  const int pv = v.pt_var[p];
  if (pv < 0) return;
  const double X0 = v.pts[3 * p], X1 = v.pts[3 * p + 1], X2 = v.pts[3 * p + 2];
  const double d0 = xV[3 * (size_t)pv], d1 = xV[3 * (size_t)pv + 1], d2 = xV[3 * (size_t)pv + 2];
  double b[3] = {0, 0, 0}, hq[3] = {0, 0, 0}, dg[3] = {0, 0, 0};
  const int e1 = v.pt_ptr[p + 1];
  for (int e = v.pt_ptr[p]; e < e1; e++) {
    const int c = v.pe_cam[e];
    double R[12];
#pragma unroll
    for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
    const Proj pr = project(v, R, X0, X1, X2, v.pe_obs[2 * e], v.pe_obs[2 * e + 1], v.pe_w[e]);
    double g0[3], g1[3];
    jac_pt<double>(v, pr, R, g0, g1);
    double q0 = g0[0] * d0 + g0[1] * d1 + g0[2] * d2;
    double q1 = g1[0] * d0 + g1[1] * d1 + g1[2] * d2;
    const int cv = v.cam_var[c];
    if (cv >= 0) {
      double j0[6], j1[6];
      jac_cam<double>(v, pr, j0, j1);
      const double* xc = xU + 6 * (size_t)cv;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        q0 += j0[i] * xc[i];
        q1 += j1[i] * xc[i];
      }
    }
    const double om = pr.om;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      b[k] += om * (g0[k] * pr.r0 + g1[k] * pr.r1);
      hq[k] += om * (g0[k] * q0 + g1[k] * q1);
      dg[k] += om * (g0[k] * g0[k] + g1[k] * g1[k]);
    }
  }
  const double rel = ctl->rel, ab = ctl->abs_;
  const double d[3] = {d0, d1, d2};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    rV[3 * (size_t)pv + k] = -b[k] - hq[k] - (rel * dg[k] + ab) * d[k];
    bV64[3 * (size_t)pv + k] = b[k];
  }
}

// one workgroup per variable camera; deterministic reduction
__global__ __launch_bounds__(256) void k_ir_res_cams(BAView v, const double* __restrict__ xU,
                                                     const double* __restrict__ xV, double* __restrict__ rU,
                                                     double* __restrict__ bU64, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][18];
  const int u = blockIdx.x;
  const int c = v.var_cam_user[u];
  double R[12];
#pragma unroll
  for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
  double xc[6];
#pragma unroll
  for (int i = 0; i < 6; i++) xc[i] = xU[6 * (size_t)u + i];
  double acc[18];
#pragma unroll
  for (int i = 0; i < 18; i++) acc[i] = 0.0;
  const int e1 = v.cam_ptr[u + 1];
  for (int e = v.cam_ptr[u] + threadIdx.x; e < e1; e += 256) {
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
  for (int i = 0; i < 18; i++) {
    const double s = wave_sum(acc[i]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int i = threadIdx.x;
    double t[3];
#pragma unroll
    for (int k = 0; k < 3; k++) t[k] = ((red[0][6 * k + i] + red[1][6 * k + i]) + red[2][6 * k + i]) + red[3][6 * k + i];
    const double rel = ctl->rel, ab = ctl->abs_;
    rU[6 * (size_t)u + i] = -t[0] - t[1] - (rel * t[2] + ab) * xc[i];
    bU64[6 * (size_t)u + i] = t[0];
  }
}

// rS_u = rU_u - sum_{k in W blocks of u} Y_k rV_{point(k)}; one workgroup per camera
template <typename T>
__global__ __launch_bounds__(256) void k_ir_reduce(BAView v, const T* __restrict__ Y, const double* __restrict__ rU,
                                                   const double* __restrict__ rV, double* __restrict__ rS,
                                                   DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][6];
  const int u = blockIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int i1 = v.cw_ptr[u + 1];
  for (int ii = v.cw_ptr[u] + threadIdx.x; ii < i1; ii += 256) {
    const int k = v.cw_idx[ii];
    const double* r = rV + 3 * (size_t)v.w_pt[k];
    const double r0 = r[0], r1 = r[1], r2 = r[2];
    const T* y = Y + 18 * (size_t)k;
#pragma unroll
    for (int a = 0; a < 6; a++) acc[a] += (double)y[3 * a] * r0 + (double)y[3 * a + 1] * r1 + (double)y[3 * a + 2] * r2;
  }
#pragma unroll
  for (int a = 0; a < 6; a++) {
    const double s = wave_sum(acc[a]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][a] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    rS[6 * (size_t)u + a] = rU[6 * (size_t)u + a] - (((red[0][a] + red[1][a]) + red[2][a]) + red[3][a]);
  }
}

// ---------------------------------------------------------------------------
// Implicit-function-theorem adjoint of the converged solve, per edge.  At x*
// the first-order condition b(x*, theta) = sum_e w_e psi_e J_e^T r_e = 0
// holds; with lambda = (db/dx)^-1 g for the loss gradient g,
//   dL/dz_e = -lambda^T db/dz_e = w_e (psi a + 2 psi'(s) (a . r) r),
//   dL/dw_e = -lambda^T db/dw_e = -psi (a . r),
// where a = J_c lambda_c + J_p lambda_p, s = |r|^2, psi = rho'(s) (1, or
// delta / sqrt(s) on the Huber branch) and psi' its derivative.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adjoint_edges(BAView v, const double* __restrict__ lamU,
                                                       const double* __restrict__ lamV, double* __restrict__ dobs,
                                                       double* __restrict__ dw) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.P) return;
  const double X0 = v.pts[3 * p], X1 = v.pts[3 * p + 1], X2 = v.pts[3 * p + 2];
  const int pv = v.pt_var[p];
  const int e1 = v.pt_ptr[p + 1];
  for (int e = v.pt_ptr[p]; e < e1; e++) {
    const int c = v.pe_cam[e];
    double R[12];
#pragma unroll
    for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
    const double we = v.pe_w[e];
    const Proj pr = project(v, R, X0, X1, X2, v.pe_obs[2 * e], v.pe_obs[2 * e + 1], we);
    double a0 = 0.0, a1 = 0.0;
    if (pv >= 0) {
      double g0[3], g1[3];
      jac_pt<double>(v, pr, R, g0, g1);
      const double* l = lamV + 3 * (size_t)pv;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        a0 += g0[k] * l[k];
        a1 += g1[k] * l[k];
      }
    }
    const int cv = v.cam_var[c];
    if (cv >= 0) {
      double j0[6], j1[6];
      jac_cam<double>(v, pr, j0, j1);
      const double* l = lamU + 6 * (size_t)cv;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        a0 += j0[i] * l[i];
        a1 += j1[i] * l[i];
      }
    }
    const double s = pr.r0 * pr.r0 + pr.r1 * pr.r1;
    double psi = 1.0, dpsi = 0.0;
    if (v.huber > 0.0 && s > v.huber * v.huber) {
      const double rs = sqrt(s);
      psi = v.huber / rs;
      dpsi = -0.5 * v.huber / (s * rs);
    }
    const double ar = a0 * pr.r0 + a1 * pr.r1;
    const int ue = v.pe_edge[e];
    dobs[2 * (size_t)ue] = we * (psi * a0 + 2.0 * dpsi * ar * pr.r0);
    dobs[2 * (size_t)ue + 1] = we * (psi * a1 + 2.0 * dpsi * ar * pr.r1);
    dw[ue] = -psi * ar;
  }
}

__global__ void k_axpby(long long n, double a, const double* x, double b, const double* y, double* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

__global__ void k_add(double* __restrict__ x, const double* __restrict__ z, long long n, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += z[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_xu_stats(int n, const double* __restrict__ xU, const T* __restrict__ bU,
                                                  double* __restrict__ part, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double inc = 0.0, den = 0.0;
  if (i < n) {
    const double x = xU[i];
    const double mu = ctl->abs_;
    inc = x * x;
    den = x * (mu * x - (double)bU[i]);
  }
  const double a = block_sum256(inc, sh);
  const double b = block_sum256(den, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// ||dx||^2, the rho denominator and the solve status; with lm_checks the
// small-increment test of LevenbergMarquardt.hpp:73-75 (1) or
// GaussNewton.hpp:72-87 (2)
__global__ __launch_bounds__(256) void k_step_finish(const double* __restrict__ part_xu, int nxu,
                                                     const double* __restrict__ part_xv, int nxv, DevCtl* ctl,
                                                     int gate, int lm_checks) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (int i = threadIdx.x; i < nxu; i += 256) { a += part_xu[2 * i]; b += part_xu[2 * i + 1]; }
  for (int i = threadIdx.x; i < nxv; i += 256) { c += part_xv[2 * i]; d += part_xv[2 * i + 1]; }
  a = block_sum256(a, sh);
  b = block_sum256(b, sh);
  c = block_sum256(c, sh);
  d = block_sum256(d, sh);
  if (threadIdx.x == 0) {
    const int ok = (ctl->fail_v == 0 && ctl->fail_s == 0 && ctl->wait_timeout == 0) ? 1 : 0;
    ctl->solve_ok = ok;
    ctl->incsq = a + c;
    ctl->den = b + d;
    ctl->total_solves++;
    if (lm_checks == 1) {
      if (ok && ctl->iter >= ctl->minNumIt && ctl->incsq <= ctl->epsIncSquare) {
        ctl->done = 1;
        ctl->reason = 1;  // SuccessSmallIncrement
      }
    } else if (lm_checks == 2) {
      if (!ok) {
        ctl->done = 1;
        ctl->reason = 4;  // SolverFailure
      } else if (ctl->iter >= ctl->minNumIt && ctl->incsq <= ctl->epsIncSquare) {
        ctl->done = 1;
        ctl->reason = 1;
      }
    }
  }
}

// oplus of the BA sample (BASchurPerfomances.cpp:163-170): cameras
// T <- exp(delta) T (R <- Exp(w) R, t <- Exp(w) t + V(w) u), points additive
__device__ __forceinline__ void so3_coeffs(double th2, double& A, double& B, double& C) {
  const double th = sqrt(th2);
  if (th < 1e-3) {
    A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    C = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
  } else {
    const double s = sin(th), h = sin(0.5 * th);
    A = s / th;
    B = 2.0 * h * h / th2;
    C = (th - s) / (th2 * th);
  }
}

__global__ __launch_bounds__(256) void k_oplus(BAView v, const double* __restrict__ xU,
                                               const double* __restrict__ xV, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < v.NU) {
    const int c = v.var_cam_user[idx];
    const double* d = xU + 6 * (size_t)idx;
    const double w0 = d[0], w1 = d[1], w2 = d[2];
    double A, B, Cc;
    so3_coeffs(w0 * w0 + w1 * w1 + w2 * w2, A, B, Cc);
    const double K[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
    double E[9], Vm[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
        const double id = (i == j) ? 1.0 : 0.0;
        E[i * 3 + j] = id + A * K[i * 3 + j] + B * k2;
        Vm[i * 3 + j] = id + B * K[i * 3 + j] + Cc * k2;
      }
    double* cm = v.cams + 12 * (size_t)c;
    double R[9], t[3];
#pragma unroll
    for (int i = 0; i < 9; i++) R[i] = cm[i];
#pragma unroll
    for (int i = 0; i < 3; i++) t[i] = cm[9 + i];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
      for (int j = 0; j < 3; j++)
        cm[i * 3 + j] = E[i * 3 + 0] * R[0 * 3 + j] + E[i * 3 + 1] * R[1 * 3 + j] + E[i * 3 + 2] * R[2 * 3 + j];
      cm[9 + i] = E[i * 3 + 0] * t[0] + E[i * 3 + 1] * t[1] + E[i * 3 + 2] * t[2] + Vm[i * 3 + 0] * d[3] +
                  Vm[i * 3 + 1] * d[4] + Vm[i * 3 + 2] * d[5];
    }
  } else if (idx < v.NU + v.NV) {
    const int q = idx - v.NU;
    const int p = v.var_pt_user[q];
#pragma unroll
    for (int i = 0; i < 3; i++) v.pts[3 * (size_t)p + i] += xV[3 * (size_t)q + i];
  }
}

__global__ __launch_bounds__(256) void k_copy_state(int nc, int np, double* __restrict__ cd, double* __restrict__ pd,
                                                    const double* __restrict__ cs, const double* __restrict__ ps,
                                                    DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) cd[i] = cs[i];
  else if (i < nc + np) pd[i - nc] = ps[i - nc];
  if (gate == GATE_ROLLBACK && i == 0) {
    // the flag is cleared by the last block to be safe against reordering:
    // every block reads it at entry, so clear only after the grid is done
  }
}

__global__ void k_clear_rollback(DevCtl* ctl) {
  if (threadIdx.x == 0) ctl->rollback_pending = 0;
}

__device__ void log_add(DevCtl* c, DevLog* log, int it, int subit, double chi2, double mu) {
  if (log && c->n_log < c->log_cap) {
    log[c->n_log].it = it;
    log[c->n_log].subit = subit;
    log[c->n_log].chi2 = chi2;
    log[c->n_log].mu = mu;
  }
  c->n_log++;
}

// after a full linearisation in the LM loop: the start of the solve
// (LevenbergMarquardt.hpp:41-61) or the end of an outer iteration (:132-141)
__global__ void k_lm_after_lin(DevCtl* c, DevLog* log) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->done || !c->need_lin) return;
  if (!c->started) {
    log_add(c, log, -1, -1, c->chi2, 0.0);
    c->started = 1;
    if (c->maxb <= c->epsBVector) {
      c->done = 1;
      c->reason = 0;  // SuccessAtStart
      c->need_lin = 0;
      return;
    }
    c->oldChi2 = c->chi2;
    c->mu = c->tau * c->maxdiag;
    c->curV = c->initV;
    c->iter = 0;
  } else {
    c->oldChi2 = c->chi2;
    c->iter++;
  }
  c->subiter = 0;
  c->need_lin = 0;
  if (c->iter >= c->maxNumIt) {
    c->done = 1;
    c->reason = 5;  // MaxIterationReached
    return;
  }
  c->rel = 0.0;
  c->abs_ = c->mu;
}

// LevenbergMarquardt.hpp:77-130: log, rho, accept/reject, damping schedule
__global__ void k_lm_control(DevCtl* c, DevLog* log) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->done) return;
  double rho = -1.0;
  if (c->solve_ok) {
    log_add(c, log, c->iter, c->subiter, c->chi2, c->mu);
    rho = (c->oldChi2 - c->chi2) / c->den;
  } else {
    log_add(c, log, c->iter, c->subiter, -1.0, c->mu);
  }
  c->rho = rho;
  c->accepted = 0;
  if (c->solve_ok && rho > 0) {
    if (c->iter > c->minNumIt && c->maxb <= c->epsBVector) {
      c->done = 1;
      c->reason = 2;  // SuccessEpsB
      return;
    }
    if (c->iter > c->minNumIt && c->chi2 < c->epsChi2) {
      c->done = 1;
      c->reason = 3;  // SuccessSmallResiduals
      return;
    }
    const double fc = 2.0 * rho - 1.0;
    double alpha = fmin(c->maxScaleFactor, 1.0 - fc * fc * fc);
    alpha = fmax(c->minScaleFactor, alpha);
    c->mu = c->mu * alpha;
    c->curV = c->initV;
    c->accepted = 1;
  } else {
    c->rollback_pending = 1;
    c->mu = c->mu * c->curV;
    c->curV = c->curV * c->penaltyV;
  }
  c->subiter++;
  if (c->subiter >= c->maxNumSubIt) {
    c->done = 1;
    c->reason = 6;  // MaxSubIterationReached
  }
  if (c->accepted && !c->done) c->need_lin = 1;
  c->rel = 0.0;
  c->abs_ = c->mu;
}

// Gauss-Newton: after the initial linearisation (GaussNewton.hpp:37-58)
__global__ void k_gn_after_lin(DevCtl* c, DevLog* log) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->done || c->started) return;
  log_add(c, log, -1, -1, c->chi2, 0.0);
  c->started = 1;
  c->need_lin = 0;
  if (c->maxb <= c->epsBVector) {
    c->done = 1;
    c->reason = 0;
    return;
  }
  c->iter = 0;
  if (c->iter >= c->maxNumIt) { c->done = 1; c->reason = 5; return; }
  c->rel = c->gn_rel;
  c->abs_ = c->gn_abs;
}

// Gauss-Newton: after oplus + update(true) (GaussNewton.hpp:97-124)
__global__ void k_gn_after_step(DevCtl* c, DevLog* log) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->done) return;
  log_add(c, log, c->iter, 0, c->chi2, 0.0);
  if (c->iter > c->minNumIt && c->maxb <= c->epsBVector) { c->done = 1; c->reason = 2; return; }
  if (c->iter > c->minNumIt && c->chi2 < c->epsChi2) { c->done = 1; c->reason = 3; return; }
  c->iter++;
  if (c->iter >= c->maxNumIt) { c->done = 1; c->reason = 5; return; }
  c->rel = c->gn_rel;
  c->abs_ = c->gn_abs;
}

__global__ void k_set_damping(DevCtl* c, double rel, double ab) {
  if (threadIdx.x == 0) {
    c->rel = rel;
    c->abs_ = ab;
    c->fail_v = 0;
    c->fail_s = 0;
    c->wait_timeout = 0;
  }
}

__global__ void k_begin_step(DevCtl* c, int gate) {
  if (!gate_open(c, gate)) return;
  if (threadIdx.x == 0) {
    c->fail_v = 0;
    c->fail_s = 0;
    c->wait_timeout = 0;
  }
}

template <typename T>
__global__ void k_zero(T* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (T)0;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
void launch_lin_points(const BAView& v, bool hess, T* V, T* bV, T* W, const Scratch& s, DevCtl* ctl, int gate,
                       hipStream_t st) {
  if (v.P == 0) return;
  if (hess)
    hipLaunchKernelGGL((k_lin_points<T, true>), dim3(nblocks(v.P, 256)), dim3(256), 0, st, v, V, bV, W,
                       s.part_chi2, s.maxes, ctl, gate);
  else
    hipLaunchKernelGGL((k_lin_points<T, false>), dim3(nblocks(v.P, 256)), dim3(256), 0, st, v, V, bV, W,
                       s.part_chi2, s.maxes, ctl, gate);
}

template <typename T>
void launch_lin_cams(const BAView& v, T* U, T* bU, const Scratch& s, DevCtl* ctl, int gate, hipStream_t st) {
  if (v.NU == 0) return;
  hipLaunchKernelGGL((k_lin_cams<T>), dim3(v.NU), dim3(256), 0, st, v, U, bU, s.maxes, ctl, gate);
}

void launch_reset_lin(const Scratch& s, DevCtl* ctl, int gate, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_lin, dim3(1), dim3(64), 0, st, s.maxes, ctl, gate);
}

void launch_finalize_lin(const Scratch& s, DevCtl* ctl, int gate, int full, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize_lin, dim3(1), dim3(256), 0, st, s.part_chi2, s.nblk_pts, s.maxes, ctl, gate, full);
}

template <typename T>
void launch_vinv_y(const BAView& v, const T* V, const T* bV, const T* W, T* Vinv, T* Y, T* tW, DevCtl* ctl,
                   int gate, hipStream_t st) {
  if (v.NV == 0) return;
  hipLaunchKernelGGL((k_vinv_y<T>), dim3(nblocks(v.NV, 256)), dim3(256), 0, st, v, V, bV, W, Vinv, Y, tW, ctl,
                     gate);
}

template <typename T>
void launch_s_diag(const BAView& v, const T* U, const T* bU, const T* Y, const T* W, const T* tW, T* Sd, T* bS,
                   DevCtl* ctl, int gate, hipStream_t st) {
  if (v.NU == 0) return;
  hipLaunchKernelGGL((k_s_diag<T>), dim3(v.NU), dim3(256), 0, st, v, U, bU, Y, W, tW, Sd, bS, ctl, gate);
}

template <typename T>
void launch_s_tiles(const TileView& tv, const int* tile_order, const int* tile_blk_ptr, const int* tile_blk,
                    const int* sb_i, const int* sb_j, const int* sb_ptr, const int* sb_pairs, const T* Sd,
                    const T* bS, const T* Y, const T* W, T* tiles, DevCtl* ctl, int gate, hipStream_t st) {
  if (tv.TB == 64)
    hipLaunchKernelGGL((k_s_tiles<T, 64>), dim3(tv.ntiles), dim3(256), 0, st, tv, tile_order, tile_blk_ptr,
                       tile_blk, sb_i, sb_j, sb_ptr, sb_pairs, Sd, bS, Y, W, tiles, ctl, gate);
  else if (tv.TB == 32)
    hipLaunchKernelGGL((k_s_tiles<T, 32>), dim3(tv.ntiles), dim3(256), 0, st, tv, tile_order, tile_blk_ptr,
                       tile_blk, sb_i, sb_j, sb_ptr, sb_pairs, Sd, bS, Y, W, tiles, ctl, gate);
}

template <typename T>
void launch_backsub(const BAView& v, const T* W, const T* Vinv, const T* bV, const double* xU, double* xV,
                    const Scratch& s, DevCtl* ctl, int gate, hipStream_t st, const double* rhs_v, int add,
                    const double* bden) {
  if (v.NV == 0) return;
  hipLaunchKernelGGL((k_backsub<T>), dim3(nblocks(v.NV, 256)), dim3(256), 0, st, v, W, Vinv, bV, xU, xV,
                     s.part_xv, ctl, gate, rhs_v, add, bden);
}

void launch_ir_residual(const BAView& v, const double* xU, const double* xV, double* rU, double* rV, double* bU64,
                        double* bV64, DevCtl* ctl, int gate, hipStream_t st) {
  if (v.P > 0)
    hipLaunchKernelGGL(k_ir_res_points, dim3(nblocks(v.P, 256)), dim3(256), 0, st, v, xU, xV, rV, bV64, ctl, gate);
  if (v.NU > 0)
    hipLaunchKernelGGL(k_ir_res_cams, dim3(v.NU), dim3(256), 0, st, v, xU, xV, rU, bU64, ctl, gate);
}

template <typename T>
void launch_ir_reduce(const BAView& v, const T* Y, const double* rU, const double* rV, double* rS, DevCtl* ctl,
                      int gate, hipStream_t st) {
  if (v.NU == 0) return;
  hipLaunchKernelGGL((k_ir_reduce<T>), dim3(v.NU), dim3(256), 0, st, v, Y, rU, rV, rS, ctl, gate);
}

void launch_axpby(long long n, double a, const double* x, double b, const double* y, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, x, b, y, out);
}

void launch_adjoint_edges(const BAView& v, const double* lamU, const double* lamV, double* dobs, double* dw,
                          hipStream_t st) {
  if (v.P == 0) return;
  hipLaunchKernelGGL(k_adjoint_edges, dim3(nblocks(v.P, 256)), dim3(256), 0, st, v, lamU, lamV, dobs, dw);
}

void launch_add(double* x, const double* z, long long n, DevCtl* ctl, int gate, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, z, n, ctl, gate);
}

template <typename T>
void launch_xu_stats(int NU, const double* xU, const T* bU, const Scratch& s, DevCtl* ctl, int gate,
                     hipStream_t st) {
  if (NU == 0) return;
  hipLaunchKernelGGL((k_xu_stats<T>), dim3(nblocks(6LL * NU, 256)), dim3(256), 0, st, 6 * NU, xU, bU, s.part_xu,
                     ctl, gate);
}

void launch_step_finish(const Scratch& s, DevCtl* ctl, int gate, int lm_checks, hipStream_t st) {
  hipLaunchKernelGGL(k_step_finish, dim3(1), dim3(256), 0, st, s.part_xu, s.nblk_cams, s.part_xv, s.nblk_xv, ctl,
                     gate, lm_checks);
}

void launch_oplus(const BAView& v, const double* xU, const double* xV, DevCtl* ctl, int gate, hipStream_t st) {
  if (v.NU + v.NV == 0) return;
  hipLaunchKernelGGL(k_oplus, dim3(nblocks(v.NU + v.NV, 256)), dim3(256), 0, st, v, xU, xV, ctl, gate);
}

void launch_copy_state(const BAView& v, double* cams_dst, double* pts_dst, const double* cams_src,
                       const double* pts_src, DevCtl* ctl, int gate, hipStream_t st) {
  const long long n = 12LL * v.C + 3LL * v.P;
  hipLaunchKernelGGL(k_copy_state, dim3(nblocks(n, 256)), dim3(256), 0, st, 12 * v.C, 3 * v.P, cams_dst, pts_dst,
                     cams_src, pts_src, ctl, gate);
  if (gate == GATE_ROLLBACK) hipLaunchKernelGGL(k_clear_rollback, dim3(1), dim3(64), 0, st, ctl);
}

void launch_lm_after_lin(DevCtl* ctl, DevLog* log, hipStream_t st) {
  hipLaunchKernelGGL(k_lm_after_lin, dim3(1), dim3(64), 0, st, ctl, log);
}
void launch_lm_control(DevCtl* ctl, DevLog* log, hipStream_t st) {
  hipLaunchKernelGGL(k_lm_control, dim3(1), dim3(64), 0, st, ctl, log);
}
void launch_gn_after_lin(DevCtl* ctl, DevLog* log, hipStream_t st) {
  hipLaunchKernelGGL(k_gn_after_lin, dim3(1), dim3(64), 0, st, ctl, log);
}
void launch_gn_after_step(DevCtl* ctl, DevLog* log, hipStream_t st) {
  hipLaunchKernelGGL(k_gn_after_step, dim3(1), dim3(64), 0, st, ctl, log);
}
void launch_set_damping(DevCtl* ctl, double rel, double abs_, hipStream_t st) {
  hipLaunchKernelGGL(k_set_damping, dim3(1), dim3(64), 0, st, ctl, rel, abs_);
}
void launch_begin_step(DevCtl* ctl, int gate, hipStream_t st) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, st, ctl, gate);
}

template <typename T>
void launch_zero(T* p, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_zero<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, p, n);
}

#define NSN_INST(T)                                                                                          \
  template void launch_lin_points<T>(const BAView&, bool, T*, T*, T*, const Scratch&, DevCtl*, int,          \
                                     hipStream_t);                                                           \
  template void launch_lin_cams<T>(const BAView&, T*, T*, const Scratch&, DevCtl*, int, hipStream_t);       \
  template void launch_vinv_y<T>(const BAView&, const T*, const T*, const T*, T*, T*, T*, DevCtl*, int,      \
                                 hipStream_t);                                                               \
  template void launch_s_diag<T>(const BAView&, const T*, const T*, const T*, const T*, const T*, T*, T*,    \
                                 DevCtl*, int, hipStream_t);                                                 \
  template void launch_s_tiles<T>(const TileView&, const int*, const int*, const int*, const int*,           \
                                  const int*, const int*, const int*, const T*, const T*, const T*,          \
                                  const T*, T*, DevCtl*, int, hipStream_t);                                  \
  template void launch_backsub<T>(const BAView&, const T*, const T*, const T*, const double*, double*,       \
                                  const Scratch&, DevCtl*, int, hipStream_t, const double*, int,             \
                                  const double*);                                                            \
  template void launch_ir_reduce<T>(const BAView&, const T*, const double*, const double*, double*, DevCtl*,  \
                                    int, hipStream_t);                                                       \
  template void launch_xu_stats<T>(int, const double*, const T*, const Scratch&, DevCtl*, int, hipStream_t); \
  template void launch_zero<T>(T*, long long, hipStream_t);

NSN_INST(double)
NSN_INST(float)
template void launch_zero<int>(int*, long long, hipStream_t);

}  // namespace nsn
