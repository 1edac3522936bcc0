// Bundle-adjustment kernels of the Schur LM step on gfx950 (wave64).
//
// Data flow of one LM sub-iteration (nelson/SolverCholeskySchur.hpp:91-157,
// nelson/LevenbergMarquardt.hpp:67-130):
//   k_lin_points / k_lin_cams  : residuals, Jacobians, V, bV, W (point-major)
//                                and U, bU (camera-major)        [update(true)]
//   k_vinv_y                   : damped V^-1, Z = W L^-T (V = L L^T), t = W V^-1 (-bV)
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

#include "ba_device.h"
#include "kernels.h"
#include "tile_device.h"

namespace nsn {

// ---------------------------------------------------------------------------
// linearisation, point-major pass: V, bV, W and the chi2 of every edge
// (nelson/BaseSection.cpp:88-169 + updateH22Block/updateH12Block), with
// kLinLanes lanes per point (one lane per point left lines of the W blocks
// partly written for the length of its edge loop, and they were filled from
// memory: 200 MB of extra reads at 200 000 points); DUP: some camera
// observes some point more than once
// ---------------------------------------------------------------------------
template <typename T, bool HESS, bool DUP>
__global__ __launch_bounds__(256) void k_lin_points(BAView v, T* __restrict__ V, T* __restrict__ bV,
                                                    T* __restrict__ W, double* __restrict__ part_chi2,
                                                    unsigned long long* __restrict__ maxes, DevCtl* ctl,
                                                    int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[8];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  double chi2 = 0.0, mb = 0.0, md = 0.0;
  lin_point_lanes<T, HESS, kLinLanes, DUP>(v, g / kLinLanes, g % kLinLanes, V, bV, W, chi2, mb, md);
  if (v.cam_prior) {  // the cameras' unary priors (cost)
    for (int c = g; c < v.C; c += gridDim.x * blockDim.x) {
      const int k = v.cam_prior[c];
      double r[6];
      if (k >= 0) chi2 += cam_prior_eval<false>(v, c, k, r, nullptr);
    }
  }
  double s = block_sum256(chi2, sh);
  if (threadIdx.x == 0) part_chi2[blockIdx.x] = s;
  if (HESS) block_max2_atomic(mb, md, sh, maxes + 1, maxes + 3);
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
  __shared__ double pri[27];  // the camera's unary prior: J^T J (21 packed), J^T r
  const int u = xcd_runs(blockIdx.x, gridDim.x);
  const int c = v.var_cam_user[u];
  if (threadIdx.x == 0) {
    for (int i = 0; i < 27; i++) pri[i] = 0.0;
    const int k = v.cam_prior ? v.cam_prior[c] : -1;
    if (k >= 0) {
      double r[6], J[36];
      cam_prior_eval<true>(v, c, k, r, J);
      for (int i = 0; i < 6; i++) {
        for (int j = i; j < 6; j++) {
          double h = 0.0;
          for (int m = 0; m < 6; m++) h += J[m * 6 + i] * J[m * 6 + j];
          pri[up6(i, j)] = h;
        }
        double g = 0.0;
        for (int m = 0; m < 6; m++) g += J[m * 6 + i] * r[m];
        pri[21 + i] = g;
      }
    }
  }
  double R[kCamRec];
  load_cam(v, c, R);
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
    const double s = (((red[0][i] + red[1][i]) + red[2][i]) + red[3][i]) + pri[i];
    if (i < 21) U[21 * (size_t)u + i] = (T)s;
    else bU[6 * (size_t)u + (i - 21)] = (T)s;
  }
  if (threadIdx.x == 0) {
    double mb = 0.0, md = 0.0;
    for (int i = 0; i < 6; i++) {
      mb = fmax(mb, fabs((double)(T)((((red[0][21 + i] + red[1][21 + i]) + red[2][21 + i]) + red[3][21 + i]) +
                                      pri[21 + i])));
      const int d = up6(i, i);
      md = fmax(md, fabs((double)(T)((((red[0][d] + red[1][d]) + red[2][d]) + red[3][d]) + pri[d])));
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
// the W V^-1 products (MatrixWVinvMultiplier::multiply,
// MatrixWVinvMultiplier.hpp:23-106) in the symmetric form Z = W L^-T, where
// V = L L^T is the damped 3x3 block's Cholesky factor (FP64): the S terms
// W_a V^-1 W_b^T = Z_a Z_b^T then read one array; and t = W V^-1 (-bV) (the
// W * V^-1 bV term of bS, SolverCholeskySchur.hpp:114).
// ---------------------------------------------------------------------------
// One thread per W block k (threads 0 .. nW-1), then one per variable point
// (nW .. nW+NV-1): a block's thread forms its point's factor itself (a few
// dozen operations, the point's V and bV shared with its neighbours in the
// wave) and writes Z_k, so a wave writes 64 consecutive blocks at once; a
// point's thread writes V_q^-1 and y_q = L_q^-1 (-bV_q) and counts a failed
// block.  The bS terms t_k = W_k V^-1 (-bV) = Z_k y_q are formed where they
// are summed (k_s_diag): 24 bytes per point instead of 48 per block.  (One thread per
// point writing all of its blocks left each line partly written for the
// length of its loop: the partial lines were filled from memory, about
// 200 MB of extra reads at the 1000-camera problem.)
template <typename T>
__global__ __launch_bounds__(256) void k_vinv_y(BAView v, const T* __restrict__ V, const T* __restrict__ bV,
                                                const T* __restrict__ W, T* __restrict__ Vinv,
                                                T* __restrict__ Z, double* __restrict__ yV, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double rel = ctl->rel, ab = ctl->abs_;
  if (g < v.nW) {
    const int k = (int)g;
    const VFactor<T, true> f = v_factor<T, true>(V, bV, v.w_pt[k], rel, ab);
    z_block<T, true>(f, W, Z, k);
  } else if (g < (long long)v.nW + v.NV) {
    const int q = (int)(g - v.nW);
    const VFactor<T, true> f = v_factor<T, true>(V, bV, q, rel, ab);
#pragma unroll
    for (int i = 0; i < 6; i++) Vinv[6 * (size_t)q + i] = (T)f.vi[i];
    yV[3 * (size_t)q] = f.y0;
    yV[3 * (size_t)q + 1] = f.y1;
    yV[3 * (size_t)q + 2] = f.y2;
    if (!f.ok) atomicAdd(&ctl->fail_v, 1);
  }
}

// ---------------------------------------------------------------------------
// diagonal S blocks and bS, one workgroup per variable camera:
//   S_uu = U_u - sum_k Z_k Z_k^T, then S_ii += rel * U_ii + abs
//   bS_u = -bU_u - sum_k t_k
// (SolverCholeskySchur.hpp:114-124; MatrixWWtMultiplier.hpp:139-153 for the
// diagonal pairs)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_s_diag(BAView v, const T* __restrict__ U, const T* __restrict__ bU,
                                                const T* __restrict__ Z, const double* __restrict__ yV,
                                                T* __restrict__ Sd,
                                                T* __restrict__ bS, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][27];
  const int u = xcd_runs(blockIdx.x, gridDim.x);
  T acc[27];
#pragma unroll
  for (int i = 0; i < 27; i++) acc[i] = (T)0;
  const int i1 = v.cw_ptr[u + 1];
  for (int ii = v.cw_ptr[u] + threadIdx.x; ii < i1; ii += 256) {
    const int k = v.cw_idx[ii], q = v.cw_pt[ii];
    T z[18];
#pragma unroll
    for (int i = 0; i < 18; i++) z[i] = Z[18 * (size_t)k + i];
    const double y0 = yV[3 * (size_t)q], y1 = yV[3 * (size_t)q + 1], y2 = yV[3 * (size_t)q + 2];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
      for (int b = a; b < 6; b++)
        acc[up6(a, b)] += z[3 * a] * z[3 * b] + z[3 * a + 1] * z[3 * b + 1] + z[3 * a + 2] * z[3 * b + 2];
#pragma unroll
    for (int a = 0; a < 6; a++)  // t_k = Z_k y_q
      acc[21 + a] += (T)((double)z[3 * a] * y0 + (double)z[3 * a + 1] * y1 + (double)z[3 * a + 2] * y2);
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
// S tiles: one workgroup per part of a lower TB x TB tile of [S; bS^T]: a
// rectangle of the tile (rows [rlo, rhi), columns [clo, chi), local) cut at
// camera boundaries, the whole tile when S has enough tiles to fill the
// device.  The part is built in LDS (zeros, identity padding, the bS row,
// diagonal blocks from Sd, off-diagonal blocks S_ij = -sum_p Z_ip Z_jp^T over
// their pair lists, MatrixWWtMultiplier.hpp:124-153) and written once,
// coalesced along the tile's columns.
// An upper block (i, j), i <= j, is stored transposed at rows 6j.., cols 6i..
// A wave builds one block at a time: lane (s, a), s < kSlots, a < 6, sums row
// a of the products of the pairs p = s, s + kSlots, ... (kSlots pairs in
// flight per step), and the kSlots partial rows are added in slot order
// through LDS (deterministic).
// ---------------------------------------------------------------------------
constexpr int kSlots = 10;

template <typename T, int TB>
__global__ __launch_bounds__(256) void k_s_tiles(TileView tv, const int* __restrict__ tile_order,
                                                 const int* __restrict__ part_tile,
                                                 const int* __restrict__ part_rect,
                                                 const int* __restrict__ part_ptr,
                                                 const int* __restrict__ part_blk, const int* __restrict__ sb_i,
                                                 const int* __restrict__ sb_j, const int* __restrict__ sb_ptr,
                                                 const int* __restrict__ sb_pairs, const T* __restrict__ Sd,
                                                 const T* __restrict__ bS, const T* __restrict__ Z,
                                                 T* __restrict__ tiles, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ T lt[TB * TB];
  __shared__ T red[4][kSlots * 36];
  const int q = xcd_order<8>(blockIdx.x, gridDim.x);
  const int t = part_tile[q];
  const int rlo = part_rect[4 * q], rhi = part_rect[4 * q + 1];
  const int clo = part_rect[4 * q + 2], chi = part_rect[4 * q + 3];
  const int hr = rhi - rlo, npe = hr * (chi - clo);
  const int ti = tile_order[2 * t], tj = tile_order[2 * t + 1];
  const int r0 = ti * TB, c0 = tj * TB, n = tv.n;
  for (int e = threadIdx.x; e < npe; e += 256) {
    const int lr = rlo + e % hr, lc = clo + e / hr;
    const int R = r0 + lr, Cc = c0 + lc;
    T val = (T)0;
    if (R == Cc && R >= n) val = (T)1;
    else if (R == n && Cc < n) val = bS[Cc];
    lt[e] = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slot = lane / 6, ra = lane - 6 * (lane / 6);  // partial-sum lanes (slot < kSlots)
  const int a = lane / 6, bb = lane % 6;                   // output lanes (lane < 36)
  T* rd = red[wv];
  const int b1 = part_ptr[q + 1];
  for (int b = part_ptr[q] + wv; b < b1; b += 4) {
    const int sb = part_blk[b];
    const int i = sb_i[sb], j = sb_j[sb];
    const int R = 6 * j + bb, Cc = 6 * i + a;
    const bool inside = lane < 36 && R >= r0 + rlo && R < r0 + rhi && Cc >= c0 + clo && Cc < c0 + chi;
    const int at = (Cc - c0 - clo) * hr + (R - r0 - rlo);
    if (i == j) {
      if (inside && bb >= a) lt[at] = Sd[21 * (size_t)i + up6(a, bb)];
      continue;
    }
    T acc[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
    if (slot < kSlots) {
      const int p0 = sb_ptr[sb], p1 = sb_ptr[sb + 1];
#pragma unroll 2
      for (int p = p0 + slot; p < p1; p += kSlots) {
        const int ka = sb_pairs[2 * p], kb = sb_pairs[2 * p + 1];
        const T* y = Z + 18 * (size_t)ka + 3 * ra;
        const T* w = Z + 18 * (size_t)kb;
        const T y0 = y[0], y1 = y[1], y2 = y[2];
#pragma unroll
        for (int c = 0; c < 6; c++) acc[c] += y0 * w[3 * c] + y1 * w[3 * c + 1] + y2 * w[3 * c + 2];
      }
#pragma unroll
      for (int c = 0; c < 6; c++) rd[lane * 6 + c] = acc[c];
    }
    wave_lds_sync();
    if (inside) {
      T sum = (T)0;
#pragma unroll
      for (int s2 = 0; s2 < kSlots; s2++) sum += rd[(s2 * 6 + a) * 6 + bb];
      lt[at] = -sum;
    }
    wave_lds_sync();
  }
  __syncthreads();
  T* out = tiles + (size_t)t * TB * TB;
  for (int e = threadIdx.x; e < npe; e += 256) out[(clo + e / hr) * TB + rlo + e % hr] = lt[e];
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
                                                 int add, const double* __restrict__ bden,
                                                 double* __restrict__ accU, int nacc) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double sh[4];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // refinement: the camera correction is added to the solution here
  // (accU += xU, grid-stride), saving a launch
  for (int g = t; g < nacc; g += gridDim.x * blockDim.x) accU[g] += xU[g];
  // kLinLanes lanes per point share its W blocks (a wave reads 16 points'
  // blocks about 4 at a time each), the sums closed by shuffles
  const int q = t / kLinLanes, l = t % kLinLanes;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (q < v.NV) wt_xu<T, kLinLanes>(v, q, l, W, xU, s0, s1, s2);
#pragma unroll
  for (int o = 1; o < kLinLanes; o <<= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
    s2 += __shfl_xor(s2, o);
  }
  double inc = 0.0, den = 0.0;
  if (q < v.NV && l == 0) backsub_finish<T>(v, q, Vinv, bV, xV, ctl->abs_, rhs_v, add, bden, s0, s1, s2, inc, den);
  const double a = block_sum256(inc, sh);
  const double b = block_sum256(den, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// the FP64 residual of the damped normal equations, res = -b - (H + D) x,
// matrix free from the edges, in one launch: the first nbp workgroups take
// the points (one thread per point), the next NU one variable camera each
__global__ __launch_bounds__(256) void k_ir_res(BAView v, const double* __restrict__ xU,
                                                const double* __restrict__ xV, double* __restrict__ rU,
                                                double* __restrict__ rV, double* __restrict__ bU64,
                                                double* __restrict__ bV64, int nbp, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  __shared__ double red[4][18];
  __shared__ double pri[18];
  __shared__ double shm[8];
  double mr = 0.0, mb = 0.0;  // max |r|, |b| of this thread's entries
  if ((int)blockIdx.x < nbp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < v.P) ir_res_point(v, xU, xV, rV, bV64, ctl, p, mr, mb);
  } else {
    ir_res_cam(v, xU, xV, rU, bU64, ctl, (int)blockIdx.x - nbp, red, pri, mr, mb);
  }
  // the residual's and the gradient's largest entries (the refinement's
  // stopping test), one atomic each per workgroup
  block_max2_atomic(mr, mb, shm, &ctl->ir_maxr, &ctl->ir_maxb);
}

__global__ void k_ir_decide(DevCtl* c, int gate, double tol) {
  if (!gate_open(c, gate)) return;
  if (threadIdx.x == 0) {
    const double mr = __longlong_as_double((long long)c->ir_maxr);
    const double mb = __longlong_as_double((long long)c->ir_maxb);
    const int more = tol <= 0.0 || mr > tol * mb;
    c->ir_more = more;
    c->ir_count += more;
    c->ir_maxr = 0ull;
    c->ir_maxb = 0ull;
  }
}

// rS_u = rU_u - sum_{k in W blocks of u} W_k V^-1 rV_{point(k)}; one workgroup per camera
template <typename T>
__global__ __launch_bounds__(256) void k_ir_reduce(BAView v, const T* __restrict__ W, const T* __restrict__ Vinv,
                                                   const double* __restrict__ rU, const double* __restrict__ rV,
                                                   double* __restrict__ rS, double* u0, double* u1, double* u2,
                                                   double* u3, int nunset, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  {
    // the solutions and destination sums of the two triangular solves that
    // follow start unset (nunset <= 256 NU)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nunset) {
      const double u = __builtin_bit_cast(double, kUnsetBits);
      u0[g] = u;
      u1[g] = u;
      u2[g] = u;
      u3[g] = u;
    }
  }
  __shared__ double red[4][6];
  const int u = xcd_runs(blockIdx.x, gridDim.x);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int i1 = v.cw_ptr[u + 1];
  for (int ii = v.cw_ptr[u] + threadIdx.x; ii < i1; ii += 256) {
    const int k = v.cw_idx[ii], q = v.w_pt[k];
    const double* r = rV + 3 * (size_t)q;
    const T* m = Vinv + 6 * (size_t)q;
    const double r0 = r[0], r1 = r[1], r2 = r[2];
    // V^-1 r (symmetric, upper triangle stored row by row)
    const double g0 = (double)m[0] * r0 + (double)m[1] * r1 + (double)m[2] * r2;
    const double g1 = (double)m[1] * r0 + (double)m[3] * r1 + (double)m[4] * r2;
    const double g2 = (double)m[2] * r0 + (double)m[4] * r1 + (double)m[5] * r2;
    const T* w = W + 18 * (size_t)k;
#pragma unroll
    for (int a = 0; a < 6; a++) acc[a] += (double)w[3 * a] * g0 + (double)w[3 * a + 1] * g1 + (double)w[3 * a + 2] * g2;
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
    double R[kCamRec];
    load_cam(v, c, R);
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

// step of the exact-Hessian refinement of the adjoint, on the device:
// eps = 1e-6 / max |lambda| (0 when lambda is 0 or not finite), one block
__global__ __launch_bounds__(256) void k_adj_eps(const double* __restrict__ lu, long long nu,
                                                 const double* __restrict__ lv, long long nv, double* eps) {
  __shared__ double sh[4];
  double m = 0.0;
  for (long long i = threadIdx.x; i < nu; i += 256) m = fmax(m, fabs(lu[i]));
  for (long long i = threadIdx.x; i < nv; i += 256) m = fmax(m, fabs(lv[i]));
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    eps[0] = (mx > 0.0 && isfinite(mx)) ? 1e-6 / mx : 0.0;
  }
}

// t = sgn * eps * lambda
__global__ void k_adj_offset(const double* __restrict__ lu, long long nu, const double* __restrict__ lv,
                             long long nv, const double* __restrict__ eps, double sgn, double* __restrict__ tu,
                             double* __restrict__ tv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double e = sgn * eps[0];
  if (i < nu) tu[i] = e * lu[i];
  if (i < nv) tv[i] = e * lv[i];
}

// r = g - (bp - bm) / (2 eps) into bp (0 when eps is 0: no correction)
__global__ void k_adj_resid(const double* __restrict__ gu, const double* __restrict__ bmu, double* __restrict__ bpu,
                            long long nu, const double* __restrict__ gv, const double* __restrict__ bmv,
                            double* __restrict__ bpv, long long nv, const double* __restrict__ eps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double e = eps[0];
  const double c = e > 0.0 ? -0.5 / e : 0.0;
  if (i < nu) bpu[i] = e > 0.0 ? gu[i] + c * (bpu[i] - bmu[i]) : 0.0;
  if (i < nv) bpv[i] = e > 0.0 ? gv[i] + c * (bpv[i] - bmv[i]) : 0.0;
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

__global__ __launch_bounds__(256) void k_oplus(BAView v, const double* __restrict__ xU,
                                               const double* __restrict__ xV, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < v.NU) oplus_cam(v, idx, xU);
  else if (idx < v.NU + v.NV) oplus_pt(v, idx - v.NU, xV);
}

__global__ __launch_bounds__(256) void k_copy_state(int nc, int np, double* __restrict__ cd, double* __restrict__ pd,
                                                    const double* __restrict__ cs, const double* __restrict__ ps,
                                                    DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) cd[i] = cs[i];
  else if (i < nc + np) pd[i - nc] = ps[i - nc];
}

__global__ void k_clear_rollback(DevCtl* ctl) {
  if (threadIdx.x == 0) ctl->rollback_pending = 0;
}

__global__ void k_lm_after_lin(DevCtl* c, DevLog* log) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lm_after_lin(c, log);
}

__global__ void k_lm_control(DevCtl* c, DevLog* log) {
  if (threadIdx.x == 0 && blockIdx.x == 0) lm_control(c, log);
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
    c->ir_more = 1;
    c->ir_maxr = 0ull;
    c->ir_maxb = 0ull;
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
  // always launched (nsn_problem_create requires n_points >= 1): the camera
  // priors' cost is added here too, so it is counted with every point fixed
  const unsigned g = nblocks((long long)v.P * kLinLanes, 256);  // = s.nblk_pts
  if (hess && v.w_dup)
    hipLaunchKernelGGL((k_lin_points<T, true, true>), dim3(g), dim3(256), 0, st, v, V, bV, W, s.part_chi2, s.maxes,
                       ctl, gate);
  else if (hess)
    hipLaunchKernelGGL((k_lin_points<T, true, false>), dim3(g), dim3(256), 0, st, v, V, bV, W, s.part_chi2, s.maxes,
                       ctl, gate);
  else
    hipLaunchKernelGGL((k_lin_points<T, false, false>), dim3(g), dim3(256), 0, st, v, V, bV, W, s.part_chi2,
                       s.maxes, ctl, gate);
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
void launch_vinv_y(const BAView& v, const T* V, const T* bV, const T* W, T* Vinv, T* Z, double* yV, DevCtl* ctl,
                   int gate, hipStream_t st) {
  if (v.NV == 0) return;
  hipLaunchKernelGGL((k_vinv_y<T>), dim3(nblocks(v.nW + v.NV, 256)), dim3(256), 0, st, v, V, bV, W, Vinv, Z, yV,
                     ctl, gate);
}

template <typename T>
void launch_s_diag(const BAView& v, const T* U, const T* bU, const T* Z, const double* yV, T* Sd, T* bS, DevCtl* ctl,
                   int gate, hipStream_t st) {
  if (v.NU == 0) return;
  hipLaunchKernelGGL((k_s_diag<T>), dim3(v.NU), dim3(256), 0, st, v, U, bU, Z, yV, Sd, bS, ctl, gate);
}

template <typename T>
void launch_s_tiles(const TileView& tv, int nparts, const int* tile_order, const int* part_tile, const int* part_rect,
                    const int* part_ptr, const int* part_blk, const int* sb_i, const int* sb_j, const int* sb_ptr,
                    const int* sb_pairs, const T* Sd, const T* bS, const T* Z, T* tiles, DevCtl* ctl, int gate,
                    hipStream_t st) {
  if (tv.TB == 64)
    hipLaunchKernelGGL((k_s_tiles<T, 64>), dim3(nparts), dim3(256), 0, st, tv, tile_order, part_tile, part_rect,
                       part_ptr, part_blk, sb_i, sb_j, sb_ptr, sb_pairs, Sd, bS, Z, tiles, ctl, gate);
  else if (tv.TB == 32)
    hipLaunchKernelGGL((k_s_tiles<T, 32>), dim3(nparts), dim3(256), 0, st, tv, tile_order, part_tile, part_rect,
                       part_ptr, part_blk, sb_i, sb_j, sb_ptr, sb_pairs, Sd, bS, Z, tiles, ctl, gate);
}

template <typename T>
void launch_backsub(const BAView& v, const T* W, const T* Vinv, const T* bV, const double* xU, double* xV,
                    const Scratch& s, DevCtl* ctl, int gate, hipStream_t st, const double* rhs_v, int add,
                    const double* bden, double* accU, int nacc) {
  if (v.NV == 0) {
    if (accU && nacc > 0) launch_add(accU, xU, nacc, ctl, gate, st);
    return;
  }
  hipLaunchKernelGGL((k_backsub<T>), dim3(nblocks((long long)v.NV * kLinLanes, 256)), dim3(256), 0, st, v, W, Vinv,
                     bV, xU, xV, s.part_xv, ctl, gate, rhs_v, add, bden, accU, accU ? nacc : 0);  // s.nblk_xv
}

void launch_ir_residual(const BAView& v, const double* xU, const double* xV, double* rU, double* rV, double* bU64,
                        double* bV64, DevCtl* ctl, int gate, hipStream_t st) {
  const int nbp = v.P > 0 ? (int)nblocks(v.P, 256) : 0;
  if (nbp + v.NU > 0)
    hipLaunchKernelGGL(k_ir_res, dim3(nbp + v.NU), dim3(256), 0, st, v, xU, xV, rU, rV, bU64, bV64, nbp, ctl, gate);
}

template <typename T>
void launch_ir_reduce(const BAView& v, const T* W, const T* Vinv, const double* rU, const double* rV, double* rS,
                      double* const unset[4], int nunset, DevCtl* ctl, int gate, hipStream_t st) {
  if (v.NU == 0) return;
  hipLaunchKernelGGL((k_ir_reduce<T>), dim3(v.NU), dim3(256), 0, st, v, W, Vinv, rU, rV, rS, unset[0], unset[1],
                     unset[2], unset[3], nunset, ctl, gate);
}

void launch_adj_eps(const double* lu, long long nu, const double* lv, long long nv, double* eps, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_eps, dim3(1), dim3(256), 0, st, lu, nu, lv, nv, eps);
}

void launch_adj_offset(const double* lu, long long nu, const double* lv, long long nv, const double* eps, double sgn,
                       double* tu, double* tv, hipStream_t st) {
  const long long n = nu > nv ? nu : nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_offset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lu, nu, lv, nv, eps, sgn, tu,
                     tv);
}

void launch_adj_resid(const double* gu, const double* bmu, double* bpu, long long nu, const double* gv,
                      const double* bmv, double* bpv, long long nv, const double* eps, hipStream_t st) {
  const long long n = nu > nv ? nu : nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_resid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gu, bmu, bpu, nu, gv, bmv, bpv,
                     nv, eps);
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

void launch_ir_decide(DevCtl* ctl, int gate, double tol, hipStream_t st) {
  hipLaunchKernelGGL(k_ir_decide, dim3(1), dim3(64), 0, st, ctl, gate, tol);
}

// ---------------------------------------------------------------------------
// off-diagonal U blocks (U-U edges of a generic section, added to S after
// the assembly: S = U - W V^-1 W^T, MatrixWWtMultiplier.hpp:124-153 copies the
// U blocks, then subtracts the products).  Block k is the upper block
// U_{i,j}, i < j, 6x6 row-major; S is stored lower, so U_{i,j}[b][a] goes to
// entry (6 j + a, 6 i + b).  Distinct blocks touch distinct entries.
// ---------------------------------------------------------------------------
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_add_uoff(TileView tv, int nb, const int* __restrict__ ui,
                                                  const int* __restrict__ uj, const T* __restrict__ Uoff,
                                                  T* __restrict__ tiles, DevCtl* ctl, int gate) {
  if (!gate_open(ctl, gate)) return;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= 36LL * nb) return;
  const int k = (int)(e / 36), a = (int)(e % 36) / 6, b = (int)(e % 6);
  const int R = 6 * uj[k] + a, Cc = 6 * ui[k] + b;
  const long long t = tv.slot(R / TB, Cc / TB);
  tiles[t * TB * TB + (long long)(Cc % TB) * TB + (R % TB)] += Uoff[36LL * k + 6 * b + a];
}

template <typename T>
void launch_add_uoff(const TileView& tv, int nb, const int* ui, const int* uj, const T* Uoff, T* tiles, DevCtl* ctl,
                     int gate, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL((k_add_uoff<T, 64>), dim3(nblocks(36LL * nb, 256)), dim3(256), 0, st, tv, nb, ui, uj, Uoff,
                     tiles, ctl, gate);
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
  template void launch_vinv_y<T>(const BAView&, const T*, const T*, const T*, T*, T*, double*, DevCtl*, int, \
                                 hipStream_t);                                                               \
  template void launch_s_diag<T>(const BAView&, const T*, const T*, const T*, const double*, T*, T*, DevCtl*, \
                                 int, hipStream_t);                                                          \
  template void launch_s_tiles<T>(const TileView&, int, const int*, const int*, const int*, const int*,      \
                                  const int*, const int*, const int*, const int*, const int*, const T*,      \
                                  const T*, const T*, T*, DevCtl*, int, hipStream_t);                        \
  template void launch_backsub<T>(const BAView&, const T*, const T*, const T*, const double*, double*,       \
                                  const Scratch&, DevCtl*, int, hipStream_t, const double*, int,             \
                                  const double*, double*, int);                                              \
  template void launch_ir_reduce<T>(const BAView&, const T*, const T*, const double*, const double*, double*, \
                                    double* const*, int, DevCtl*, int, hipStream_t);                         \
  template void launch_xu_stats<T>(int, const double*, const T*, const Scratch&, DevCtl*, int, hipStream_t); \
  template void launch_zero<T>(T*, long long, hipStream_t);                                                 \
  template void launch_add_uoff<T>(const TileView&, int, const int*, const int*, const T*, T*, DevCtl*, int,  \
                                   hipStream_t);

NSN_INST(double)
NSN_INST(float)
template void launch_zero<int>(int*, long long, hipStream_t);

}  // namespace nsn
