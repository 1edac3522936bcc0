// Device helpers of the BA kernels shared by the single-problem kernels
// (ba_kernels.hip) and the batched-window solver (batch_kernels.hip): the
// reprojection edge, its Jacobians, the BA oplus, reductions and the LM
// control steps.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nsn {

// Workgroups are dealt to the device's 8 XCDs in turn (workgroup b runs on
// XCD b % 8), each XCD with its own L2.  xcd_order<C>(b, n) maps workgroup b
// of n to a work item so that every XCD takes runs of C consecutive items:
// neighbouring cameras or S tiles, which read the same points' blocks (a
// point's W / Z blocks are adjacent in memory), then share one L2 instead of
// eight, while the runs still rotate over the XCDs (8 C items at a time), so
// uneven items spread over all of them.  A bijection on [0, n) (the last
// n mod 8C items keep their order); every item is computed as before.
template <int C>
__device__ __forceinline__ int xcd_order(int b, int n) {
  constexpr int kXcd = 8, S = kXcd * C;
  if (b >= (n / S) * S) return b;
  const int r = b % S;
  return b - r + (r % kXcd) * C + r / kXcd;
}

// one contiguous run per XCD (items 0 .. n/8 on the first, and so on), for
// launches whose workgroups are all resident at once and whose items cost
// about the same (one workgroup per camera)
__device__ __forceinline__ int xcd_runs(int b, int n) {
  constexpr int kXcd = 8;
  const int per = (n + kXcd - 1) / kXcd, rem = n % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * per - (rem ? max(0, x - rem) : 0) + k;
}

__device__ __forceinline__ bool gate_open(const DevCtl* c, int gate) {
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

// the workgroup's largest x and y (non-negative) folded into *a and *b with
// one atomic each, from thread 0 (blockDim.x == 256; sh: 8 doubles of LDS).
// One atomic per wave on one address serialised the point pass: 6 250 of
// them per launch at 200 000 points.
__device__ __forceinline__ void block_max2_atomic(double x, double y, double* sh, unsigned long long* a,
                                                  unsigned long long* b) {
  x = wave_max(x);
  y = wave_max(y);
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = x;
    sh[4 + (threadIdx.x >> 6)] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomic_max_nonneg(a, fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3])));
    atomic_max_nonneg(b, fmax(fmax(sh[4], sh[5]), fmax(sh[6], sh[7])));
  }
}

// packed upper index of a 6x6 symmetric block, i <= j
__host__ __device__ constexpr int up6(int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); }

// ---------------------------------------------------------------------------
// the reprojection edge
// ---------------------------------------------------------------------------
// camera record in registers: R (row-major), t, then (f, k1, k2) of the
// per-camera model when the problem has one (zeros otherwise)
constexpr int kCamRec = 15;

__device__ __forceinline__ void load_cam(const BAView& v, int c, double* R) {
#pragma unroll
  for (int i = 0; i < 12; i++) R[i] = v.cams[12 * c + i];
  if (v.intr) {
#pragma unroll
    for (int i = 0; i < 3; i++) R[12 + i] = v.intr[3 * c + i];
  } else {
    R[12] = R[13] = R[14] = 0.0;
  }
}

// Xc = R X + t, 1 / Xc_z, the residual, the loss terms, and for the BAL
// model D = d(pixel)/d(Xc) (2 x 3; the pinhole's D is formed where it is used)
struct Proj {
  double Xc0, Xc1, Xc2, iz, r0, r1, cost, om;
  double d00, d01, d02, d10, d11, d12;
};

// Two camera models:
//  * shared pinhole (v.intr == null): pixel = f Xc_xy / Xc_z + (cx, cy);
//  * the BAL model (v.intr: f, k1, k2 per camera, the camera looks down -z):
//    p = -Xc_xy / Xc_z, pixel = f (1 + k1 |p|^2 + k2 |p|^4) p.
__device__ __forceinline__ Proj project(const BAView& v, const double* __restrict__ R, double X0,
                                        double X1, double X2, double o0, double o1, double we) {
  Proj p;
  p.Xc0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + R[9];
  p.Xc1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + R[10];
  p.Xc2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + R[11];
  p.iz = 1.0 / p.Xc2;
  const double iz = p.iz;
  if (!v.intr) {
    p.r0 = v.f * p.Xc0 * p.iz + v.cx - o0;
    p.r1 = v.f * p.Xc1 * p.iz + v.cy - o1;
  } else {
    const double f = R[12], k1 = R[13], k2 = R[14];
    const double px = -p.Xc0 * iz, py = -p.Xc1 * iz;
    const double s2 = px * px + py * py;
    const double dist = 1.0 + s2 * (k1 + k2 * s2);
    p.r0 = f * dist * px - o0;
    p.r1 = f * dist * py - o1;
    // D = f (dist I + g p p^T) dp/dXc, g = d(dist)/d(s2) * 2, dp/dXc = -iz [I | p]
    const double g = 2.0 * k1 + 4.0 * k2 * s2;
    const double m00 = f * (dist + g * px * px), m01 = f * g * px * py, m11 = f * (dist + g * py * py);
    p.d00 = -iz * m00;
    p.d01 = -iz * m01;
    p.d02 = -iz * (m00 * px + m01 * py);
    p.d10 = -iz * m01;
    p.d11 = -iz * m11;
    p.d12 = -iz * (m01 * px + m11 * py);
  }
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

// d(r)/d(omega, upsilon) for the left SE(3) perturbation: D [-[Xc]x | I].
// The pinhole's D = [[a, 0, b0], [0, a, b1]] (a = f / z, b = -a Xc_xy / z).
template <typename T>
__device__ __forceinline__ void jac_cam(const BAView& v, const Proj& p, T* j0, T* j1) {
  if (!v.intr) {
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
    return;
  }
  j0[0] = (T)(p.d02 * p.Xc1 - p.d01 * p.Xc2);
  j0[1] = (T)(p.d00 * p.Xc2 - p.d02 * p.Xc0);
  j0[2] = (T)(p.d01 * p.Xc0 - p.d00 * p.Xc1);
  j0[3] = (T)p.d00;
  j0[4] = (T)p.d01;
  j0[5] = (T)p.d02;
  j1[0] = (T)(p.d12 * p.Xc1 - p.d11 * p.Xc2);
  j1[1] = (T)(p.d10 * p.Xc2 - p.d12 * p.Xc0);
  j1[2] = (T)(p.d11 * p.Xc0 - p.d10 * p.Xc1);
  j1[3] = (T)p.d10;
  j1[4] = (T)p.d11;
  j1[5] = (T)p.d12;
}

// d(r)/d(X) = D R
template <typename T>
__device__ __forceinline__ void jac_pt(const BAView& v, const Proj& p, const double* __restrict__ R,
                                       T* g0, T* g1) {
  if (!v.intr) {
    const double a = v.f * p.iz;
    const double b0 = -a * p.Xc0 * p.iz;
    const double b1 = -a * p.Xc1 * p.iz;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      g0[k] = (T)(a * R[k] + b0 * R[6 + k]);
      g1[k] = (T)(a * R[3 + k] + b1 * R[6 + k]);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    g0[k] = (T)(p.d00 * R[k] + p.d01 * R[3 + k] + p.d02 * R[6 + k]);
    g1[k] = (T)(p.d10 * R[k] + p.d11 * R[3 + k] + p.d12 * R[6 + k]);
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

// oplus of the BA sample (BASchurPerfomances.cpp:163-170) for variable
// camera idx: T <- exp(delta) T (R <- Exp(w) R, t <- Exp(w) t + V(w) u)
__device__ __forceinline__ void oplus_cam(const BAView& v, int idx, const double* __restrict__ xU) {
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
}

// points are additive
// principal log of a rotation (R row-major): theta = atan2(|vee(R - R^T)| / 2,
// (tr R - 1) / 2), omega = theta / (2 sin theta) vee(R - R^T); near theta = pi
// the axis comes from a a^T = ((R + R^T) / 2 - cos theta I) / (1 - cos theta)
__device__ inline void so3_log(const double* R, double* w) {
  const double v0 = R[7] - R[5], v1 = R[2] - R[6], v2 = R[3] - R[1];
  const double c = 0.5 * (R[0] + R[4] + R[8] - 1.0);
  const double s = 0.5 * sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  const double th = atan2(s, c);
  if (th < 1e-4) {
    const double k = 0.5 + th * th / 12.0;
    w[0] = k * v0; w[1] = k * v1; w[2] = k * v2;
  } else if (3.14159265358979323846 - th > 1e-4) {
    const double k = th / (2.0 * s);
    w[0] = k * v0; w[1] = k * v1; w[2] = k * v2;
  } else {
    double B[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) B[i * 3 + k] = (0.5 * (R[i * 3 + k] + R[k * 3 + i]) - (i == k ? c : 0.0)) / (1.0 - c);
    // the column j of the largest diagonal entry (the first on ties), taken
    // with selects: a runtime index into B would put B in scratch memory
    const int j1 = B[4] > B[0] ? 1 : 0;
    const int j = B[8] > (j1 ? B[4] : B[0]) ? 2 : j1;
    const double b0 = j == 0 ? B[0] : (j == 1 ? B[1] : B[2]);
    const double b1 = j == 0 ? B[3] : (j == 1 ? B[4] : B[5]);
    const double b2 = j == 0 ? B[6] : (j == 1 ? B[7] : B[8]);
    const double d = sqrt(j == 0 ? B[0] : (j == 1 ? B[4] : B[8]));
    const double a0 = b0 / d, a1 = b1 / d, a2 = b2 / d;
    const double sg = a0 * v0 + a1 * v1 + a2 * v2 < 0.0 ? -1.0 : 1.0;
    w[0] = sg * th * a0; w[1] = sg * th * a1; w[2] = sg * th * a2;
  }
}

// inverse left Jacobian of SO(3): I - K / 2 + (1 / th^2 - (1 + cos th) / (2 th sin th)) K^2, K = [w]x
__device__ inline void so3_jl_inv(const double* w, double* J) {
  const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double th = sqrt(th2);
  const double g = th < 1e-4 ? 1.0 / 12.0 + th2 / 720.0 : 1.0 / th2 - (1.0 + cos(th)) / (2.0 * th * sin(th));
  const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
      J[i * 3 + j] = (i == j ? 1.0 : 0.0) - 0.5 * K[i * 3 + j] + g * k2;
    }
}

// unary prior k on user camera c (EdgeUnary, nelson/EdgeUnary.h:76-93):
// r = w .* [log(R) - omega0; t - t0]; with JAC the weighted Jacobian (6x6
// row-major) w.r.t. the left update exp(d) T: [[Jl^-1(log R), 0], [-[t]x, I]].
// Returns |r|^2.
template <bool JAC>
__device__ inline double cam_prior_eval(const BAView& v, int c, int k, double* r, double* J) {
  const double* cm = v.cams + 12 * (size_t)c;
  const double* p0 = v.cam_prior_p + 6 * (size_t)k;
  const double* w = v.cam_prior_w + 6 * (size_t)k;
  double R[9], om[3];
  for (int i = 0; i < 9; i++) R[i] = cm[i];
  so3_log(R, om);
  double cost = 0.0;
  for (int i = 0; i < 3; i++) {
    r[i] = w[i] * (om[i] - p0[i]);
    r[3 + i] = w[3 + i] * (cm[9 + i] - p0[3 + i]);
  }
  for (int i = 0; i < 6; i++) cost += r[i] * r[i];
  if (JAC) {
    double Jl[9];
    so3_jl_inv(om, Jl);
    const double t0 = cm[9], t1 = cm[10], t2 = cm[11];
    const double Tx[9] = {0, t2, -t1, -t2, 0, t0, t1, -t0, 0};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        J[i * 6 + j] = w[i] * Jl[i * 3 + j];
        J[i * 6 + 3 + j] = 0.0;
        J[(3 + i) * 6 + j] = w[3 + i] * Tx[i * 3 + j];
        J[(3 + i) * 6 + 3 + j] = (i == j) ? w[3 + i] : 0.0;
      }
  }
  return cost;
}

__device__ __forceinline__ void oplus_pt(const BAView& v, int q, const double* __restrict__ xV) {
  const int p = v.var_pt_user[q];
#pragma unroll
  for (int i = 0; i < 3; i++) v.pts[3 * (size_t)p + i] += xV[3 * (size_t)q + i];
}

__device__ inline void log_add(DevCtl* c, DevLog* log, int it, int subit, double chi2, double mu) {
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
__device__ inline void lm_after_lin(DevCtl* c, DevLog* log) {
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
__device__ inline void lm_control(DevCtl* c, DevLog* log) {
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


// one point of the point-major linearisation pass: V_p, bV_p, the W blocks
// of its edges and the chi2 of every edge of p, with max |bV| and max diag V
// of the point (nelson/BaseSection.cpp:88-169, updateH22Block / updateH12Block)
template <typename T, bool HESS>
__device__ __forceinline__ void lin_point(const BAView& v, int p, T* __restrict__ V, T* __restrict__ bV,
                                          T* __restrict__ W, double& chi2, double& mb, double& md) {
  const double X0 = v.pts[3 * p], X1 = v.pts[3 * p + 1], X2 = v.pts[3 * p + 2];
  const int pv = v.pt_var[p];
  T vv[6] = {0, 0, 0, 0, 0, 0};
  T bv[3] = {0, 0, 0};
  T w[18];
  int cur = -1;
  const int e0 = v.pt_ptr[p], e1 = v.pt_ptr[p + 1];
  for (int e = e0; e < e1; e++) {
    const int c = v.pe_cam[e];
    double R[kCamRec];
    load_cam(v, c, R);
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
  // the point's unary prior (after its edges): r = w .* (X - X0), J = diag(w)
  if (v.pt_prior) {
    const int k = v.pt_prior[p];
    if (k >= 0) {
      const double* x0 = v.pt_prior_p + 3 * (size_t)k;
      const double* wp = v.pt_prior_w + 3 * (size_t)k;
      const double Xp[3] = {X0, X1, X2};
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double e = Xp[i] - x0[i], w2 = wp[i] * wp[i];
        chi2 += w2 * e * e;
        if (HESS && pv >= 0) {
          vv[i == 0 ? 0 : (i == 1 ? 3 : 5)] += (T)w2;
          bv[i] += (T)(w2 * e);
        }
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

// The point-major linearisation pass with L lanes per point (a group of L
// consecutive lanes of one wave; L divides 64): lane l takes the edges
// e0 + l, e0 + l + L, ... of point p (p >= v.P: a padding group with no
// edges, which still joins the shuffles).  Each lane sums its edges' V and bV
// terms; the group closes them with xor shuffles (every lane ends with the
// same sum), then lane 0 adds the point's unary prior and writes V_p, bV_p.
// The lane of a W block's first edge forms the block: its own edge, then
// (DUP: the problem has them) the camera's repeated observations of the
// point, which are adjacent; without DUP the kernel needs no registers for
// them (124 instead of 204 VGPRs in FP64).  The chi2 of each lane's edges
// goes into that lane's chi2.  Same terms as lin_point, with the sum over a
// point's edges grouped by lane.
template <typename T, bool HESS, int L, bool DUP>
__device__ __forceinline__ void lin_point_lanes(const BAView& v, int p, int l, T* __restrict__ V,
                                                T* __restrict__ bV, T* __restrict__ W, double& chi2, double& mb,
                                                double& md) {
  const bool valid = p < v.P;
  double X0 = 0.0, X1 = 0.0, X2 = 0.0;
  int pv = -1, e0 = 0, e1 = 0;
  if (valid) {
    X0 = v.pts[3 * p];
    X1 = v.pts[3 * p + 1];
    X2 = v.pts[3 * p + 2];
    pv = v.pt_var[p];
    e0 = v.pt_ptr[p];
    e1 = v.pt_ptr[p + 1];
  }
  T vv[6] = {0, 0, 0, 0, 0, 0};
  T bv[3] = {0, 0, 0};
  for (int e = e0 + l; e < e1; e += L) {
    const int c = v.pe_cam[e];
    double R[kCamRec];
    load_cam(v, c, R);
    const Proj pr = project(v, R, X0, X1, X2, v.pe_obs[2 * e], v.pe_obs[2 * e + 1], v.pe_w[e]);
    chi2 += pr.cost;
    if (HESS && pv >= 0) {
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
      const int blk = v.pe_wblk[e];
      if (blk >= 0 && (e == e0 || v.pe_wblk[e - 1] != blk)) {
        T w[18];
        {
          T j0[6], j1[6];
          jac_cam<T>(v, pr, j0, j1);
#pragma unroll
          for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) w[i * 3 + j] = (T)0 + om * (j0[i] * g0[j] + j1[i] * g1[j]);
        }
        if constexpr (DUP) {
          for (int f = e + 1; f < e1 && v.pe_wblk[f] == blk; f++) {
            const Proj pf = project(v, R, X0, X1, X2, v.pe_obs[2 * f], v.pe_obs[2 * f + 1], v.pe_w[f]);
            T h0[3], h1[3], j0[6], j1[6];
            jac_pt<T>(v, pf, R, h0, h1);
            jac_cam<T>(v, pf, j0, j1);
            const T of = (T)pf.om;
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
              for (int j = 0; j < 3; j++) w[i * 3 + j] += of * (j0[i] * h0[j] + j1[i] * h1[j]);
          }
        }
#pragma unroll
        for (int i = 0; i < 18; i++) W[18 * (size_t)blk + i] = w[i];
      }
    }
  }
  if (HESS) {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) {
#pragma unroll
      for (int i = 0; i < 6; i++) vv[i] += (T)__shfl_xor((double)vv[i], o);
#pragma unroll
      for (int i = 0; i < 3; i++) bv[i] += (T)__shfl_xor((double)bv[i], o);
    }
  }
  if (l != 0 || !valid) return;
  // the point's unary prior (after its edges): r = w .* (X - X0), J = diag(w)
  if (v.pt_prior) {
    const int k = v.pt_prior[p];
    if (k >= 0) {
      const double* x0 = v.pt_prior_p + 3 * (size_t)k;
      const double* wp = v.pt_prior_w + 3 * (size_t)k;
      const double Xp[3] = {X0, X1, X2};
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double e = Xp[i] - x0[i], w2 = wp[i] * wp[i];
        chi2 += w2 * e * e;
        if (HESS && pv >= 0) {
          vv[i == 0 ? 0 : (i == 1 ? 3 : 5)] += (T)w2;
          bv[i] += (T)(w2 * e);
        }
      }
    }
  }
  if (HESS && pv >= 0) {
#pragma unroll
    for (int i = 0; i < 6; i++) V[6 * (size_t)pv + i] = vv[i];
#pragma unroll
    for (int i = 0; i < 3; i++) bV[3 * (size_t)pv + i] = bv[i];
    mb = fmax(fabs((double)bv[0]), fmax(fabs((double)bv[1]), fabs((double)bv[2])));
    md = fmax(fabs((double)vv[0]), fmax(fabs((double)vv[3]), fabs((double)vv[5])));
  }
}

// The damped 3x3 block of point q (V + rel diag V + abs I) through its
// Cholesky factor L in FP64 (MatrixDiagInv.hpp:25-140): V^-1 (packed upper,
// zero when the block is not positive definite), x = V^-1 (-bV), and the
// right factor M of the W products: M = V^-1, or with ZF M = L^-T (zero when
// the block failed), so that Y = W M gives W V^-1 W'^T = Y Y'^T; with ZF
// also y = L^-1 (-bV) (zero when the block failed), so that t = W V^-1 (-bV)
// = Z y.
template <typename T, bool ZF>
struct VFactor {
  double vi[6];
  double x0, x1, x2;
  double y0, y1, y2;
  T M[9];
  bool ok;
};

template <typename T, bool ZF>
__device__ __forceinline__ VFactor<T, ZF> v_factor(const T* __restrict__ V, const T* __restrict__ bV, int q,
                                                   double rel, double ab) {
  VFactor<T, ZF> f;
  double a00 = V[6 * (size_t)q + 0], a01 = V[6 * (size_t)q + 1], a02 = V[6 * (size_t)q + 2];
  double a11 = V[6 * (size_t)q + 3], a12 = V[6 * (size_t)q + 4], a22 = V[6 * (size_t)q + 5];
  if (rel != 0.0 || ab != 0.0) {
    a00 = a00 * (1.0 + rel) + ab;
    a11 = a11 * (1.0 + rel) + ab;
    a22 = a22 * (1.0 + rel) + ab;
  }
  double* vi = f.vi;
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
  if (!ok)
    for (int i = 0; i < 6; i++) vi[i] = 0.0;
  f.ok = ok;
  const double b0 = bV[3 * (size_t)q], b1 = bV[3 * (size_t)q + 1], b2 = bV[3 * (size_t)q + 2];
  // x = V^-1 (-bV)
  f.x0 = -(vi[0] * b0 + vi[1] * b1 + vi[2] * b2);
  f.x1 = -(vi[1] * b0 + vi[3] * b1 + vi[4] * b2);
  f.x2 = -(vi[2] * b0 + vi[4] * b1 + vi[5] * b2);
  T* M = f.M;
  f.y0 = f.y1 = f.y2 = 0.0;
  if constexpr (ZF) {
    const double z = ok ? 1.0 : 0.0;
    if (ok) {
      f.y0 = -(m00 * b0);
      f.y1 = -(m10 * b0 + m11 * b1);
      f.y2 = -(m20 * b0 + m21 * b1 + m22 * b2);
    }
    M[0] = (T)(z * m00); M[1] = (T)(z * m10); M[2] = (T)(z * m20);
    M[3] = (T)0;         M[4] = (T)(z * m11); M[5] = (T)(z * m21);
    M[6] = (T)0;         M[7] = (T)0;         M[8] = (T)(z * m22);
  } else {
    M[0] = (T)vi[0]; M[1] = (T)vi[1]; M[2] = (T)vi[2];
    M[3] = (T)vi[1]; M[4] = (T)vi[3]; M[5] = (T)vi[4];
    M[6] = (T)vi[2]; M[7] = (T)vi[4]; M[8] = (T)vi[5];
  }
  return f;
}

// Y_k = W_k M (6x3) for the W block k
template <typename T, bool ZF>
__device__ __forceinline__ void z_block(const VFactor<T, ZF>& f, const T* __restrict__ W, T* __restrict__ Y, int k) {
  const T* w = W + 18 * (size_t)k;
  const T* M = f.M;
  T y[18];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const T w0 = w[3 * i], w1 = w[3 * i + 1], w2 = w[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; j++) y[3 * i + j] = w0 * M[j] + w1 * M[3 + j] + w2 * M[6 + j];
  }
#pragma unroll
  for (int i = 0; i < 18; i++) Y[18 * (size_t)k + i] = y[i];
}

// Y_k = W_k M (6x3) and t_k = W_k x (6) for the W block k
template <typename T, bool ZF>
__device__ __forceinline__ void y_block(const VFactor<T, ZF>& f, const T* __restrict__ W, T* __restrict__ Y,
                                        T* __restrict__ tW, int k) {
  const T* w = W + 18 * (size_t)k;
  const T* M = f.M;
  T y[18];
  T t[6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const T w0 = w[3 * i], w1 = w[3 * i + 1], w2 = w[3 * i + 2];
#pragma unroll
    for (int j = 0; j < 3; j++) y[3 * i + j] = w0 * M[j] + w1 * M[3 + j] + w2 * M[6 + j];
    t[i] = (T)((double)w0 * f.x0 + (double)w1 * f.x1 + (double)w2 * f.x2);
  }
#pragma unroll
  for (int i = 0; i < 18; i++) Y[18 * (size_t)k + i] = y[i];
#pragma unroll
  for (int i = 0; i < 6; i++) tW[6 * (size_t)k + i] = t[i];
}

// V_q^-1 into Vinv and Y, t for the W blocks of point q; false if the damped
// block is not positive definite (V^-1 is then zero)
template <typename T, bool ZF = false>
__device__ __forceinline__ bool vinv_y_point(const BAView& v, int q, const T* __restrict__ V,
                                             const T* __restrict__ bV, const T* __restrict__ W,
                                             T* __restrict__ Vinv, T* __restrict__ Y, T* __restrict__ tW, double rel,
                                             double ab) {
  const VFactor<T, ZF> f = v_factor<T, ZF>(V, bV, q, rel, ab);
#pragma unroll
  for (int i = 0; i < 6; i++) Vinv[6 * (size_t)q + i] = (T)f.vi[i];
  const int k1 = v.wptr[q + 1];
  for (int k = v.wptr[q]; k < k1; k++) y_block<T, ZF>(f, W, Y, tW, k);
  return f.ok;
}

// back-substitution of point q: xV_q = V_q^-1 (rhs_q - sum_k W_k^T xU_c(k)),
// rhs = -bV or rhs_v, added to xV with `add`; inc = |xV_q|^2 and
// den = xV_q^T (mu xV_q - b_q) with b = bden (FP64) or bV
template <typename T>
__device__ __forceinline__ void backsub_finish(const BAView& v, int q, const T* __restrict__ Vinv,
                                               const T* __restrict__ bV, double* __restrict__ xV, double mu,
                                               const double* __restrict__ rhs_v, int add,
                                               const double* __restrict__ bden, double s0, double s1, double s2,
                                               double& inc, double& den);

// (W^T xU)_q over the W blocks wptr[q] + l, wptr[q] + l + L, ... of point q
template <typename T, int L>
__device__ __forceinline__ void wt_xu(const BAView& v, int q, int l, const T* __restrict__ W,
                                      const double* __restrict__ xU, double& s0, double& s1, double& s2) {
  const int k1 = v.wptr[q + 1];
  for (int k = v.wptr[q] + l; k < k1; k += L) {
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
}

template <typename T>
__device__ __forceinline__ void backsub_point(const BAView& v, int q, const T* __restrict__ W,
                                              const T* __restrict__ Vinv, const T* __restrict__ bV,
                                              const double* __restrict__ xU, double* __restrict__ xV, double mu,
                                              const double* __restrict__ rhs_v, int add,
                                              const double* __restrict__ bden, double& inc, double& den) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  wt_xu<T, 1>(v, q, 0, W, xU, s0, s1, s2);
  backsub_finish<T>(v, q, Vinv, bV, xV, mu, rhs_v, add, bden, s0, s1, s2, inc, den);
}

// the rest of point q's back-substitution from s = (W^T xU)_q
template <typename T>
__device__ __forceinline__ void backsub_finish(const BAView& v, int q, const T* __restrict__ Vinv,
                                               const T* __restrict__ bV, double* __restrict__ xV, double mu,
                                               const double* __restrict__ rhs_v, int add,
                                               const double* __restrict__ bden, double s0, double s1, double s2,
                                               double& inc, double& den) {
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
  inc = x0 * x0 + x1 * x1 + x2 * x2;
  den = x0 * (mu * x0 - bv0) + x1 * (mu * x1 - bv1) + x2 * (mu * x2 - bv2);
}

// ---- the FP64 residual of the refinement, per point (single problems and
// batched windows) and per camera (single problems) ----
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
__device__ __forceinline__ void ir_res_point(const BAView& v, const double* __restrict__ xU,
                                             const double* __restrict__ xV, double* __restrict__ rV,
                                             double* __restrict__ bV64, const DevCtl* ctl, int p, double& mr,
                                             double& mb) {
  const int pv = v.pt_var[p];
  if (pv < 0) return;
  const double X0 = v.pts[3 * p], X1 = v.pts[3 * p + 1], X2 = v.pts[3 * p + 2];
  const double d0 = xV[3 * (size_t)pv], d1 = xV[3 * (size_t)pv + 1], d2 = xV[3 * (size_t)pv + 2];
  double b[3] = {0, 0, 0}, hq[3] = {0, 0, 0}, dg[3] = {0, 0, 0};
  const int e1 = v.pt_ptr[p + 1];
  for (int e = v.pt_ptr[p]; e < e1; e++) {
    const int c = v.pe_cam[e];
    double R[kCamRec];
    load_cam(v, c, R);
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
  const double d[3] = {d0, d1, d2};
  if (v.pt_prior) {  // the point's unary prior: J = diag(w)
    const int k = v.pt_prior[p];
    if (k >= 0) {
      const double Xp[3] = {X0, X1, X2};
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double w2 = v.pt_prior_w[3 * (size_t)k + i] * v.pt_prior_w[3 * (size_t)k + i];
        b[i] += w2 * (Xp[i] - v.pt_prior_p[3 * (size_t)k + i]);
        hq[i] += w2 * d[i];
        dg[i] += w2;
      }
    }
  }
  const double rel = ctl->rel, ab = ctl->abs_;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double r = -b[k] - hq[k] - (rel * dg[k] + ab) * d[k];
    rV[3 * (size_t)pv + k] = r;
    bV64[3 * (size_t)pv + k] = b[k];
    mr = fmax(mr, fabs(r));
    mb = fmax(mb, fabs(b[k]));
  }
}

// one workgroup per variable camera; deterministic reduction
// red: 4 x 18 doubles of LDS; pri: 18 (the camera's unary prior: J^T r,
// J^T J x, diag(J^T J)); the whole workgroup calls it for one camera
__device__ __forceinline__ void ir_res_cam(const BAView& v, const double* __restrict__ xU,
                                           const double* __restrict__ xV, double* __restrict__ rU,
                                           double* __restrict__ bU64, const DevCtl* ctl, int u, double (*red)[18],
                                           double* pri, double& mr, double& mb) {
  const int c = v.var_cam_user[u];
  double R[kCamRec];
  load_cam(v, c, R);
  double xc[6];
#pragma unroll
  for (int i = 0; i < 6; i++) xc[i] = xU[6 * (size_t)u + i];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 18; i++) pri[i] = 0.0;
    const int kp = v.cam_prior ? v.cam_prior[c] : -1;
    if (kp >= 0) {
      double r[6], J[36];
      cam_prior_eval<true>(v, c, kp, r, J);
#pragma unroll
      for (int m = 0; m < 6; m++) {
        double jx = 0.0;
#pragma unroll
        for (int n = 0; n < 6; n++) jx += J[m * 6 + n] * xc[n];
#pragma unroll
        for (int i = 0; i < 6; i++) {
          pri[i] += J[m * 6 + i] * r[m];
          pri[6 + i] += J[m * 6 + i] * jx;
          pri[12 + i] += J[m * 6 + i] * J[m * 6 + i];
        }
      }
    }
  }
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
    t[0] += pri[i];
    t[1] += pri[6 + i];
    t[2] += pri[12 + i];
    const double rel = ctl->rel, ab = ctl->abs_;
    const double r = -t[0] - t[1] - (rel * t[2] + ab) * xc[i];
    rU[6 * (size_t)u + i] = r;
    bU64[6 * (size_t)u + i] = t[0];
    mr = fabs(r);
    mb = fabs(t[0]);
  }
}


}  // namespace nsn
