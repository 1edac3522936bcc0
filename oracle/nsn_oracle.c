/*
 * nsn_oracle.c -- CPU FP64 restatement of nelson's GN/LM Schur-complement
 * step for bundle adjustment.  TEST INFRASTRUCTURE ONLY (the checker): only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The reference (nelson + Eigen3 + mat@847e63e) cannot be built here, so this
 * file restates its algorithm.  What pins each part independently of this
 * file (tests/test_oracle*.py):
 *   or_ldlt_factor/solve  : residual ||A x - b|| and L D L^T reconstruction
 *                           vs scipy.linalg.ldl / numpy.linalg.solve;
 *                           pivot order = argmax |diag| (hand-checked case).
 *   or_schur_step         : equals the dense solve of (H + diag(rel*H_ii+abs)) dx = -b
 *                           assembled by numpy from the same blocks (brute force).
 *   or_diag_inv, or_wvinv, or_wvinv_mult_sub, or_wwt, or_wtx_sub
 *                         : the reference's own property tests
 *                           (nelson/tests/testMatrix*.cpp) recomputed with
 *                           dense numpy products on the same sparsity patterns.
 *   or_rho                : closed form; equals 1 for a linear problem.
 *   or_lm_solve/or_gn_solve: linear toy problem of
 *                           nelson/tests/testDoubleSectionCommon.h converges to
 *                           chi2 < 1e-12; mu schedule known answers (x1/3 per
 *                           accepted linear step, x2,x4,x8 on forced rejects);
 *                           MaxSubIterationReached / SuccessAtStart / SolverFailure.
 *   or_ba_*               : residuals vs an independent numpy projection;
 *                           Jacobians vs central finite differences;
 *                           H, b vs dense J^T J / J^T r assembled by numpy.
 */
#include "nsn_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Hessian storage                                                          */
/* ------------------------------------------------------------------------ */
int or_hessian_alloc(or_hessian* h, int BU, int BV, int NU, int NV, int nW) {
  memset(h, 0, sizeof(*h));
  h->BU = BU; h->BV = BV; h->NU = NU; h->NV = NV; h->nW = nW;
  size_t nu = (size_t)BU * NU;
  h->U = (double*)calloc(nu * nu + 1, sizeof(double));
  h->V = (double*)calloc((size_t)NV * BV * BV + 1, sizeof(double));
  h->W = (double*)calloc((size_t)nW * BU * BV + 1, sizeof(double));
  h->bU = (double*)calloc(nu + 1, sizeof(double));
  h->bV = (double*)calloc((size_t)NV * BV + 1, sizeof(double));
  h->Wi = (int*)calloc((size_t)nW + 1, sizeof(int));
  h->Wj = (int*)calloc((size_t)nW + 1, sizeof(int));
  return h->U && h->V && h->W && h->bU && h->bV && h->Wi && h->Wj;
}

int or_hessian_alloc_voff(or_hessian* h, int nVo) {
  h->nVo = nVo;
  h->Vo = (double*)calloc((size_t)nVo * h->BV * h->BV + 1, sizeof(double));
  h->Voi = (int*)calloc((size_t)nVo + 1, sizeof(int));
  h->Voj = (int*)calloc((size_t)nVo + 1, sizeof(int));
  return h->Vo && h->Voi && h->Voj;
}

void or_hessian_free(or_hessian* h) {
  free(h->U); free(h->V); free(h->W); free(h->bU); free(h->bV); free(h->Wi); free(h->Wj);
  free(h->Vo); free(h->Voi); free(h->Voj);
  memset(h, 0, sizeof(*h));
}

/* nelson/DoubleSectionHessian.hpp:51-58 */
void or_hessian_clear(or_hessian* h) {
  size_t nu = (size_t)h->BU * h->NU;
  h->chi2 = 0;
  memset(h->U, 0, nu * nu * sizeof(double));
  memset(h->V, 0, (size_t)h->NV * h->BV * h->BV * sizeof(double));
  memset(h->W, 0, (size_t)h->nW * h->BU * h->BV * sizeof(double));
  memset(h->bU, 0, nu * sizeof(double));
  memset(h->bV, 0, (size_t)h->NV * h->BV * sizeof(double));
  if (h->nVo) memset(h->Vo, 0, (size_t)h->nVo * h->BV * h->BV * sizeof(double));
}

/* nelson/DoubleSectionHessian.hpp:60-63 */
double or_hessian_max_abs_b(const or_hessian* h) {
  double m = 0;
  size_t nu = (size_t)h->BU * h->NU, nv = (size_t)h->BV * h->NV;
  for (size_t i = 0; i < nu; i++) if (fabs(h->bU[i]) > m) m = fabs(h->bU[i]);
  double mv = 0;
  for (size_t i = 0; i < nv; i++) if (fabs(h->bV[i]) > mv) mv = fabs(h->bV[i]);
  return m > mv ? m : mv;
}

/* ------------------------------------------------------------------------ */
/* Dense LDL^T with symmetric diagonal pivoting: the algorithm of Eigen's   */
/* LDLT (ldlt_inplace<Lower>::unblocked), which nelson calls through         */
/* Eigen::LDLT<M, Upper> (nelson/SolverCholeskyDense.h:14,47;               */
/* nelson/SolverCholeskyDense.hpp:15-18,55).  a is column-major, lower.      */
/* ------------------------------------------------------------------------ */
#define AC(i, j) a[(size_t)(j) * (size_t)n + (size_t)(i)]

#ifdef _OPENMP
#include <omp.h>
static int host_threads(void) { return omp_get_max_threads(); }
int or_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
#else
static int host_threads(void) { return 1; }
int or_set_threads(int n) {
  (void)n;
  return 1;
}
#endif

int or_ldlt_factor(double* a, int n, int* tp) {
  if (n <= 0) return 1;
  if (n == 1) { tp[0] = 0; return 1; }
  double* temp = (double*)malloc((size_t)n * sizeof(double));
  int ret = 1, found_zero = 0;
  for (int k = 0; k < n; k++) {
    /* largest |diagonal| in the trailing corner (first occurrence) */
    int idx = k;
    double big = fabs(AC(k, k));
    for (int i = k + 1; i < n; i++) {
      double v = fabs(AC(i, i));
      if (v > big) { big = v; idx = i; }
    }
    tp[k] = idx;
    if (idx != k) {
      /* symmetric swap of k and idx restricted to the lower triangle */
      for (int j = 0; j < k; j++) { double t = AC(k, j); AC(k, j) = AC(idx, j); AC(idx, j) = t; }
      for (int i = idx + 1; i < n; i++) { double t = AC(i, k); AC(i, k) = AC(i, idx); AC(i, idx) = t; }
      { double t = AC(k, k); AC(k, k) = AC(idx, idx); AC(idx, idx) = t; }
      for (int i = k + 1; i < idx; i++) { double t = AC(i, k); AC(i, k) = AC(idx, i); AC(idx, i) = t; }
    }
    int rs = n - k - 1;
    if (k > 0) {
      for (int j = 0; j < k; j++) temp[j] = AC(j, j) * AC(k, j);
      double dot = 0;
      for (int j = 0; j < k; j++) dot += AC(k, j) * temp[j];
      AC(k, k) -= dot;
      if (rs > 0) {
        /* colk[i] -= sum_j colj[i] temp[j], each entry over j in ascending
         * order (Eigen's A21 -= A20 * temp).  Rows are split in blocks over
         * the host threads (OpenMP): every entry sees the same operations in
         * the same order whatever the thread count, so the factor is the
         * same bit for bit as the single-threaded one. */
        double* colk = &AC(0, k);
        const int nt = host_threads();
        if (nt == 1 || rs < 256) {
          for (int j = 0; j < k; j++) {
            const double tj = temp[j];
            const double* colj = &AC(0, j);
            for (int i = k + 1; i < n; i++) colk[i] -= colj[i] * tj;
          }
        } else {
          /* 64-row blocks: a block of colk stays in L1 over the j loop */
          const int bs = 64;
          const int nblk = (rs + bs - 1) / bs;
#pragma omp parallel for schedule(static) num_threads(nt)
          for (int ib = 0; ib < nblk; ib++) {
            const int i0 = k + 1 + bs * ib, i1 = i0 + bs < n ? i0 + bs : n;
            for (int j = 0; j < k; j++) {
              const double tj = temp[j];
              const double* colj = &AC(0, j);
              for (int i = i0; i < i1; i++) colk[i] -= colj[i] * tj;
            }
          }
        }
      }
    }
    double akk = AC(k, k);
    int valid = fabs(akk) > 0.0;
    if (k == 0 && !valid) {
      for (int j = 0; j < n; j++) {
        tp[j] = j;
        for (int i = j + 1; i < n; i++) if (AC(i, j) != 0.0) ret = 0;
      }
      free(temp);
      return ret;
    }
    if (rs > 0 && valid) {
      for (int i = k + 1; i < n; i++) AC(i, k) /= akk;
    } else if (rs > 0) {
      for (int i = k + 1; i < n; i++) if (AC(i, k) != 0.0) ret = 0;
    }
    if (found_zero && valid) ret = 0;
    else if (!valid) found_zero = 1;
  }
  free(temp);
  return ret;
}

/* LDLT::solve: P^T L^-T D^+ L^-1 P b, D^+ the pseudo-inverse with tolerance DBL_MIN */
void or_ldlt_solve(const double* a, int n, const int* tp, double* x) {
  for (int k = 0; k < n; k++) { int j = tp[k]; if (j != k) { double t = x[k]; x[k] = x[j]; x[j] = t; } }
  for (int j = 0; j < n; j++) {
    const double xj = x[j];
    for (int i = j + 1; i < n; i++) x[i] -= AC(i, j) * xj;
  }
  for (int i = 0; i < n; i++) {
    double d = AC(i, i);
    if (fabs(d) > DBL_MIN) x[i] /= d; else x[i] = 0.0;
  }
  for (int j = n - 1; j >= 0; j--) {
    double s = x[j];
    for (int i = j + 1; i < n; i++) s -= AC(i, j) * x[i];
    x[j] = s;
  }
  for (int k = n - 1; k >= 0; k--) { int j = tp[k]; if (j != k) { double t = x[k]; x[k] = x[j]; x[j] = t; } }
}
#undef AC

/* ------------------------------------------------------------------------ */
/* L4 kernels                                                               */
/* ------------------------------------------------------------------------ */

/* small dense inverse by LDL^T with pivoting (stands for Eigen's fixed-size
 * .inverse(), nelson/MatrixDiagInv.hpp:41).  m: B x B row-major symmetric
 * (upper read).  Returns 1 on success. */
static int small_sym_inverse(int B, const double* m, double* inv) {
  double a[64]; int tp[8];
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      /* column-major lower element (i,j), i>=j, = upper element (j,i) */
      a[j * B + i] = (i >= j) ? m[j * B + i] : 0.0;
    }
  int ok = or_ldlt_factor(a, B, tp);
  for (int c = 0; c < B; c++) {
    double x[8];
    for (int r = 0; r < B; r++) x[r] = (r == c) ? 1.0 : 0.0;
    or_ldlt_solve(a, B, tp, x);
    for (int r = 0; r < B; r++) inv[r * B + c] = x[r];
  }
  return ok;
}

/* nelson/MatrixDiagInv.hpp:25-140: Vinv_p = (sym_upper(V_p), diag*(1+rel)+abs)^-1 */
int or_diag_inv(int NV, int BV, const double* V, double rel, double abs_, double* Vinv) {
  int ok = 1;
  double m[64];
  for (int p = 0; p < NV; p++) {
    const double* v = V + (size_t)p * BV * BV;
    for (int i = 0; i < BV; i++)
      for (int j = 0; j < BV; j++) m[i * BV + j] = (j >= i) ? v[i * BV + j] : v[j * BV + i];
    if (rel != 0 || abs_ != 0)
      for (int i = 0; i < BV; i++) { m[i * BV + i] *= (1 + rel); m[i * BV + i] += abs_; }
    ok &= small_sym_inverse(BV, m, Vinv + (size_t)p * BV * BV);
  }
  return ok;
}

/* nelson/MatrixDiagInv.hpp:143-213 */
void or_diag_inv_mult(int NV, int BV, const double* Vinv, const double* v, double* res) {
  for (int p = 0; p < NV; p++) {
    const double* M = Vinv + (size_t)p * BV * BV;
    for (int i = 0; i < BV; i++) {
      double s = 0;
      for (int j = 0; j < BV; j++) s += M[i * BV + j] * v[p * BV + j];
      res[p * BV + i] = s;
    }
  }
}

/* nelson/MatrixWVinvMultiplier.hpp:23-106: Y_k = W_k * Vinv_{j(k)} */
void or_wvinv(int nW, int BU, int BV, const int* Wj, const double* W, const double* Vinv, double* Y) {
  for (int k = 0; k < nW; k++) {
    const double* w = W + (size_t)k * BU * BV;
    const double* vi = Vinv + (size_t)Wj[k] * BV * BV;
    double* y = Y + (size_t)k * BU * BV;
    for (int r = 0; r < BU; r++)
      for (int c = 0; c < BV; c++) {
        double s = 0;
        for (int m = 0; m < BV; m++) s += w[r * BV + m] * vi[m * BV + c];
        y[r * BV + c] = s;
      }
  }
}

/* nelson/MatrixWVinvMultiplier.hpp:110-193: res_i -= sum_k Y_k v_{j(k)} */
void or_wvinv_mult_sub(int nW, int BU, int BV, const int* Wi, const int* Wj, const double* Y,
                       const double* v, double* res) {
  for (int k = 0; k < nW; k++) {
    const double* y = Y + (size_t)k * BU * BV;
    for (int r = 0; r < BU; r++) {
      double s = 0;
      for (int c = 0; c < BV; c++) s += y[r * BV + c] * v[(size_t)Wj[k] * BV + c];
      res[(size_t)Wi[k] * BU + r] -= s;
    }
  }
}

/* nelson/MatrixWWtMultiplier.hpp:124-284: S = upper(U) - sum_{i<=j} A_ik B_jk^T.
 * Blocks of the same V column k pair up; W is in UID (row-major) order. */
void or_wwt(int NU, int BU, int BV, int nW, const int* Wi, const int* Wj,
            const double* A, const double* B, const double* U, double* S) {
  size_t nu = (size_t)NU * BU;
  for (size_t i = 0; i < nu; i++)
    for (size_t j = 0; j < nu; j++) S[i * nu + j] = (j >= i) ? U[i * nu + j] : 0.0;
  /* column lists */
  int NV = 0;
  for (int k = 0; k < nW; k++) if (Wj[k] + 1 > NV) NV = Wj[k] + 1;
  int* cnt = (int*)calloc((size_t)NV + 1, sizeof(int));
  for (int k = 0; k < nW; k++) cnt[Wj[k] + 1]++;
  for (int p = 0; p < NV; p++) cnt[p + 1] += cnt[p];
  int* fill = (int*)malloc(((size_t)NV + 1) * sizeof(int));
  memcpy(fill, cnt, ((size_t)NV + 1) * sizeof(int));
  int* lst = (int*)malloc(((size_t)nW + 1) * sizeof(int));
  for (int k = 0; k < nW; k++) lst[fill[Wj[k]]++] = k;  /* ascending row within each column */
  for (int p = 0; p < NV; p++) {
    for (int a = cnt[p]; a < cnt[p + 1]; a++) {
      int ka = lst[a];
      int bi = Wi[ka];
      const double* Ab = A + (size_t)ka * BU * BV;
      for (int b = a; b < cnt[p + 1]; b++) {
        int kb = lst[b];
        int bj = Wi[kb];
        const double* Bb = B + (size_t)kb * BU * BV;
        for (int r = 0; r < BU; r++)
          for (int c = 0; c < BU; c++) {
            if (bi == bj && c < r) continue;
            double s = 0;
            for (int m = 0; m < BV; m++) s += Ab[r * BV + m] * Bb[c * BV + m];
            S[((size_t)bi * BU + r) * nu + (size_t)bj * BU + c] -= s;
          }
      }
    }
  }
  free(cnt); free(fill); free(lst);
}

/* nelson/MatrixWtXMultiplier.hpp:49-129: res_j -= sum_k W_k^T x_{i(k)} */
void or_wtx_sub(int nW, int BU, int BV, const int* Wi, const int* Wj, const double* W,
                const double* x, double* res) {
  for (int k = 0; k < nW; k++) {
    const double* w = W + (size_t)k * BU * BV;
    for (int c = 0; c < BV; c++) {
      double s = 0;
      for (int r = 0; r < BU; r++) s += w[r * BV + c] * x[(size_t)Wi[k] * BU + r];
      res[(size_t)Wj[k] * BV + c] -= s;
    }
  }
}

/* simplicial LDL^T in natural order (no pivoting), column-major lower;
 * fails on an exactly zero pivot like Eigen::SimplicialLDLT */
static int ldlt_nopiv_factor(double* a, int n) {
  for (int k = 0; k < n; k++) {
    double d = a[k * n + k];
    for (int j = 0; j < k; j++) d -= a[j * n + k] * a[j * n + k] * a[j * n + j];
    if (d == 0.0) return 0;
    a[k * n + k] = d;
    for (int i = k + 1; i < n; i++) {
      double s = a[k * n + i];
      for (int j = 0; j < k; j++) s -= a[j * n + i] * a[j * n + k] * a[j * n + j];
      a[k * n + i] = s / d;
    }
  }
  return 1;
}

static void ldlt_nopiv_solve(const double* a, int n, double* x) {
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) x[i] -= a[j * n + i] * x[j];
  for (int i = 0; i < n; i++) x[i] /= a[i * n + i];
  for (int j = n - 1; j >= 0; j--)
    for (int i = j + 1; i < n; i++) x[j] -= a[j * n + i] * x[i];
}

/* ------------------------------------------------------------------------ */
/* Schur step: nelson/SolverCholeskySchur.hpp:91-157                        */
/* ------------------------------------------------------------------------ */
double or_max_abs_h_diag(const or_hessian* h) {  /* SolverCholeskySchur.hpp:75-78 */
  double mv = 0, mu = 0;
  for (int p = 0; p < h->NV; p++)
    for (int i = 0; i < h->BV; i++) {
      double v = fabs(h->V[(size_t)p * h->BV * h->BV + i * h->BV + i]);
      if (v > mv) mv = v;
    }
  size_t nu = (size_t)h->BU * h->NU;
  for (size_t i = 0; i < nu; i++) { double v = fabs(h->U[i * nu + i]); if (v > mu) mu = v; }
  return mv > mu ? mv : mu;
}

/* V with off-diagonal blocks (V-V edges, nelson/DoubleSection.h:498-510):
 * SolverCholeskySchur's V solver is then a sparse LDL^T of the whole damped
 * V (solverVType, SolverCholeskySchur.h:82; SolverCholeskySparse.hpp:41-69),
 * restated as the simplicial LDL^T without pivoting in natural order over
 * the nv x nv matrix (column-major lower, upper blocks read as symmetric).
 * The reference orders V by AMD first; an ordering changes the rounding
 * only.  Returns the factor (NULL if a pivot is exactly zero). */
static double* v_full_factor(const or_hessian* h, double rel, double abs_, int* ok) {
  const int BV = h->BV, NV = h->NV;
  const size_t nv = (size_t)BV * NV;
  double* a = (double*)calloc(nv * nv + 1, sizeof(double));
  for (int p = 0; p < NV; p++) {
    const double* v = h->V + (size_t)p * BV * BV;
    for (int i = 0; i < BV; i++)
      for (int j = 0; j <= i; j++) a[((size_t)p * BV + j) * nv + (size_t)p * BV + i] = v[j * BV + i];
  }
  for (int k = 0; k < h->nVo; k++) {   /* block (Voi, Voj), Voi < Voj: lower entry (Voj row, Voi col) */
    const double* v = h->Vo + (size_t)k * BV * BV;
    const size_t r0 = (size_t)h->Voj[k] * BV, c0 = (size_t)h->Voi[k] * BV;
    for (int i = 0; i < BV; i++)
      for (int j = 0; j < BV; j++) a[(c0 + i) * nv + r0 + j] = v[i * BV + j];
  }
  if (rel != 0 || abs_ != 0)    /* the V solver's damping, MatrixDiagInv.hpp:25-140 */
    for (size_t i = 0; i < nv; i++) a[i * nv + i] = a[i * nv + i] * (1 + rel) + abs_;
  *ok = ldlt_nopiv_factor(a, (int)nv);
  return a;
}

/* the reduction with a V factored whole: bS = -bU - W x0, S = (U - W V^-1
 * W^T).upper + damping, Y = V^-1 W^T formed column by column
 * (SolverCholeskySchur.hpp:114-124) */
static void schur_reduce_full_v(const or_hessian* h, double rel, double abs_, const double* vf,
                                const double* xv0, double* S, double* bS) {
  const int BU = h->BU, BV = h->BV, nW = h->nW;
  const size_t nu = (size_t)BU * h->NU, nv = (size_t)BV * h->NV;
  for (size_t i = 0; i < nu; i++) {
    double t = 0;
    for (int k = 0; k < nW; k++) {
      if ((size_t)h->Wi[k] != i / BU) continue;
      const double* w = h->W + (size_t)k * BU * BV + (i % BU) * BV;
      for (int c = 0; c < BV; c++) t += w[c] * xv0[(size_t)h->Wj[k] * BV + c];
    }
    bS[i] = -h->bU[i] - t;
  }
  for (size_t i = 0; i < nu; i++)
    for (size_t j = 0; j < nu; j++) S[i * nu + j] = (j >= i) ? h->U[i * nu + j] : 0.0;
  double* y = (double*)malloc(nv * sizeof(double) + 8);
  for (size_t j = 0; j < nu; j++) {     /* column j of V^-1 W^T */
    memset(y, 0, nv * sizeof(double));
    for (int k = 0; k < nW; k++) {
      if ((size_t)h->Wi[k] != j / BU) continue;
      const double* w = h->W + (size_t)k * BU * BV + (j % BU) * BV;
      for (int c = 0; c < BV; c++) y[(size_t)h->Wj[k] * BV + c] += w[c];
    }
    ldlt_nopiv_solve(vf, (int)nv, y);
    for (int k = 0; k < nW; k++) {      /* S_ij -= (W y)_i for i <= j */
      const size_t i0 = (size_t)h->Wi[k] * BU;
      const double* w = h->W + (size_t)k * BU * BV;
      for (int r = 0; r < BU; r++) {
        if (i0 + r > j) continue;
        double t = 0;
        for (int c = 0; c < BV; c++) t += w[r * BV + c] * y[(size_t)h->Wj[k] * BV + c];
        S[(i0 + r) * nu + j] -= t;
      }
    }
  }
  free(y);
  if (rel != 0 || abs_ != 0)
    for (size_t i = 0; i < nu; i++) { S[i * nu + i] += rel * h->U[i * nu + i]; S[i * nu + i] += abs_; }
}

/* The reduction half of the Schur step (nelson/SolverCholeskySchur.hpp:91-124):
 * the damped V factor per point block (vf, BV x BV column-major lower LDL^T),
 * xv0 = V^-1 (-bV), bS = -bU - W xv0 and the damped S = (U - W V^-1 W^T).upper,
 * row-major nu x nu.  Returns 0 if a V block fails; the caller frees
 * *vf, *xv0, *S and *bS (NULL on failure).  or_schur_assemble and
 * or_schur_step both call it, so the S the CPU baseline times is the S the
 * step factorises. */
static int schur_reduce(const or_hessian* h, double rel, double abs_, double** vf_out, double** xv0_out,
                        double** S_out, double** bS_out) {
  const int BU = h->BU, BV = h->BV, NV = h->NV, nW = h->nW;
  const size_t nu = (size_t)BU * h->NU, nv = (size_t)BV * NV;
  int ok = 1;
  if (h->nVo > 0) {
    double* vf = v_full_factor(h, rel, abs_, &ok);
    double* xv0 = (double*)malloc(nv * sizeof(double) + 8);
    double* S = NULL; double* bS = NULL;
    if (ok) {
      for (size_t i = 0; i < nv; i++) xv0[i] = -h->bV[i];
      ldlt_nopiv_solve(vf, (int)nv, xv0);
      S = (double*)malloc(nu * nu * sizeof(double) + 8);
      bS = (double*)malloc(nu * sizeof(double) + 8);
      schur_reduce_full_v(h, rel, abs_, vf, xv0, S, bS);
    }
    *vf_out = vf; *xv0_out = xv0; *S_out = S; *bS_out = bS;
    return ok;
  }
  /* V solver: per diagonal block, damped LDL^T without pivoting in natural
   * order, failing on an exactly zero pivot: the simplicial LDL^T that
   * SolverCholeskySchur uses as its block-diagonal V solver
   * (SolverCholeskySparse.hpp:41-69, SolverCholeskySparse.h:16):
   * x0 = V^-1 (-bV) */
  double* vf = (double*)malloc((size_t)NV * BV * BV * sizeof(double) + 8);
  double* xv0 = (double*)malloc(nv * sizeof(double) + 8);
  for (int p = 0; p < NV && ok; p++) {
    const double* v = h->V + (size_t)p * BV * BV;
    double* a = vf + (size_t)p * BV * BV;
    for (int i = 0; i < BV; i++)
      for (int j = 0; j < BV; j++) a[j * BV + i] = (i >= j) ? v[j * BV + i] : 0.0;
    if (rel != 0 || abs_ != 0)
      for (int i = 0; i < BV; i++) a[i * BV + i] = a[i * BV + i] * (1 + rel) + abs_;
    ok = ldlt_nopiv_factor(a, BV);
    if (ok) {
      for (int i = 0; i < BV; i++) xv0[(size_t)p * BV + i] = -h->bV[(size_t)p * BV + i];
      ldlt_nopiv_solve(a, BV, xv0 + (size_t)p * BV);
    }
  }
  double* S = NULL; double* bS = NULL;
  if (ok) {
    /* bS = -bU - W * x0   (SolverCholeskySchur.hpp:114) */
    bS = (double*)malloc(nu * sizeof(double) + 8);
    double* t = (double*)calloc(nu + 1, sizeof(double));
    for (int k = 0; k < nW; k++) {
      const double* w = h->W + (size_t)k * BU * BV;
      for (int r = 0; r < BU; r++) {
        double s = 0;
        for (int c = 0; c < BV; c++) s += w[r * BV + c] * xv0[(size_t)h->Wj[k] * BV + c];
        t[(size_t)h->Wi[k] * BU + r] += s;
      }
    }
    for (size_t i = 0; i < nu; i++) bS[i] = -h->bU[i] - t[i];
    free(t);
    /* S = (U - W * V^-1 W^T).upper  (:118), Z_k = V^-1 W_k^T via the V factor */
    double* Z = (double*)malloc((size_t)nW * BU * BV * sizeof(double) + 8); /* Z_k stored BU x BV = (V^-1 W_k^T)^T */
    for (int k = 0; k < nW; k++) {
      const double* w = h->W + (size_t)k * BU * BV;
      int p = h->Wj[k];
      for (int r = 0; r < BU; r++) {
        double x[8];
        for (int c = 0; c < BV; c++) x[c] = w[r * BV + c];
        ldlt_nopiv_solve(vf + (size_t)p * BV * BV, BV, x);
        for (int c = 0; c < BV; c++) Z[(size_t)k * BU * BV + r * BV + c] = x[c];
      }
    }
    S = (double*)malloc(nu * nu * sizeof(double) + 8);
    or_wwt(h->NU, BU, BV, nW, h->Wi, h->Wj, h->W, Z, h->U, S);
    free(Z);
    /* damping (:121-124) */
    if (rel != 0 || abs_ != 0)
      for (size_t i = 0; i < nu; i++) { S[i * nu + i] += rel * h->U[i * nu + i]; S[i * nu + i] += abs_; }
  }
  *vf_out = vf; *xv0_out = xv0; *S_out = S; *bS_out = bS;
  return ok;
}

/* the reduction alone (no factorisation or solves): the CPU baseline times
 * the factorisation of the assembled S separately.  S_out (nu x nu) and
 * bS_out may not be NULL. */
int or_schur_assemble(const or_hessian* h, double rel, double abs_, double* S_out, double* bS_out) {
  const size_t nu = (size_t)h->BU * h->NU;
  double *vf, *xv0, *S, *bS;
  int ok = schur_reduce(h, rel, abs_, &vf, &xv0, &S, &bS);
  if (ok) {
    memcpy(S_out, S, nu * nu * sizeof(double));
    memcpy(bS_out, bS, nu * sizeof(double));
  }
  free(vf); free(xv0); free(S); free(bS);
  return ok;
}

int or_schur_step(const or_hessian* h, double rel, double abs_, double* xU, double* xV,
                  double* S_out, double* bS_out) {
  const int BU = h->BU, BV = h->BV, NV = h->NV, nW = h->nW;
  const size_t nu = (size_t)BU * h->NU, nv = (size_t)BV * NV;
  double *vf, *xv0, *S, *bS;
  int* stp = NULL;
  int ok = schur_reduce(h, rel, abs_, &vf, &xv0, &S, &bS);
  if (ok) {
    if (S_out) memcpy(S_out, S, nu * nu * sizeof(double));
    if (bS_out) memcpy(bS_out, bS, nu * sizeof(double));
    /* factorize (:136); row-major upper == column-major lower */
    stp = (int*)malloc(nu * sizeof(int) + 8);
    ok = or_ldlt_factor(S, (int)nu, stp);
  }
  if (ok) {
    /* xU = S^-1 bS (:140) */
    memcpy(xU, bS, nu * sizeof(double));
    or_ldlt_solve(S, (int)nu, stp, xU);
    /* bVt = -bV - W^T xU (:144); xV = V^-1 bVt (:147) */
    for (size_t i = 0; i < nv; i++) xV[i] = 0.0;
    or_wtx_sub(nW, BU, BV, h->Wi, h->Wj, h->W, xU, xV);
    for (size_t i = 0; i < nv; i++) xV[i] = -h->bV[i] + xV[i];
    if (h->nVo > 0)
      ldlt_nopiv_solve(vf, (int)nv, xV);
    else
      for (int p = 0; p < NV; p++)
        ldlt_nopiv_solve(vf + (size_t)p * BV * BV, BV, xV + (size_t)p * BV);
  } else {
    memset(xU, 0, nu * sizeof(double));        /* :151-154 */
    memset(xV, 0, nv * sizeof(double));
  }
  free(vf); free(xv0); free(S); free(bS); free(stp);
  return ok;
}

/* nelson/DoubleSection.h:300-311 */
double or_rho(const or_hessian* h, double mu, const double* xU, const double* xV,
              double old_chi2, double new_chi2) {
  size_t nu = (size_t)h->BU * h->NU, nv = (size_t)h->BV * h->NV;
  double num = old_chi2 - new_chi2;
  double du = 0, dv = 0;
  for (size_t i = 0; i < nu; i++) du += xU[i] * (mu * xU[i] - h->bU[i]);
  for (size_t i = 0; i < nv; i++) dv += xV[i] * (mu * xV[i] - h->bV[i]);
  return num / (du + dv);
}

/* ------------------------------------------------------------------------ */
/* LM / GN drivers                                                          */
/* ------------------------------------------------------------------------ */
void or_lm_default_settings(or_lm_settings* s) {    /* nelson/LevenbergMarquardt.cpp:6-21 */
  s->maxNumIt = 100; s->maxNumSubIt = 10; s->minNumIt = 0;
  s->epsBVector = 1e-15; s->epsIncSquare = 1e-15 * 1e-15; s->epsChi2 = 1e-15;
  s->tau = 1e-3; s->penaltyV = 2.0; s->initV = 2.0;
  s->maxScaleFactor = 2.0 / 3.0; s->minScaleFactor = 1.0 / 3.0;
}

void or_gn_default_settings(or_gn_settings* s) {    /* nelson/GaussNewton.cpp:12-22 */
  s->maxNumIt = 100; s->minNumIt = 0;
  s->epsBVector = 1e-15; s->epsIncSquare = 1e-15 * 1e-15; s->epsChi2 = 1e-15;
  s->absLambda = 0; s->relLambda = 0;
}

static void log_add(or_lm_result* r, int it, int subit, double chi2, double mu) {
  if (r->log && r->n_log < r->log_cap) {
    r->log[r->n_log].it = it; r->log[r->n_log].subit = subit;
    r->log[r->n_log].chi2 = chi2; r->log[r->n_log].mu = mu;
  }
  r->n_log++;
}

static void dx_add(or_lm_result* r, int* count, const double* xU, size_t nu, const double* xV, size_t nv) {
  if (r->dx_log && (size_t)(*count + 1) * (nu + nv) <= (size_t)r->dx_log_cap) {
    memcpy(r->dx_log + (size_t)(*count) * (nu + nv), xU, nu * sizeof(double));
    memcpy(r->dx_log + (size_t)(*count) * (nu + nv) + nu, xV, nv * sizeof(double));
  }
  (*count)++;
  r->dx_len = (int)(nu + nv);
}

/* nelson/LevenbergMarquardt.hpp:35-146 */
int or_lm_solve(const or_lm_settings* s, or_ops* op, or_lm_result* res) {
  int iter = -1, subiter = -1, ndx = 0;
  res->n_log = 0;
  op->update(op->ctx, 1);
  or_hessian* h = op->hessian(op->ctx);
  log_add(res, iter, subiter, h->chi2, 0.0);
  res->final_chi2 = h->chi2;
  res->iterations = 0;
  if (or_hessian_max_abs_b(h) <= s->epsBVector) { res->reason = OR_SUCCESS_AT_START; return res->reason; }
  op->backup(op->ctx);
  double oldChi2 = h->chi2;
  size_t nu = (size_t)h->BU * h->NU, nv = (size_t)h->BV * h->NV;
  double* xU = (double*)calloc(nu + 1, sizeof(double));
  double* xV = (double*)calloc(nv + 1, sizeof(double));
  double mu = s->tau * or_max_abs_h_diag(h);
  double curV = s->initV;
  int reason = OR_MAX_ITERATION_REACHED;
  iter = 0;
  while (iter < s->maxNumIt) {
    subiter = 0;
    int toRepeat;
    do {
      int solveOk = or_schur_step(h, 0.0, mu, xU, xV, NULL, NULL);
      dx_add(res, &ndx, xU, nu, xV, nv);
      double n2 = 0;
      for (size_t i = 0; i < nu; i++) n2 += xU[i] * xU[i];
      double n2v = 0;
      for (size_t i = 0; i < nv; i++) n2v += xV[i] * xV[i];
      n2 += n2v;
      if (solveOk && iter >= s->minNumIt && n2 <= s->epsIncSquare) { reason = OR_SUCCESS_SMALL_INCREMENT; goto done; }
      double rho = -1;
      if (solveOk) {
        op->oplus(op->ctx, xU, xV);
        double bchi2 = h->chi2;
        (void)bchi2;
        op->update(op->ctx, 0);
        log_add(res, iter, subiter, h->chi2, mu);
        rho = or_rho(h, mu, xU, xV, oldChi2, h->chi2);
      } else {
        log_add(res, iter, subiter, -1.0, mu);
      }
      if (solveOk && rho > 0) {
        if (iter > s->minNumIt && or_hessian_max_abs_b(h) <= s->epsBVector) { reason = OR_SUCCESS_EPS_B; goto done; }
        if (iter > s->minNumIt && h->chi2 < s->epsChi2) { reason = OR_SUCCESS_SMALL_RESIDUALS; goto done; }
        double fc = 2 * rho - 1;
        double alpha = fmin(s->maxScaleFactor, 1 - fc * fc * fc);
        alpha = fmax(s->minScaleFactor, alpha);
        mu = mu * alpha;
        curV = s->initV;
      } else {
        op->rollback(op->ctx);
        mu = mu * curV;
        curV = curV * s->penaltyV;
      }
      subiter++;
      if (subiter >= s->maxNumSubIt) { reason = OR_MAX_SUBITERATION_REACHED; goto done; }
      toRepeat = (rho <= 0 || !solveOk);
    } while (toRepeat);
    op->update(op->ctx, 1);
    op->backup(op->ctx);
    oldChi2 = h->chi2;
    iter++;
  }
done:
  res->reason = reason;
  res->iterations = iter;
  res->final_chi2 = h->chi2;
  free(xU); free(xV);
  return reason;
}

/* nelson/GaussNewton.hpp:31-127 */
int or_gn_solve(const or_gn_settings* s, or_ops* op, or_lm_result* res) {
  int iter = -1, ndx = 0;
  res->n_log = 0;
  op->update(op->ctx, 1);
  or_hessian* h = op->hessian(op->ctx);
  log_add(res, iter, -1, h->chi2, 0.0);
  res->final_chi2 = h->chi2;
  res->iterations = 0;
  if (or_hessian_max_abs_b(h) <= s->epsBVector) { res->reason = OR_SUCCESS_AT_START; return res->reason; }
  size_t nu = (size_t)h->BU * h->NU, nv = (size_t)h->BV * h->NV;
  double* xU = (double*)calloc(nu + 1, sizeof(double));
  double* xV = (double*)calloc(nv + 1, sizeof(double));
  int reason = OR_MAX_ITERATION_REACHED;
  iter = 0;
  while (iter < s->maxNumIt) {
    int solveOk = or_schur_step(h, s->relLambda, s->absLambda, xU, xV, NULL, NULL);
    dx_add(res, &ndx, xU, nu, xV, nv);
    if (!solveOk) { reason = OR_SOLVER_FAILURE; break; }
    double n2 = 0, n2v = 0;
    for (size_t i = 0; i < nu; i++) n2 += xU[i] * xU[i];
    for (size_t i = 0; i < nv; i++) n2v += xV[i] * xV[i];
    if (iter >= s->minNumIt && n2 + n2v <= s->epsIncSquare) { reason = OR_SUCCESS_SMALL_INCREMENT; break; }
    op->oplus(op->ctx, xU, xV);
    op->update(op->ctx, 1);
    log_add(res, iter, 0, h->chi2, 0.0);
    if (iter > s->minNumIt && or_hessian_max_abs_b(h) <= s->epsBVector) { reason = OR_SUCCESS_EPS_B; break; }
    if (iter > s->minNumIt && h->chi2 < s->epsChi2) { reason = OR_SUCCESS_SMALL_RESIDUALS; break; }
    iter++;
  }
  res->reason = reason;
  res->iterations = iter;
  res->final_chi2 = h->chi2;
  free(xU); free(xV);
  return reason;
}

/* ------------------------------------------------------------------------ */
/* Rotations: Rodrigues and the SE(3) left update exp(delta) * T            */
/* (the sample's camera update, nelson/samples/BASchurPerfomances.cpp:168)  */
/* ------------------------------------------------------------------------ */
static void so3_coeffs(double th2, double* A, double* B, double* Cc) {
  double th = sqrt(th2);
  if (th < 1e-3) {
    *A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    *B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    *Cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
  } else {
    double s = sin(th), h = sin(0.5 * th);
    *A = s / th;
    *B = 2.0 * h * h / th2;
    *Cc = (th - s) / (th2 * th);
  }
}

void or_rodrigues(const double* w, double* R) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double A, B, Cc;
  so3_coeffs(th2, &A, &B, &Cc);
  double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
      R[i * 3 + j] = (i == j ? 1.0 : 0.0) + A * K[i * 3 + j] + B * k2;
    }
}

/* T <- exp(delta) * T with delta = (omega, upsilon):
 * R <- Exp(omega) R,  t <- Exp(omega) t + V(omega) upsilon */
void or_se3_left_update(const double* d, double* R, double* t) {
  const double* w = d;
  const double* u = d + 3;
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double A, B, Cc;
  so3_coeffs(th2, &A, &B, &Cc);
  double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double E[9], Vm[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
      double id = (i == j ? 1.0 : 0.0);
      E[i * 3 + j] = id + A * K[i * 3 + j] + B * k2;
      Vm[i * 3 + j] = id + B * K[i * 3 + j] + Cc * k2;
    }
  double Rn[9], tn[3];
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      Rn[i * 3 + j] = E[i * 3 + 0] * R[0 * 3 + j] + E[i * 3 + 1] * R[1 * 3 + j] + E[i * 3 + 2] * R[2 * 3 + j];
    tn[i] = E[i * 3 + 0] * t[0] + E[i * 3 + 1] * t[1] + E[i * 3 + 2] * t[2]
          + Vm[i * 3 + 0] * u[0] + Vm[i * 3 + 1] * u[1] + Vm[i * 3 + 2] * u[2];
  }
  memcpy(R, Rn, sizeof(Rn));
  memcpy(t, tn, sizeof(tn));
}

/* ------------------------------------------------------------------------ */
/* BA section                                                               */
/* ------------------------------------------------------------------------ */
struct or_ba {
  int C, P, E;
  int *cam, *pt;
  double *obs, *w;
  double f, cx, cy, huber;
  int *cam_var, *pt_var;      /* user index -> variable index or -1 */
  int NU, NV;
  double *cams, *pts, *cams_bak, *pts_bak;   /* cams: 12 per camera (R, t) */
  int* edge_w;                /* edge -> W block or -1 */
  double *er, *eJc, *eJp, *eom; /* per-edge r (2), Jc (12), Jp (6), IRLS weight */
  int ncp, npp;                 /* unary priors (or_ba_set_priors) */
  int *cp_id, *pp_id;
  double *cp_p0, *cp_w, *pp_p0, *pp_w;
  double* intr;                 /* per camera (f, k1, k2) of the BAL model, or NULL */
  or_hessian H;
};

static int cmp_pair(const void* a, const void* b) {
  const long long* x = (const long long*)a;
  const long long* y = (const long long*)b;
  return (*x < *y) ? -1 : (*x > *y);
}

or_ba* or_ba_create(int C, int P, int E, const int* cam, const int* pt, const double* obs,
                    const double* weight, const unsigned char* cam_fixed,
                    const unsigned char* pt_fixed, double f, double cx, double cy,
                    double huber_delta, const double* cams0, const double* pts0) {
  or_ba* b = (or_ba*)calloc(1, sizeof(or_ba));
  b->C = C; b->P = P; b->E = E;
  b->f = f; b->cx = cx; b->cy = cy; b->huber = huber_delta;
  b->cam = (int*)malloc(sizeof(int) * (size_t)(E + 1));
  b->pt = (int*)malloc(sizeof(int) * (size_t)(E + 1));
  b->obs = (double*)malloc(sizeof(double) * 2 * (size_t)(E + 1));
  b->w = (double*)malloc(sizeof(double) * (size_t)(E + 1));
  memcpy(b->cam, cam, sizeof(int) * (size_t)E);
  memcpy(b->pt, pt, sizeof(int) * (size_t)E);
  memcpy(b->obs, obs, sizeof(double) * 2 * (size_t)E);
  for (int e = 0; e < E; e++) b->w[e] = weight ? weight[e] : 1.0;
  b->cam_var = (int*)malloc(sizeof(int) * (size_t)(C + 1));
  b->pt_var = (int*)malloc(sizeof(int) * (size_t)(P + 1));
  int nu = 0, nv = 0;
  for (int c = 0; c < C; c++) b->cam_var[c] = (cam_fixed && cam_fixed[c]) ? -1 : nu++;
  for (int p = 0; p < P; p++) b->pt_var[p] = (pt_fixed && pt_fixed[p]) ? -1 : nv++;
  b->NU = nu; b->NV = nv;
  b->cams = (double*)malloc(sizeof(double) * 12 * (size_t)(C + 1));
  b->cams_bak = (double*)malloc(sizeof(double) * 12 * (size_t)(C + 1));
  b->pts = (double*)malloc(sizeof(double) * 3 * (size_t)(P + 1));
  b->pts_bak = (double*)malloc(sizeof(double) * 3 * (size_t)(P + 1));
  for (int c = 0; c < C; c++) {
    or_rodrigues(cams0 + 6 * c, b->cams + 12 * c);
    memcpy(b->cams + 12 * c + 9, cams0 + 6 * c + 3, 3 * sizeof(double));
  }
  memcpy(b->pts, pts0, sizeof(double) * 3 * (size_t)P);
  /* W structure: unique (variable cam, variable point) pairs in row-major
   * UID order (nelson/DoubleSection.hpp:157-191) */
  long long* keys = (long long*)malloc(sizeof(long long) * (size_t)(E + 1));
  int nk = 0;
  for (int e = 0; e < E; e++) {
    int ci = b->cam_var[cam[e]], pj = b->pt_var[pt[e]];
    if (ci >= 0 && pj >= 0) keys[nk++] = (long long)ci * (long long)(nv + 1) + pj;
  }
  qsort(keys, (size_t)nk, sizeof(long long), cmp_pair);
  int nW = 0;
  for (int k = 0; k < nk; k++) if (k == 0 || keys[k] != keys[k - 1]) keys[nW++] = keys[k];
  or_hessian_alloc(&b->H, 6, 3, nu, nv, nW);
  for (int k = 0; k < nW; k++) { b->H.Wi[k] = (int)(keys[k] / (nv + 1)); b->H.Wj[k] = (int)(keys[k] % (nv + 1)); }
  b->edge_w = (int*)malloc(sizeof(int) * (size_t)(E + 1));
  for (int e = 0; e < E; e++) {
    int ci = b->cam_var[cam[e]], pj = b->pt_var[pt[e]];
    b->edge_w[e] = -1;
    if (ci >= 0 && pj >= 0) {
      long long key = (long long)ci * (long long)(nv + 1) + pj;
      long long* hit = (long long*)bsearch(&key, keys, (size_t)nW, sizeof(long long), cmp_pair);
      b->edge_w[e] = (int)(hit - keys);
    }
  }
  free(keys);
  b->er = (double*)malloc(sizeof(double) * 2 * (size_t)(E + 1));
  b->eJc = (double*)malloc(sizeof(double) * 12 * (size_t)(E + 1));
  b->eJp = (double*)malloc(sizeof(double) * 6 * (size_t)(E + 1));
  b->eom = (double*)malloc(sizeof(double) * (size_t)(E + 1));
  return b;
}

void or_ba_destroy(or_ba* b) {
  if (!b) return;
  free(b->cam); free(b->pt); free(b->obs); free(b->w); free(b->cam_var); free(b->pt_var);
  free(b->cams); free(b->pts); free(b->cams_bak); free(b->pts_bak); free(b->edge_w);
  free(b->er); free(b->eJc); free(b->eJp); free(b->eom);
  free(b->cp_id); free(b->cp_p0); free(b->cp_w); free(b->pp_id); free(b->pp_p0); free(b->pp_w);
  free(b->intr);
  or_hessian_free(&b->H);
  free(b);
}

/* principal log of a rotation: theta = atan2(|vee(R - R^T)| / 2, (tr R - 1) / 2),
 * omega = theta / (2 sin theta) vee(R - R^T); near theta = pi the axis comes
 * from a a^T = ((R + R^T) / 2 - cos theta I) / (1 - cos theta) (its largest
 * diagonal entry's column), signed like vee(R - R^T) */
void or_so3_log(const double* R, double* w) {
  double v[3] = {R[7] - R[5], R[2] - R[6], R[3] - R[1]};
  double c = 0.5 * (R[0] + R[4] + R[8] - 1.0);
  double s = 0.5 * sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double th = atan2(s, c);
  if (th < 1e-4) {
    double k = 0.5 + th * th / 12.0;
    for (int i = 0; i < 3; i++) w[i] = k * v[i];
  } else if (3.14159265358979323846 - th > 1e-4) {
    double k = th / (2.0 * s);
    for (int i = 0; i < 3; i++) w[i] = k * v[i];
  } else {
    double B[9];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) B[i * 3 + k] = (0.5 * (R[i * 3 + k] + R[k * 3 + i]) - (i == k ? c : 0.0)) / (1.0 - c);
    int j = 0;
    for (int i = 1; i < 3; i++) if (B[i * 4] > B[j * 4]) j = i;
    double a[3];
    double d = sqrt(B[j * 4]);
    for (int i = 0; i < 3; i++) a[i] = B[i * 3 + j] / d;
    double sg = a[0] * v[0] + a[1] * v[1] + a[2] * v[2] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < 3; i++) w[i] = sg * th * a[i];
  }
}

/* Jl^-1(w) = I - K / 2 + (1 / th^2 - (1 + cos th) / (2 th sin th)) K^2, K = [w]x */
void or_so3_jl_inv(const double* w, double* J) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = sqrt(th2);
  double g = th < 1e-4 ? 1.0 / 12.0 + th2 / 720.0 : 1.0 / th2 - (1.0 + cos(th)) / (2.0 * th * sin(th));
  double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
      J[i * 3 + j] = (i == j ? 1.0 : 0.0) - 0.5 * K[i * 3 + j] + g * k2;
    }
}

void or_ba_set_priors(or_ba* b, int ncp, const int* cam_ids, const double* cam_p0, const double* cam_w,
                      int npp, const int* pt_ids, const double* pt_p0, const double* pt_w) {
  free(b->cp_id); free(b->cp_p0); free(b->cp_w); free(b->pp_id); free(b->pp_p0); free(b->pp_w);
  b->ncp = ncp; b->npp = npp;
  b->cp_id = (int*)malloc(sizeof(int) * (size_t)(ncp + 1));
  b->cp_p0 = (double*)malloc(sizeof(double) * 6 * (size_t)(ncp + 1));
  b->cp_w = (double*)malloc(sizeof(double) * 6 * (size_t)(ncp + 1));
  b->pp_id = (int*)malloc(sizeof(int) * (size_t)(npp + 1));
  b->pp_p0 = (double*)malloc(sizeof(double) * 3 * (size_t)(npp + 1));
  b->pp_w = (double*)malloc(sizeof(double) * 3 * (size_t)(npp + 1));
  if (ncp) {
    memcpy(b->cp_id, cam_ids, sizeof(int) * (size_t)ncp);
    memcpy(b->cp_p0, cam_p0, sizeof(double) * 6 * (size_t)ncp);
    memcpy(b->cp_w, cam_w, sizeof(double) * 6 * (size_t)ncp);
  }
  if (npp) {
    memcpy(b->pp_id, pt_ids, sizeof(int) * (size_t)npp);
    memcpy(b->pp_p0, pt_p0, sizeof(double) * 3 * (size_t)npp);
    memcpy(b->pp_w, pt_w, sizeof(double) * 3 * (size_t)npp);
  }
}

/* camera prior k: residual r (6) and, with jac, J (6x6 row-major, weighted) */
static double cam_prior(const or_ba* b, int k, int jac, double* r, double* J) {
  const double* R = b->cams + 12 * b->cp_id[k];
  const double* t = R + 9;
  const double* p0 = b->cp_p0 + 6 * k;
  const double* w = b->cp_w + 6 * k;
  double om[3];
  or_so3_log(R, om);
  double cost = 0.0;
  for (int i = 0; i < 3; i++) r[i] = w[i] * (om[i] - p0[i]);
  for (int i = 0; i < 3; i++) r[3 + i] = w[3 + i] * (t[i] - p0[3 + i]);
  for (int i = 0; i < 6; i++) cost += r[i] * r[i];
  if (jac) {
    double Jl[9];
    or_so3_jl_inv(om, Jl);
    /* d t / d omega = -[t]x, d t / d upsilon = I */
    double Tx[9] = {0, t[2], -t[1], -t[2], 0, t[0], t[1], -t[0], 0};
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

void or_ba_set_intrinsics(or_ba* b, const double* intr) {
  free(b->intr);
  b->intr = NULL;
  if (intr) {
    b->intr = (double*)malloc(sizeof(double) * 3 * (size_t)b->C);
    memcpy(b->intr, intr, sizeof(double) * 3 * (size_t)b->C);
  }
}

/* the BA reprojection edge: r = pi(R X + t) - z.  pi is the pinhole with
 * focal f and principal point (cx, cy), or with per-camera intrinsics the
 * BAL model: q = -(x, y) / z, pi = f (1 + k1 |q|^2 + k2 |q|^4) q.
 * Jacobians w.r.t. the left SE(3) perturbation (omega, upsilon) of the camera
 * and w.r.t. the point.  Cost per edge: w * rho_huber(|r|^2) (rho = identity
 * without Huber); IRLS weight returned. */
static double ba_edge(const or_ba* b, int e, int jac, double* r, double* Jc, double* Jp, double* om) {
  const double* R = b->cams + 12 * b->cam[e];
  const double* t = R + 9;
  const double* X = b->pts + 3 * b->pt[e];
  double Xc[3];
  for (int i = 0; i < 3; i++) Xc[i] = R[i * 3 + 0] * X[0] + R[i * 3 + 1] * X[1] + R[i * 3 + 2] * X[2] + t[i];
  double iz = 1.0 / Xc[2];
  double u, v, D[6];
  if (!b->intr) {
    u = b->f * Xc[0] * iz + b->cx;
    v = b->f * Xc[1] * iz + b->cy;
    double a = b->f * iz;
    D[0] = a; D[1] = 0.0; D[2] = -a * Xc[0] * iz;
    D[3] = 0.0; D[4] = a; D[5] = -a * Xc[1] * iz;
  } else {
    const double* K = b->intr + 3 * b->cam[e];
    double qx = -Xc[0] * iz, qy = -Xc[1] * iz;
    double s = qx * qx + qy * qy;
    double dist = 1.0 + K[1] * s + K[2] * s * s;
    double ddist = K[1] + 2.0 * K[2] * s;       /* d dist / d s */
    u = K[0] * dist * qx;
    v = K[0] * dist * qy;
    /* chain rule: d pi / d q, then d q / d Xc */
    double du_dqx = K[0] * (dist + qx * ddist * 2.0 * qx), du_dqy = K[0] * qx * ddist * 2.0 * qy;
    double dv_dqx = K[0] * qy * ddist * 2.0 * qx, dv_dqy = K[0] * (dist + qy * ddist * 2.0 * qy);
    double dqx_dx = -iz, dqx_dz = Xc[0] * iz * iz, dqy_dy = -iz, dqy_dz = Xc[1] * iz * iz;
    D[0] = du_dqx * dqx_dx; D[1] = du_dqy * dqy_dy; D[2] = du_dqx * dqx_dz + du_dqy * dqy_dz;
    D[3] = dv_dqx * dqx_dx; D[4] = dv_dqy * dqy_dy; D[5] = dv_dqx * dqx_dz + dv_dqy * dqy_dz;
  }
  r[0] = u - b->obs[2 * e];
  r[1] = v - b->obs[2 * e + 1];
  double s = r[0] * r[0] + r[1] * r[1];
  double cost, psi;
  if (b->huber > 0 && s > b->huber * b->huber) {
    double n = sqrt(s);
    cost = 2.0 * b->huber * n - b->huber * b->huber;
    psi = b->huber / n;
  } else {
    cost = s;
    psi = 1.0;
  }
  *om = b->w[e] * psi;
  if (jac) {
    /* d Xc / d (omega, upsilon) = [ -[Xc]x | I ] */
    double M[18] = {0.0, Xc[2], -Xc[1], 1.0, 0.0, 0.0,
                    -Xc[2], 0.0, Xc[0], 0.0, 1.0, 0.0,
                    Xc[1], -Xc[0], 0.0, 0.0, 0.0, 1.0};
    for (int i = 0; i < 2; i++) {
      for (int k = 0; k < 6; k++)
        Jc[i * 6 + k] = D[i * 3 + 0] * M[0 * 6 + k] + D[i * 3 + 1] * M[1 * 6 + k] + D[i * 3 + 2] * M[2 * 6 + k];
      for (int k = 0; k < 3; k++)
        Jp[i * 3 + k] = D[i * 3 + 0] * R[0 * 3 + k] + D[i * 3 + 1] * R[1 * 3 + k] + D[i * 3 + 2] * R[2 * 3 + k];
    }
  }
  return b->w[e] * cost;
}

/* DoubleSection::update(bool) (nelson/DoubleSection.hpp:452-467):
 * clear (hessian only), evaluate every edge with the chi2 sum in edge order
 * (nelson/BaseSection.cpp:88-169), then run every block's updater list, which
 * holds edges in reverse insertion order (emplace_front,
 * nelson/DoubleSection.hpp:349-363), using the edge's
 * updateH11/H22/H12 (nelson/samples/BASchurPerfomances.cpp:217-232). */
void or_ba_update(or_ba* b, int hessian) {
  or_hessian* H = &b->H;
  if (hessian) or_hessian_clear(H);
  double chi2 = 0;
  for (int e = 0; e < b->E; e++)
    chi2 += ba_edge(b, e, hessian, b->er + 2 * e, b->eJc + 12 * e, b->eJp + 6 * e, b->eom + e);
  /* unary priors after the binary edges: cameras, then points */
  double rc[6], Jcp[36];
  for (int k = 0; k < b->ncp; k++) chi2 += cam_prior(b, k, 0, rc, Jcp);
  for (int k = 0; k < b->npp; k++) {
    const double* X = b->pts + 3 * b->pp_id[k];
    for (int i = 0; i < 3; i++) {
      double ri = b->pp_w[3 * k + i] * (X[i] - b->pp_p0[3 * k + i]);
      chi2 += ri * ri;
    }
  }
  H->chi2 = chi2;
  if (!hessian) return;
  size_t nu = (size_t)6 * H->NU;
  for (int e = b->E - 1; e >= 0; e--) {
    int ci = b->cam_var[b->cam[e]], pj = b->pt_var[b->pt[e]];
    const double* J = b->eJc + 12 * e;
    const double* G = b->eJp + 6 * e;
    const double* r = b->er + 2 * e;
    double om = b->eom[e];
    if (ci >= 0) {
      for (int i = 0; i < 6; i++) {
        for (int j = i; j < 6; j++)
          H->U[((size_t)ci * 6 + i) * nu + (size_t)ci * 6 + j] += om * (J[i] * J[j] + J[6 + i] * J[6 + j]);
        H->bU[(size_t)ci * 6 + i] += om * (J[i] * r[0] + J[6 + i] * r[1]);
      }
    }
    if (pj >= 0) {
      double* V = H->V + (size_t)pj * 9;
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) V[i * 3 + j] += om * (G[i] * G[j] + G[3 + i] * G[3 + j]);
        H->bV[(size_t)pj * 3 + i] += om * (G[i] * r[0] + G[3 + i] * r[1]);
      }
    }
    if (ci >= 0 && pj >= 0) {
      double* W = H->W + (size_t)b->edge_w[e] * 18;
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 3; j++) W[i * 3 + j] += om * (J[i] * G[j] + J[6 + i] * G[3 + j]);
    }
  }
  /* the priors' diagonal blocks and gradient segments (variable nodes) */
  for (int k = 0; k < b->ncp; k++) {
    int ci = b->cam_var[b->cp_id[k]];
    if (ci < 0) continue;
    cam_prior(b, k, 1, rc, Jcp);
    for (int i = 0; i < 6; i++) {
      for (int j = i; j < 6; j++) {
        double h = 0.0;
        for (int m = 0; m < 6; m++) h += Jcp[m * 6 + i] * Jcp[m * 6 + j];
        H->U[((size_t)ci * 6 + i) * nu + (size_t)ci * 6 + j] += h;
      }
      double g = 0.0;
      for (int m = 0; m < 6; m++) g += Jcp[m * 6 + i] * rc[m];
      H->bU[(size_t)ci * 6 + i] += g;
    }
  }
  for (int k = 0; k < b->npp; k++) {
    int pj = b->pt_var[b->pp_id[k]];
    if (pj < 0) continue;
    const double* X = b->pts + 3 * b->pp_id[k];
    for (int i = 0; i < 3; i++) {
      double w = b->pp_w[3 * k + i];
      H->V[(size_t)pj * 9 + i * 4] += w * w;
      H->bV[(size_t)pj * 3 + i] += w * w * (X[i] - b->pp_p0[3 * k + i]);
    }
  }
}

/* oplus of the sample (nelson/samples/BASchurPerfomances.cpp:163-170):
 * points additive, cameras exp(delta) * T; identity camera ordering */
void or_ba_oplus(or_ba* b, const double* xU, const double* xV) {
  for (int p = 0; p < b->P; p++) {
    int v = b->pt_var[p];
    if (v < 0) continue;
    for (int i = 0; i < 3; i++) b->pts[3 * p + i] += xV[3 * v + i];
  }
  for (int c = 0; c < b->C; c++) {
    int u = b->cam_var[c];
    if (u < 0) continue;
    or_se3_left_update(xU + 6 * u, b->cams + 12 * c, b->cams + 12 * c + 9);
  }
}

static void ba_update_cb(void* ctx, int hessian) { or_ba_update((or_ba*)ctx, hessian); }
static void ba_oplus_cb(void* ctx, const double* xU, const double* xV) { or_ba_oplus((or_ba*)ctx, xU, xV); }
static void ba_backup_cb(void* ctx) {
  or_ba* b = (or_ba*)ctx;
  memcpy(b->cams_bak, b->cams, sizeof(double) * 12 * (size_t)b->C);
  memcpy(b->pts_bak, b->pts, sizeof(double) * 3 * (size_t)b->P);
}
static void ba_rollback_cb(void* ctx) {
  or_ba* b = (or_ba*)ctx;
  memcpy(b->cams, b->cams_bak, sizeof(double) * 12 * (size_t)b->C);
  memcpy(b->pts, b->pts_bak, sizeof(double) * 3 * (size_t)b->P);
}
static or_hessian* ba_hessian_cb(void* ctx) { return &((or_ba*)ctx)->H; }

or_ops or_ba_ops(or_ba* b) {
  or_ops o;
  o.ctx = b; o.update = ba_update_cb; o.oplus = ba_oplus_cb;
  o.backup = ba_backup_cb; o.rollback = ba_rollback_cb; o.hessian = ba_hessian_cb;
  return o;
}

or_hessian* or_ba_hessian(or_ba* b) { return &b->H; }
int or_ba_num_var_cams(const or_ba* b) { return b->NU; }
int or_ba_num_var_pts(const or_ba* b) { return b->NV; }

void or_ba_get_state(const or_ba* b, double* cams_rt, double* pts) {
  memcpy(cams_rt, b->cams, sizeof(double) * 12 * (size_t)b->C);
  memcpy(pts, b->pts, sizeof(double) * 3 * (size_t)b->P);
}

void or_ba_set_state_rt(or_ba* b, const double* cams_rt, const double* pts) {
  memcpy(b->cams, cams_rt, sizeof(double) * 12 * (size_t)b->C);
  memcpy(b->pts, pts, sizeof(double) * 3 * (size_t)b->P);
}

void or_ba_edge_jacobians(or_ba* b, double* r, double* Jc, double* Jp) {
  double om;
  for (int e = 0; e < b->E; e++) ba_edge(b, e, 1, r + 2 * e, Jc + 12 * e, Jp + 6 * e, &om);
}
