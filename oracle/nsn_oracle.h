/*
 * nsn_oracle.h -- CPU FP64 restatement of nelson's Gauss-Newton /
 * Levenberg-Marquardt Schur-complement step for bundle adjustment.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and
 * by nothing else.  The product (nelson_amd/) never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the nelson repository root).  Pinning of each function (what checks it
 * independently of this file) is listed at the top of nsn_oracle.c.
 */
#ifndef NSN_ORACLE_H
#define NSN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generic block Hessian (nelson/DoubleSectionHessian.h:37-141) -------
 * U : dense nu x nu (nu = BU*NU), row-major; only the upper triangle (j >= i)
 *     is written and read, like the reference's selfadjointView<Upper>.
 * V : NV diagonal blocks of BV x BV, row-major each (full block stored).
 * W : nW blocks of BU x BV, row-major each, block (Wi[k], Wj[k]); blocks are
 *     kept in the reference's UID order (row-major: by U row, then V column,
 *     nelson/DoubleSection.hpp:157-191).
 * bU: nu, bV: BV*NV, chi2 = sum of edge costs (no 1/2).                     */
typedef struct {
  int BU, BV, NU, NV, nW;
  double *U, *V, *W, *bU, *bV;
  int *Wi, *Wj;
  double chi2;
  /* off-diagonal V blocks of V-V edges (nelson/DoubleSection.h:498-510):
   * Vo holds BV x BV row-major blocks V_(Voi[k], Voj[k]), Voi[k] < Voj[k];
   * nVo = 0 (V block diagonal) unless or_hessian_alloc_voff was called */
  int nVo;
  double* Vo;
  int *Voi, *Voj;
} or_hessian;

int  or_hessian_alloc(or_hessian* h, int BU, int BV, int NU, int NV, int nW);
void or_hessian_free(or_hessian* h);
int  or_hessian_alloc_voff(or_hessian* h, int nVo);          /* after or_hessian_alloc */
void or_hessian_clear(or_hessian* h);                     /* DoubleSectionHessian.hpp:51-58 */
double or_hessian_max_abs_b(const or_hessian* h);         /* DoubleSectionHessian.hpp:60-63 */

/* ---- dense pivoted LDL^T (Eigen::LDLT<M,Upper>, used by
 *      nelson/SolverCholeskyDense.h:14,47) --------------------------------
 * a: n x n column-major holding the lower triangle (== upper of the row-major
 * symmetric matrix).  On return a holds L (unit lower, strictly below diag)
 * and D on the diagonal; perm holds the transposition sequence.
 * Returns 1 on success (Eigen ComputationInfo::Success), 0 otherwise.       */
int  or_ldlt_factor(double* a, int n, int* transp);
void or_ldlt_solve(const double* a, int n, const int* transp, double* x);

/* ---- Schur step (nelson/SolverCholeskySchur.hpp:91-157) ----------------
 * Writes the increment (xU in internal order, xV); returns 1 on success, 0
 * on factorisation failure (increment zeroed, SolverCholeskySchur.hpp:151-154).
 * work/phase outputs: if S_out != NULL it receives the damped Schur matrix
 * (nu x nu row-major upper), if bS_out != NULL the reduced rhs.            */
int or_schur_assemble(const or_hessian* h, double rel, double abs_, double* S_out, double* bS_out);
int or_schur_step(const or_hessian* h, double rel, double abs_, double* xU, double* xV,
                  double* S_out, double* bS_out);

/* maxAbsHDiag (nelson/SolverCholeskySchur.hpp:75-78) */
double or_max_abs_h_diag(const or_hessian* h);

/* computeRhoChi2Change (nelson/DoubleSection.h:300-311) */
double or_rho(const or_hessian* h, double mu, const double* xU, const double* xV,
              double old_chi2, double new_chi2);

/* ---- the L4 kernels, as separate functions (property-test targets) ------ */
/* MatrixDiagInv::compute (nelson/MatrixDiagInv.hpp:25-140): Vinv_p = inv(sym(V_p) damped) */
int  or_diag_inv(int NV, int BV, const double* V, double rel, double abs_, double* Vinv);
/* MatrixDiagInv::rightMultVector (MatrixDiagInv.hpp:143-213): res_p = Vinv_p v_p */
void or_diag_inv_mult(int NV, int BV, const double* Vinv, const double* v, double* res);
/* MatrixWVinvMultiplier::multiply (MatrixWVinvMultiplier.hpp:23-106): Y_k = W_k Vinv_{Wj[k]} */
void or_wvinv(int nW, int BU, int BV, const int* Wj, const double* W, const double* Vinv, double* Y);
/* MatrixWVinvMultiplier::rightMultVectorSub (:110-193): res_c -= sum_k Y_k v_{Wj[k]} */
void or_wvinv_mult_sub(int nW, int BU, int BV, const int* Wi, const int* Wj, const double* Y,
                       const double* v, double* res);
/* MatrixWWtMultiplier::multiply (MatrixWWtMultiplier.hpp:124-284):
 * S(upper, nu x nu row-major) = U(upper) - sum over pairs A_ik B_jk^T, i<=j   */
void or_wwt(int NU, int BU, int BV, int nW, const int* Wi, const int* Wj,
            const double* A, const double* B, const double* U, double* S);
/* MatrixWtXMultiplier::rightMultVectorSub (MatrixWtXMultiplier.hpp:49-129):
 * res_p -= sum_c W_cp^T x_c                                                  */
void or_wtx_sub(int nW, int BU, int BV, const int* Wi, const int* Wj, const double* W,
                const double* x, double* res);

/* ---- problem callbacks + LM / GN drivers ------------------------------- */
typedef struct {
  void* ctx;
  void (*update)(void* ctx, int hessian);          /* DoubleSection::update(bool) */
  void (*oplus)(void* ctx, const double* xU, const double* xV);
  void (*backup)(void* ctx);
  void (*rollback)(void* ctx);
  or_hessian* (*hessian)(void* ctx);
} or_ops;

typedef struct {                 /* nelson/LevenbergMarquardt.h:18-34, defaults LevenbergMarquardt.cpp:6-21 */
  int maxNumIt, maxNumSubIt, minNumIt;
  double epsBVector, epsIncSquare, epsChi2, tau, penaltyV, initV, maxScaleFactor, minScaleFactor;
} or_lm_settings;
void or_lm_default_settings(or_lm_settings* s);

typedef struct {                 /* nelson/GaussNewton.h:18-30, defaults GaussNewton.cpp:12-22 */
  int maxNumIt, minNumIt;
  double epsBVector, epsIncSquare, epsChi2, absLambda, relLambda;
} or_gn_settings;
void or_gn_default_settings(or_gn_settings* s);

/* termination codes: nelson/LevenbergMarquardt.h:36-44 / GaussNewton.h:34-41 */
enum { OR_SUCCESS_AT_START = 0, OR_SUCCESS_SMALL_INCREMENT, OR_SUCCESS_EPS_B,
       OR_SUCCESS_SMALL_RESIDUALS, OR_SOLVER_FAILURE, OR_MAX_ITERATION_REACHED,
       OR_MAX_SUBITERATION_REACHED };

/* per-sub-iteration log (LevenbergMarquardtStats, LevenbergMarquardt.h:57-85):
 * entry k = (it, subit, chi2); chi2 == -1 marks a solver failure.           */
typedef struct { int it, subit; double chi2, mu; } or_log_entry;

typedef struct {
  int reason, iterations, n_log;
  double final_chi2;
  or_log_entry* log; int log_cap;            /* caller-provided buffer */
  double* dx_log; int dx_log_cap; int dx_len; /* optional: increments of every solve */
} or_lm_result;

/* LevenbergMarquardt::solve (nelson/LevenbergMarquardt.hpp:35-146) */
int or_lm_solve(const or_lm_settings* s, or_ops* op, or_lm_result* res);
/* GaussNewton::solve (nelson/GaussNewton.hpp:31-127) */
int or_gn_solve(const or_gn_settings* s, or_ops* op, or_lm_result* res);

/* ---- bundle-adjustment section (builder-defined BA edge; contract of
 *      nelson/samples/BASchurPerfomances.cpp:163-232) -------------------- */
typedef struct or_ba or_ba;
/* cams0: 6 per camera (angle-axis omega, t), pts0: 3 per point.
 * weight may be NULL (all 1); cam_fixed / pt_fixed may be NULL (none fixed).
 * huber_delta <= 0 selects the plain squared loss.                          */
or_ba* or_ba_create(int C, int P, int E, const int* cam, const int* pt, const double* obs,
                    const double* weight, const unsigned char* cam_fixed,
                    const unsigned char* pt_fixed, double f, double cx, double cy,
                    double huber_delta, const double* cams0, const double* pts0);
void or_ba_destroy(or_ba* b);
or_ops or_ba_ops(or_ba* b);
or_hessian* or_ba_hessian(or_ba* b);
void or_ba_update(or_ba* b, int hessian);
void or_ba_oplus(or_ba* b, const double* xU, const double* xV);
/* state out: cams as 12 per camera (R row-major 3x3, t) and 3 per point */
void or_ba_get_state(const or_ba* b, double* cams_rt, double* pts);
void or_ba_set_state_rt(or_ba* b, const double* cams_rt, const double* pts);
/* per-edge residuals (2E) and Jacobians (Jc 12E row-major 2x6, Jp 6E 2x3) at the current state */
void or_ba_edge_jacobians(or_ba* b, double* r, double* Jc, double* Jp);
int  or_ba_num_var_cams(const or_ba* b);
int  or_ba_num_var_pts(const or_ba* b);

/* unary prior edges (the BA instance of nelson's EdgeUnary,
 * nelson/EdgeUnary.h:76-93: one residual on one node, accumulated into that
 * node's diagonal block and gradient segment).  Camera prior k on user
 * camera cam_ids[k]: r = w .* [log(R) - omega0; t - t0] (6 values of cam_p0
 * and cam_w per prior); point prior k: r = w .* (X - X0).  Jacobians w.r.t.
 * the section's perturbations: camera [[Jl^-1(log R), 0], [-[t]x, I]]
 * (left SE(3) update), point I.  A new call replaces the previous priors;
 * counts of 0 remove them. */
void or_ba_set_priors(or_ba* b, int ncp, const int* cam_ids, const double* cam_p0, const double* cam_w,
                      int npp, const int* pt_ids, const double* pt_p0, const double* pt_w);
/* per-camera intrinsics (f, k1, k2; 3 per camera) of the BAL camera model:
 * q = -(x, y) / z of Xc = R X + t, pixel = f (1 + k1 |q|^2 + k2 |q|^4) q
 * (the camera looks down -z; focal / cx / cy unused).  NULL restores the
 * shared pinhole. */
void or_ba_set_intrinsics(or_ba* b, const double* intr);
/* log of a rotation matrix (principal angle-axis) and the inverse left
 * Jacobian of SO(3) at omega (row-major 3x3) */
void or_so3_log(const double* R, double* w);
void or_so3_jl_inv(const double* w, double* J);

/* rotation helpers (own restatement of Rodrigues / SE(3) exp) */
void or_rodrigues(const double* w, double* R);
void or_se3_left_update(const double* delta, double* R, double* t);

/* host threads of the LDL^T row updates (OpenMP; results do not depend on
 * the count); n <= 0 only reads it.  Returns the count in use. */
int or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
