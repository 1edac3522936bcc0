/*
 * nelson_amd.h -- C ABI of the MI355X (gfx950) Schur-complement LM step for
 * bundle adjustment.  Plain pointers and sizes only; every call returns an
 * int status (NSN_OK = 0) and nsn_last_error() describes the last failure of
 * the calling thread.
 *
 * The reference exposes a C++ template API and no C ABI.  Each entry point
 * below names the reference member it replaces (paths relative to the nelson
 * repository root).  One handle owns all device buffers of one problem; a
 * handle is used by one host thread at a time (the reference's section and
 * solver are not re-entrant either).
 *
 * Increment vectors (dxu, dxv) follow the reference's
 * DoubleSectionHessianVectors: 6 values per VARIABLE camera (angle, then
 * translation of the left SE(3) perturbation) in camera order with the fixed
 * cameras skipped, then 3 per VARIABLE point.
 */
#ifndef NELSON_AMD_H
#define NELSON_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NSN_OK 0
#define NSN_NOT_PD 1      /* factorisation failed; increment zeroed (SolverCholeskySchur.hpp:151-154) */
#define NSN_EINVAL (-1)   /* bad argument */
#define NSN_EHIP (-2)     /* HIP runtime error or no device */
#define NSN_ENOMEM (-3)   /* allocation failed */
#define NSN_ESTATE (-4)   /* call out of order (e.g. schur step before linearize) */
#define NSN_EDEVICE (-5)  /* a device-side check failed (e.g. a bounded wait expired) */

/* precision of the solve */
#define NSN_FP64 0        /* everything in FP64 */
#define NSN_MIXED 1       /* FP32 Jacobians / U,V,W,S assembly / Cholesky; FP64 pivots, rho, refinement */

/* termination reasons: nelson/LevenbergMarquardt.h:36-44 (GaussNewton.h:34-41 is the prefix) */
#define NSN_SUCCESS_AT_START 0
#define NSN_SUCCESS_SMALL_INCREMENT 1
#define NSN_SUCCESS_EPS_B 2
#define NSN_SUCCESS_SMALL_RESIDUALS 3
#define NSN_SOLVER_FAILURE 4
#define NSN_MAX_ITERATION_REACHED 5
#define NSN_MAX_SUBITERATION_REACHED 6

typedef struct nsn_problem nsn_problem;

/* Problem description.  Replaces the user's DoubleSection subclass set-up:
 * parametersReady() + addEdge(NodeId cam, NodeId pt, EdgeBinaryUV*) per
 * observation + structureReady() (nelson/DoubleSection.hpp:67-80,83-246,338-365);
 * the fixed masks replace NodeType::Fixed (nelson/EdgeInterface.h:35-55).
 * All arrays are copied; the caller may free them after nsn_problem_create. */
typedef struct {
  int n_cameras, n_points, n_edges;
  const int* cam_idx;              /* [n_edges] camera of each observation            */
  const int* pt_idx;               /* [n_edges] point of each observation             */
  const double* obs;               /* [2*n_edges] pixel observations (u, v)           */
  const double* weight;            /* [n_edges] per-edge weight, or NULL (all 1)      */
  const unsigned char* cam_fixed;  /* [n_cameras] 1 = fixed, or NULL (none fixed)    */
  const unsigned char* pt_fixed;   /* [n_points] 1 = fixed, or NULL                   */
  double focal, cx, cy;            /* pinhole intrinsics shared by all cameras        */
  double huber_delta;              /* > 0: Huber loss with this threshold (pixels)    */
  int precision;                   /* NSN_FP64 or NSN_MIXED                           */
  int device;                      /* HIP device ordinal                              */
  /* [3*n_cameras] per-camera (f, k1, k2) of the BAL camera model, or NULL for
   * the shared pinhole above.  BAL model: the camera looks down -z,
   * p = -(R X + t)_xy / (R X + t)_z, pixel = f (1 + k1 |p|^2 + k2 |p|^4) p;
   * focal / cx / cy are then unused.  The intrinsics stay constant. */
  const double* intrinsics;
  /* order of the variable cameras inside the solver (the rows of S):
   * NSN_ORDER_AS_GIVEN, NSN_ORDER_RCM (reverse Cuthill-McKee on the camera
   * graph: a narrow envelope of S for the sparse-S path, the role of the
   * reference's AMD ordering, DoubleSection::permuteAMD_SchurU,
   * nelson/DoubleSection.hpp:30-63), or NSN_ORDER_ND (nested dissection of
   * the camera graph into independent parts and separators, aligned to the
   * 64-row tiles of S: the factorisation runs the independent parts'
   * diagonal chains at the same time).  Arrays indexed by variable camera at
   * this interface (dxu, the exported / imported Hessian) stay in camera
   * order either way. */
  int camera_order;
} nsn_problem_desc;

#define NSN_ORDER_AS_GIVEN 0
#define NSN_ORDER_RCM 1
#define NSN_ORDER_ND 2

typedef struct {
  int n_var_cameras, n_var_points, n_wblocks;  /* NU, NV, non-zero W blocks */
  int n_s;                                     /* 6 * NU, order of the reduced camera system */
  int n_s_blocks;                              /* non-zero upper 6x6 blocks of S */
  long long n_s_pairs;                         /* block products of the S assembly */
  int n_tiles;                                 /* stored 64x64 tiles of [S; bS^T]: its envelope (the
                                                  sparse S path; fill stays inside it) */
  int n_tiles_dense;                           /* NT (NT + 1) / 2, the dense count */
  double factor_flops;                         /* floating-point operations of the tile Cholesky over
                                                  the envelope (tile updates 2*64^3 each, triangular
                                                  solves 64^3, tile factorisations 64^3/3) */
  int n_chains;                                /* diagonal chains the factorisation runs at the same
                                                  time (independent parts of the camera order) */
  int chain_steps;                             /* its critical path in 64-column diagonal steps (the
                                                  number of tile rows for a single chain) */
  int lookahead;                               /* columns by which the factorisation's worker queue
                                                  starts the chains' pre-updates early (the largest
                                                  up to 12 its launch is checked to complete with) */
  int chunk_tiles;                             /* scratch tiles of the sums split into chunks (sums
                                                  of more than 128 tile products) */
} nsn_problem_info;

/* LevenbergMarquardtSettings (nelson/LevenbergMarquardt.h:18-34) */
typedef struct {
  int maxNumIt, maxNumSubIt, minNumIt;
  double epsBVector, epsIncSquare, epsChi2, tau, penaltyV, initV, maxScaleFactor, minScaleFactor;
} nsn_lm_settings;

/* GaussNewtonSettings (nelson/GaussNewton.h:18-30) */
typedef struct {
  int maxNumIt, minNumIt;
  double epsBVector, epsIncSquare, epsChi2, absLambda, relLambda;
} nsn_gn_settings;

/* one LevenbergMarquardtStats::InStats record (nelson/LevenbergMarquardt.h:58-61):
 * chi2 == -1 marks a solver failure; mu is the damping used for that solve. */
typedef struct { int it, subit; double chi2, mu; } nsn_log_entry;

typedef struct {
  int reason;          /* NSN_SUCCESS_* / NSN_MAX_* */
  int iterations;      /* outer iterations completed (the reference's iter) */
  int sub_iterations;  /* total Schur solves */
  int n_log;           /* log entries written (including the initial one) */
  double final_chi2;   /* hessian().chi2() at return */
  double elapsed_ms;   /* wall-clock of the solve */
} nsn_lm_result;

/* ---- library ------------------------------------------------------------ */
const char* nsn_last_error(void);
int nsn_version(int* major, int* minor);
int nsn_device_count(int* n);

/* defaults: nelson/LevenbergMarquardt.cpp:6-21, nelson/GaussNewton.cpp:12-22 */
void nsn_lm_default_settings(nsn_lm_settings* s);
void nsn_gn_default_settings(nsn_gn_settings* s);

/* ---- synthetic BA generator (host only) -----------------------------------
 * Seeded SplitMix64 + Box-Muller; the samples' BA_Problem constructor and
 * addNoise (nelson/samples/BASchurPerfomances.cpp:40-100) play this role in
 * the reference.  layout 0: cameras on a circle looking at a 4 m cube of
 * points; layout 1: cameras on a ring, each point observed from a contiguous
 * arc of `arc` cameras; layout 2: the same with arcs that do not wrap round
 * the ring (cameras along an open path: a banded S).  Camera 0 and point 0
 * are left unperturbed (the gauge
 * of the sample, BASchurPerfomances.cpp:245-246).  Writes at most edge_cap
 * observations; *n_edges receives the count (NSN_EINVAL if it would exceed). */
typedef struct {
  int n_cameras, n_points;
  int k_min, k_max;            /* observations per point, uniform in [k_min, k_max] */
  int layout, arc;
  double obs_sigma;            /* pixel noise */
  double outlier_frac, outlier_sigma;
  double rot_sigma, trans_sigma, point_sigma;   /* initial-state perturbation */
  double focal, cx, cy;
  unsigned long long seed;
} nsn_gen_params;

int nsn_generate_ba(const nsn_gen_params* p, int edge_cap,
                    double* cams_gt, double* pts_gt, double* cams0, double* pts0,
                    int* cam_idx, int* pt_idx, double* obs, int* n_edges);

/* Synthetic BA on a given visibility graph (camera-point pairs), with the
 * geometry of the reference's BA sample (BASchurPerfomances.cpp:40-92, which
 * builds its problem on the ladybug graph at :236-248): points uniform in
 * [-point_radius, point_radius]^3, cameras at cam_radius from the origin
 * looking at it with a random roll; obs = projection + N(0, obs_sigma);
 * initial state = ground truth perturbed except camera 0 and point 0.
 * cam_radius must exceed sqrt(3) point_radius.  The sample's values:
 * focal 100, principal point (50, 50), point_radius 1, cam_radius 2. */
typedef struct {
  double focal, cx, cy;
  double point_radius, cam_radius;
  double obs_sigma;
  double rot_sigma, trans_sigma, point_sigma;
  unsigned long long seed;
} nsn_graph_gen_params;

int nsn_generate_on_graph(const nsn_graph_gen_params* p, int n_cameras, int n_points, int n_edges,
                          const int* cam_idx, const int* pt_idx, double* cams_gt, double* pts_gt,
                          double* cams0, double* pts0, double* obs);

/* Problem::load (nelson/samples/Problems.cpp:8-30): reads a visibility-graph
 * text file "nCameras nPoints nEdges" + nEdges "camera point" pairs.  With
 * cam_idx and pt_idx NULL it reads the header only (sizes); otherwise it
 * fills up to capacity pairs.  Unlike the reference it rejects malformed or
 * short files and out-of-range pairs (NSN_EINVAL, message names the entry). */
int nsn_visibility_load(const char* path, int capacity, int* n_cameras, int* n_points, int* n_edges,
                        int* cam_idx, int* pt_idx);

/* A Bundle Adjustment in the Large (BAL) problem file, the data format of the
 * ladybug problem the reference's visibility graph comes from
 * (nelson/samples/Problems.cpp:38): header "nCameras nPoints nObservations",
 * the observations "camera point x y", 9 numbers per camera (angle-axis,
 * translation, f, k1, k2), 3 per point.  With every array NULL it reads the
 * header only; otherwise it fills obs (2 per observation), cams (6 per camera,
 * the nsn_set_state layout), intrinsics (3 per camera, nsn_problem_desc's
 * BAL model) and pts (3 per point).  Rejects short, malformed or
 * out-of-range input (NSN_EINVAL, message names the entry). */
int nsn_bal_load(const char* path, int capacity, int* n_cameras, int* n_points, int* n_obs,
                 int* cam_idx, int* pt_idx, double* obs, double* cams, double* intrinsics, double* pts);

/* ---- problem -------------------------------------------------------------- */
/* parametersReady + addEdge* + structureReady (DoubleSection.hpp:67-246,338-365) */
int nsn_problem_create(const nsn_problem_desc* d, nsn_problem** out);
int nsn_problem_destroy(nsn_problem* h);
int nsn_problem_info_get(nsn_problem* h, nsn_problem_info* info);
/* The structure stage on the host alone (no device is touched): the figures
 * nsn_problem_info_get would report for a problem created from d (camera
 * order, tile envelope, chains, operations), and, when cam_order is not
 * NULL, the order of the variable cameras inside the solver
 * (cam_order[k] = the camera placed k-th, [number of variable cameras]).
 * Lets a caller compare camera orders before allocating anything. */
int nsn_problem_plan(const nsn_problem_desc* d, nsn_problem_info* info, int* cam_order);

/* state: cameras 6 per camera (angle-axis, translation), points 3 per point;
 * replaces the user's parameterU/parameterV (DoubleSection.h:226-229) */
int nsn_set_state(nsn_problem* h, const double* cams, const double* pts);
int nsn_get_state(nsn_problem* h, double* cams, double* pts);
/* cameras as 12 per camera: R (row-major 3x3) then t */
int nsn_get_state_rt(nsn_problem* h, double* cams_rt, double* pts);
int nsn_set_state_rt(nsn_problem* h, const double* cams_rt, const double* pts);
/* backupSolution / rollbackSolution (driven by LevenbergMarquardt.hpp:51,118,136) */
int nsn_backup(nsn_problem* h);
int nsn_rollback(nsn_problem* h);

/* DoubleSection::update(bool) (DoubleSection.hpp:452-467); chi2 may be NULL */
int nsn_linearize(nsn_problem* h, int with_hessian, double* chi2);
/* DoubleSectionHessian::maxAbsValBVect (DoubleSectionHessian.hpp:60-63) */
int nsn_max_abs_b(nsn_problem* h, double* out);

/* Hessian export in the reference's block order (U, V by diagonal block;
 * W row-major by camera, then point: DoubleSection.hpp:157-191).
 * U: 36 per variable camera (full 6x6 row-major), V: 9 per variable point,
 * W: 18 per block (6x3 row-major), Wi/Wj: block coordinates (variable indices).
 * Any output may be NULL. */
int nsn_export_hessian(nsn_problem* h, double* U, double* V, double* W, int* Wi, int* Wj,
                       double* bU, double* bV, double* chi2);
/* host-assembled Hessian in the same layout (generic edges evaluated by the
 * caller); the structure (Wi, Wj) must equal the problem's. */
int nsn_import_hessian(nsn_problem* h, const double* U, const double* V, const double* W,
                       const double* bU, const double* bV, double chi2);

/* A problem given by its block structure instead of BA edges, for sections
 * whose edges the caller evaluates on the host (the reference's generic
 * DoubleSection with user edges, nelson/DoubleSection.h:181-545): n_u
 * parameters of the first section (6x6 diagonal blocks of U), n_v of the
 * second (3x3 diagonal blocks of V, which must be block diagonal), the
 * non-zero 6x3 W blocks (u, v) and the non-zero off-diagonal 6x6 U blocks
 * (i, j), i < j, of U-U edges.  A section with V-V edges (V not block
 * diagonal) is given as one first section of n_v + n_u parameters, the V
 * parameters first, with n_v = 0: its V-V blocks, W blocks (transposed) and
 * U-U blocks are all off-diagonal blocks; the factorisation then eliminates
 * V first and S = U - W V^-1 W^T is its trailing part (include/nelson_amd.hpp
 * does this for EdgeBinaryVV).  Every parameter is variable (the
 * caller leaves fixed ones out).  Smaller blocks are embedded by the caller
 * (include/nelson_amd.hpp pads them with an identity diagonal and zero
 * gradient, which leaves the solution of the real part unchanged).  The
 * Hessian is always imported (nsn_import_hessian_blocks); calls that need
 * edges or a BA state return NSN_ESTATE.  Mixed precision solves the
 * imported system without iterative refinement. */
typedef struct {
  int n_u, n_v;
  int n_w;
  const int* w_u;      /* [n_w] */
  const int* w_v;      /* [n_w] */
  int n_uoff;
  const int* uoff_i;   /* [n_uoff] */
  const int* uoff_j;   /* [n_uoff] */
  int precision;       /* NSN_FP64 or NSN_MIXED */
  int device;
} nsn_block_desc;
int nsn_problem_create_blocks(const nsn_block_desc* d, nsn_problem** out);
/* nsn_import_hessian plus the off-diagonal U blocks: Uoff holds 36 values
 * (6x6 row-major U_ij) per block, in row-major (i, j) order of the blocks
 * (the UID order of nelson/DoubleSection.hpp:157-191); NULL when the problem
 * has none.  The W blocks are in row-major (u, v) order as well. */
int nsn_import_hessian_blocks(nsn_problem* h, const double* U, const double* Uoff, const double* V,
                              const double* W, const double* bU, const double* bV, double chi2);

/* Test hook, no counterpart in the reference: builds the structure of d
 * (structureReady: edge orders, W blocks, the S pattern and its (W, W)
 * pairs, d->camera_order honoured) twice, on the host and on the GPU (the
 * path nsn_problem_create takes), and sets *n_diff to the number of fields
 * in which they differ (0: equal; nsn_last_error names the first).
 * host_ms / device_ms (may be NULL) receive the two build times. */
int nsn_structure_check(const nsn_problem_desc* d, long long* n_diff, double* host_ms, double* device_ms);

/* ---- SolverCholeskySchur (nelson/SolverCholeskySchur.h:53-117) ---------- */
int nsn_schur_init(nsn_problem* h);                                  /* init (.hpp:44-62) */
int nsn_max_abs_h_diag(nsn_problem* h, double* out);                 /* maxAbsHDiag (.hpp:75-78) */
/* computeIncrement(H, b, rel, abs) (.hpp:91-157): NSN_OK or NSN_NOT_PD.
 * dxu / dxv (may be NULL) receive incrementVector(); *sq_norm (may be NULL)
 * incrementVectorSquaredNorm(). */
int nsn_schur_step(nsn_problem* h, double rel, double abs_, double* dxu, double* dxv, double* sq_norm);
/* the user's oplus with the last increment (BASchurPerfomances.cpp:163-170) */
int nsn_oplus(nsn_problem* h);
/* computeRhoChi2Change(mu, inc, oldChi2) with the current chi2 (DoubleSection.h:300-311) */
int nsn_rho(nsn_problem* h, double mu, double old_chi2, double* rho);

/* ---- drivers -------------------------------------------------------------- */
/* LevenbergMarquardt<SolverCholeskySchur>::solve (LevenbergMarquardt.hpp:35-146),
 * with the loop state (mu, v, rho, accept/reject, stats) kept on the device.
 * log (may be NULL) receives up to log_cap entries. */
int nsn_lm_solve(nsn_problem* h, const nsn_lm_settings* s, nsn_lm_result* res,
                 nsn_log_entry* log, int log_cap);
/* GaussNewton<SolverCholeskySchur>::solve (GaussNewton.hpp:31-127) */
int nsn_gn_solve(nsn_problem* h, const nsn_gn_settings* s, nsn_lm_result* res,
                 nsn_log_entry* log, int log_cap);

/* per-phase device time of the last Schur step, milliseconds, in the order of
 * SolverCholeskySchurIterationTimeStat (SolverCholeskySchur.h:25-38):
 * [0] V^-1 and W V^-1, [1] bS + S assembly, [2] factorisation,
 * [3] triangular solves, [4] back-substitution, and with n >= 6
 * [5] iterative refinement (mixed precision).  n >= 5. */
int nsn_phase_times(nsn_problem* h, double* ms, int n);

/* ---- mixed precision (no reference counterpart) ---------------------------
 * Iterative refinement of every NSN_MIXED increment: each step evaluates the
 * FP64 residual r = -b - (H + D) dx of the damped normal equations from the
 * edges and adds the FP32 Schur solve of it.  The gain-ratio denominator
 * uses the FP64 gradient.  Ignored for NSN_FP64.
 * nsn_set_refinement_tol: at most max_steps steps (0..8); a step stops after
 * its residual when max |r| <= tol * max |b| (tol = 0: every step runs), the
 * test made on the device.  Default: 2 steps, tol 1e-10 (the work of
 * two fixed steps, less the second solve when the first reaches the
 * tolerance).
 * nsn_set_refinement: exactly `steps` steps (tol 0).
 * nsn_refinement_count: refinement solves run since the problem was created. */
int nsn_set_refinement(nsn_problem* h, int steps);
int nsn_set_refinement_tol(nsn_problem* h, int max_steps, double tol);
int nsn_refinement_count(nsn_problem* h, int* steps);

/* replaces the observations [2*n_edges] and / or the per-edge weights
 * [n_edges] (either may be NULL) of an existing problem, in the user's edge
 * order; the next nsn_linearize / solve uses them.  The reference keeps the
 * measurements inside its user edges; this is the equivalent of editing them
 * between solves. */
int nsn_set_observations(nsn_problem* h, const double* obs, const double* weight);

/* ---- implicit-function-theorem adjoint (no reference counterpart) --------
 * At a converged state x* (gradient b(x*) = 0) and for a loss L(x*) with
 * gradient g given per camera [6*n_cameras] (left-perturbation coordinates
 * (omega, upsilon) of the BA sample's oplus, BASchurPerfomances.cpp:163-170)
 * and per point [3*n_points], in user order (entries of fixed nodes are
 * ignored; either array may be NULL for zero), computes
 *   lambda = (db/dx)^-1 g,  dL/d(obs) = -lambda^T db/d(obs),
 *   dL/d(weight) = -lambda^T db/d(weight)
 * into dL_dobs [2*n_edges] and dL_dw [n_edges] (either may be NULL), in the
 * user's edge order.  db/dx is the Gauss-Newton Hessian, factorised once
 * (undamped) with the Schur solver; hessian_steps > 0 refines lambda against
 * the exact Hessian, whose products are central differences of the FP64
 * gradient along lambda, each step reusing the same factorisation.
 * Cost: one linearisation, one Schur factorisation and 1 + hessian_steps
 * Schur solves.  NSN_NOT_PD if the Hessian is not positive definite. */
int nsn_adjoint(nsn_problem* h, const double* g_cams, const double* g_pts, int hessian_steps,
                double* dL_dobs, double* dL_dw);

/* The same adjoint without allocating device memory.  No counterpart in the
 * reference; for callers that cannot allocate at the point of the call
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc after set-up).  The work vectors live in a static device array
 * of 32 MiB (4 Mi doubles) that each device holds for the life of the
 * process; the problem must fit: 7 x 6 NU + 7 x 3 NV + 3 E doubles plus
 * rounding of each vector to 256 bytes (NSN_EINVAL otherwise, before any
 * work).  Calls on one device share the array: calls that overlap on one
 * device race.  The results equal nsn_adjoint's. */
int nsn_adjoint_noalloc(nsn_problem* h, const double* g_cams, const double* g_pts, int hessian_steps,
                        double* dL_dobs, double* dL_dw);

/* ---- batched independent windows (no reference counterpart) -------------
 * Many independent BA problems ("windows", uneven sizes, each with at most 42
 * variable cameras) solved by LevenbergMarquardt<SolverCholeskySchur>
 * (LevenbergMarquardt.hpp:35-146) in one persistent launch: each workgroup
 * runs whole windows.  All windows share precision and device. */
typedef struct nsn_batch nsn_batch;
int nsn_batch_create(const nsn_problem_desc* windows, int n_windows, nsn_batch** out);
int nsn_batch_destroy(nsn_batch* b);
/* cams / pts: concatenation over windows of 6*C_w (angle-axis, t) and 3*P_w values */
int nsn_batch_set_state(nsn_batch* b, const double* cams, const double* pts);
int nsn_batch_get_state(nsn_batch* b, double* cams, double* pts);
/* LM with the settings s on every window from the current state; chi2_out
 * [n_windows] (may be NULL) final costs, *elapsed_ms (may be NULL) device
 * time of the launch.  The first call runs nsn_batch_tune unless a launch
 * shape was set. */
int nsn_batch_lm_solve(nsn_batch* b, const nsn_lm_settings* s, double* chi2_out, double* elapsed_ms);
/* per-window termination reason (NSN_* reasons), outer iterations and Schur
 * solves of the last nsn_batch_lm_solve; any array may be NULL */
int nsn_batch_results(nsn_batch* b, int* reasons, int* iterations, int* sub_iterations);
/* autotuner: times every launch shape -- factorisation variant (8 or 16:
 * column panels of that width; 64: 64x64 tiles with MFMA products) x 256 /
 * 512 / 1024 persistent workgroups -- for up to 2 LM iterations on this batch
 * (state restored after each) and keeps the fastest */
int nsn_batch_tune(nsn_batch* b, const nsn_lm_settings* s, int* variant, int* grid);
int nsn_batch_set_config(nsn_batch* b, int variant, int grid);
/* the launch shape in use and the tuner's timings (ms, in candidate order) */
int nsn_batch_config(nsn_batch* b, int* variant, int* grid, double* tune_ms, int n_tune);
/* mixed-precision windows: at most max_steps (0..8) refinement steps per
 * increment, each adding the FP32 factor's solution for the window's FP64
 * residual of the damped normal equations, stopping once max |r| <= tol *
 * max |b| (as nsn_set_refinement_tol for one problem; tol 0: every step
 * runs).  Default 0 steps (the unrefined kernel): a step evaluates the
 * FP64 residual from every edge twice, which costs more than an FP64
 * batch's whole iteration saves.  FP64 batches accept and ignore it.  The
 * launch shape stays as tuned. */
int nsn_batch_set_refinement(nsn_batch* b, int max_steps, double tol);
/* refinement solves per window [n_windows] in the last nsn_batch_lm_solve */
int nsn_batch_refinement_counts(nsn_batch* b, int* steps);

/* ---- unary priors (nelson's EdgeUnary, EdgeUnary.h:76-93, for the BA
 * section) ----
 * Camera prior k on user camera cam_ids[k]: residual w .* [log(R) - omega0;
 * t - t0] with cam_p0 = (omega0 angle-axis, t0) and cam_w 6 weights per
 * prior; point prior k: w .* (X - X0), pt_p0 / pt_w 3 per prior.  Each adds
 * |r|^2 to the cost and J^T J, J^T r to its node's diagonal block and
 * gradient (variable nodes), J w.r.t. the section's perturbations (cameras:
 * left SE(3) update).  At most one prior per node; a new call replaces all
 * priors (counts 0 remove them); the next linearisation uses them.  Batched
 * windows take no priors. */
int nsn_set_priors(nsn_problem* h, int n_cam, const int* cam_ids, const double* cam_p0, const double* cam_w,
                   int n_pt, const int* pt_ids, const double* pt_p0, const double* pt_w);

/* ---- measured ceilings (no counterpart in the reference) ----
 * For stating roofline fractions against the device itself as well as
 * against its datasheet: the FP64 and FP32 rate of the 16x16x4 MFMA
 * instructions on register operands (TFLOP/s) and the bandwidth of a 1 GiB
 * streaming copy (GB/s, bytes read + written).  Any output may be NULL. */
int nsn_measure_peaks(int device, double* mfma_f64_tflops, double* mfma_f32_tflops, double* copy_gbps);

#ifdef __cplusplus
}
#endif
#endif
