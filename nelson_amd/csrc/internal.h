// Internal declarations shared by the host code and the HIP kernels.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>

namespace nsn {

// Sets the calling thread's error message and returns `code`.
int fail(int code, const char* fmt, ...);
// Clears the calling thread's error message.
void clear_error();

// Host rotation helpers (angle-axis <-> rotation matrix, row-major).
void rot_exp(const double w[3], double R[9]);
void rot_log(const double R[9], double w[3]);

// Device-side loop state of the Levenberg-Marquardt / Gauss-Newton drivers.
// One instance per problem, in device memory; the host reads it back with a
// single small copy per sub-iteration.
struct DevCtl {
  // damping used by the next Schur step: diag <- diag * (1 + rel) + abs
  double rel, abs_;
  // LM state (nelson/LevenbergMarquardt.hpp:35-146)
  double mu, curV, oldChi2;
  double chi2;            // chi2 of the current linearisation / last cost evaluation
  double chi2_lin;        // chi2 at the last full linearisation
  double maxb;            // max |b| of the last full linearisation
  double maxdiag;         // max |diag U|, |diag V| of the last full linearisation
  double rho;
  double incsq;           // ||dx||^2 of the last increment
  double den;             // dx^T (mu dx - b)
  int iter, subiter;
  int done, reason;
  int need_lin;           // 1: the next sequence starts with a full linearisation
  int solve_ok;           // last Schur step factorised
  int fail_v, fail_s;     // V-block / S factorisation failures (counts)
  int wait_timeout;       // a bounded device wait expired
  int total_solves;
  int rollback_pending;   // rejected LM step: restore the backup
  // iterative refinement (mixed precision): max |r| and max |b| of the last
  // FP64 residual (bit patterns of non-negative doubles), whether the step
  // goes on, and the refinement steps run so far
  unsigned long long ir_maxr, ir_maxb;
  int ir_more, ir_count;
  int accepted;           // last LM step accepted
  int n_log, log_cap;
  int mode;               // 0 = primitives, 1 = LM loop, 2 = GN loop
  int started;
  // LM settings (copied in at solve start)
  int maxNumIt, maxNumSubIt, minNumIt;
  double epsBVector, epsIncSquare, epsChi2, tau, penaltyV, initV, maxScaleFactor, minScaleFactor;
  double gn_rel, gn_abs;
};

// A roctx range (rocprofiler-sdk-roctx) around a library call: a rocprofv3
// run with --marker-trace shows the C ABI calls beside the kernels they
// launch.  Host-side markers only; no device work.
struct TraceRange {
  explicit TraceRange(const char* name);
  ~TraceRange();
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

struct DevLog {
  int it, subit;
  double chi2, mu;
};

}  // namespace nsn
