// nelson_amd.hpp -- C++ host interface of the MI355X Schur LM step, over the
// C ABI of nelson_amd.h (header only; link libnelson_amd.so).
//
// It keeps the names, call order and result conventions of the reference's
// API for this path so that code written against nelson's BA sample reads
// the same:
//   BASection                  DoubleSection<...> of the BA sample with its
//                              reprojection edges (nelson/DoubleSection.h:181-545,
//                              nelson/samples/BASchurPerfomances.cpp:163-246)
//   SolverCholeskySchur        nelson/SolverCholeskySchur.h:53-117
//   LevenbergMarquardt<S>      nelson/LevenbergMarquardt.h:90-132
//   GaussNewton<S>             nelson/GaussNewton.h
// Lifecycle, as in the reference: set the parameters (parameterU / parameterV),
// parametersReady(), addEdge() per observation, structureReady(), then
// update(true) / solvers / drivers.  computeIncrement returns false (and a
// zero increment) when the factorisation fails, like the reference; misuse
// and device errors, which the reference only asserts, throw nelson_amd::Error.
#ifndef NELSON_AMD_HPP
#define NELSON_AMD_HPP

#include <array>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nelson_amd.h"

namespace nelson_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* call) {
  if (rc != NSN_OK && rc != NSN_NOT_PD)
    throw Error(rc, std::string(call) + ": " + nsn_last_error());
}

// nelson/LevenbergMarquardt.h:36-44 (GaussNewton uses the prefix up to
// MaxIterationReached)
enum class TerminationReason {
  SuccessAtStart = NSN_SUCCESS_AT_START,
  SuccessSmallIncrement = NSN_SUCCESS_SMALL_INCREMENT,
  SuccessEpsB = NSN_SUCCESS_EPS_B,
  SuccessSmallResiduals = NSN_SUCCESS_SMALL_RESIDUALS,
  SolverFailure = NSN_SOLVER_FAILURE,
  MaxIterationReached = NSN_MAX_ITERATION_REACHED,
  MaxSubIterationReached = NSN_MAX_SUBITERATION_REACHED,
};

inline std::string toString(TerminationReason r) {
  switch (r) {
    case TerminationReason::SuccessAtStart: return "SuccessAtStart";
    case TerminationReason::SuccessSmallIncrement: return "SuccessSmallIncrement";
    case TerminationReason::SuccessEpsB: return "SuccessEpsB";
    case TerminationReason::SuccessSmallResiduals: return "SuccessSmallResiduals";
    case TerminationReason::SolverFailure: return "SolverFailure";
    case TerminationReason::MaxIterationReached: return "MaxIterationReached";
    case TerminationReason::MaxSubIterationReached: return "MaxSubIterationReached";
  }
  return "?";
}

inline bool success(TerminationReason r) {
  return r == TerminationReason::SuccessAtStart || r == TerminationReason::SuccessSmallIncrement ||
         r == TerminationReason::SuccessEpsB || r == TerminationReason::SuccessSmallResiduals;
}

enum class Precision { FP64 = NSN_FP64, Mixed = NSN_MIXED };

// ---------------------------------------------------------------------------
// The BA section: cameras (angle-axis, translation; X_cam = R X + t) and
// points, pinhole intrinsics shared by all cameras.
// ---------------------------------------------------------------------------
class BASection {
 public:
  struct Camera {
    std::array<double, 3> w{};  // angle-axis
    std::array<double, 3> t{};  // translation
  };

  BASection(double focal, double cx, double cy) : f_(focal), cx_(cx), cy_(cy) {}
  ~BASection() {
    if (h_) nsn_problem_destroy(h_);
  }
  BASection(const BASection&) = delete;
  BASection& operator=(const BASection&) = delete;

  // sizes of the two parameter sets (numParametersU / numParametersV)
  void setNumParameters(int numU, int numV) {
    cams_.assign(numU, Camera());
    pts_.assign(numV, std::array<double, 3>{});
    camFixed_.assign(numU, 0);
    ptFixed_.assign(numV, 0);
  }
  int numParametersU() const { return (int)cams_.size(); }
  int numParametersV() const { return (int)pts_.size(); }
  Camera& parameterU(int i) { return cams_.at(i); }
  const Camera& parameterU(int i) const { return cams_.at(i); }
  std::array<double, 3>& parameterV(int i) { return pts_.at(i); }
  const std::array<double, 3>& parameterV(int i) const { return pts_.at(i); }
  // NodeType::Fixed of the reference's edges (nelson/EdgeInterface.h:35-55)
  void setFixedU(int i, bool fixed = true) { camFixed_.at(i) = fixed ? 1 : 0; }
  void setFixedV(int i, bool fixed = true) { ptFixed_.at(i) = fixed ? 1 : 0; }
  void setHuber(double delta) { huber_ = delta; }

  void parametersReady() { ready_ = true; }
  // one reprojection edge (EdgeBinaryUV) between camera i and point j
  void addEdge(int i, int j, double u, double v, double weight = 1.0) {
    if (!ready_) throw Error(NSN_ESTATE, "addEdge before parametersReady");
    if (h_) throw Error(NSN_ESTATE, "addEdge after structureReady");
    cam_.push_back(i);
    pt_.push_back(j);
    obs_.push_back(u);
    obs_.push_back(v);
    w_.push_back(weight);
  }
  // a unary prior edge (EdgeUnaryU / EdgeUnaryV, nelson/DoubleSection.h:361-366,
  // nelson/EdgeUnary.h:76-93) with a diagonal weight: on camera i
  // r = weight .* [log R - prior.w; t - prior.t], on point j r = weight .* (X - prior).
  // At most one per node (NSN_EINVAL at structureReady otherwise).
  void addEdgeUnaryU(int i, const Camera& prior, const std::array<double, 6>& weight) {
    if (!ready_) throw Error(NSN_ESTATE, "addEdgeUnaryU before parametersReady");
    if (h_) throw Error(NSN_ESTATE, "addEdgeUnaryU after structureReady");
    priCam_.push_back(i);
    for (int k = 0; k < 3; k++) priCamP_.push_back(prior.w[k]);
    for (int k = 0; k < 3; k++) priCamP_.push_back(prior.t[k]);
    priCamW_.insert(priCamW_.end(), weight.begin(), weight.end());
  }
  void addEdgeUnaryV(int j, const std::array<double, 3>& prior, const std::array<double, 3>& weight) {
    if (!ready_) throw Error(NSN_ESTATE, "addEdgeUnaryV before parametersReady");
    if (h_) throw Error(NSN_ESTATE, "addEdgeUnaryV after structureReady");
    priPt_.push_back(j);
    priPtP_.insert(priPtP_.end(), prior.begin(), prior.end());
    priPtW_.insert(priPtW_.end(), weight.begin(), weight.end());
  }
  void structureReady(Precision p = Precision::FP64, int device = 0) {
    if (!ready_) throw Error(NSN_ESTATE, "structureReady before parametersReady");
    nsn_problem_desc d{};
    d.n_cameras = numParametersU();
    d.n_points = numParametersV();
    d.n_edges = (int)cam_.size();
    d.cam_idx = cam_.data();
    d.pt_idx = pt_.data();
    d.obs = obs_.data();
    d.weight = w_.data();
    d.cam_fixed = camFixed_.data();
    d.pt_fixed = ptFixed_.data();
    d.focal = f_;
    d.cx = cx_;
    d.cy = cy_;
    d.huber_delta = huber_;
    d.precision = (int)p;
    d.device = device;
    check(nsn_problem_create(&d, &h_), "nsn_problem_create");
    if (!priCam_.empty() || !priPt_.empty())
      check(nsn_set_priors(h_, (int)priCam_.size(), priCam_.data(), priCamP_.data(), priCamW_.data(),
                           (int)priPt_.size(), priPt_.data(), priPtP_.data(), priPtW_.data()),
            "nsn_set_priors");
    upload();
  }

  // DoubleSection::update(bool): cost, and with `hessians` the blocks and b
  double update(bool hessians) {
    check(nsn_linearize(handle(), hessians ? 1 : 0, &chi2_), "nsn_linearize");
    return chi2_;
  }
  double chi2() const { return chi2_; }  // hessian().chi2()
  // the BA sample's oplus with the solver's last increment, then the state
  // is read back into parameterU / parameterV
  void oplus() {
    check(nsn_oplus(handle()), "nsn_oplus");
    download();
  }
  void backupSolution() { check(nsn_backup(handle()), "nsn_backup"); }
  void rollbackSolution() {
    check(nsn_rollback(handle()), "nsn_rollback");
    download();
  }
  // parameterU / parameterV <-> device state
  void upload() {
    std::vector<double> c(6 * cams_.size()), x(3 * pts_.size());
    for (size_t i = 0; i < cams_.size(); i++)
      for (int k = 0; k < 3; k++) {
        c[6 * i + k] = cams_[i].w[k];
        c[6 * i + 3 + k] = cams_[i].t[k];
      }
    for (size_t i = 0; i < pts_.size(); i++)
      for (int k = 0; k < 3; k++) x[3 * i + k] = pts_[i][k];
    check(nsn_set_state(handle(), c.data(), x.data()), "nsn_set_state");
  }
  void download() {
    std::vector<double> c(6 * cams_.size()), x(3 * pts_.size());
    check(nsn_get_state(handle(), c.data(), x.data()), "nsn_get_state");
    for (size_t i = 0; i < cams_.size(); i++)
      for (int k = 0; k < 3; k++) {
        cams_[i].w[k] = c[6 * i + k];
        cams_[i].t[k] = c[6 * i + 3 + k];
      }
    for (size_t i = 0; i < pts_.size(); i++)
      for (int k = 0; k < 3; k++) pts_[i][k] = x[3 * i + k];
  }
  nsn_problem* handle() const {
    if (!h_) throw Error(NSN_ESTATE, "BASection used before structureReady");
    return h_;
  }
  nsn_problem_info info() const {
    nsn_problem_info in{};
    check(nsn_problem_info_get(handle(), &in), "nsn_problem_info_get");
    return in;
  }

 private:
  double f_, cx_, cy_, huber_ = 0.0, chi2_ = 0.0;
  bool ready_ = false;
  std::vector<Camera> cams_;
  std::vector<std::array<double, 3>> pts_;
  std::vector<unsigned char> camFixed_, ptFixed_;
  std::vector<int> cam_, pt_;
  std::vector<double> obs_, w_;
  std::vector<int> priCam_, priPt_;
  std::vector<double> priCamP_, priCamW_, priPtP_, priPtW_;
  nsn_problem* h_ = nullptr;
};

// ---------------------------------------------------------------------------
// SolverCholeskySchur (nelson/SolverCholeskySchur.h:53-117)
// ---------------------------------------------------------------------------
class SolverCholeskySchur {
 public:
  struct TimingStats {  // SolverCholeskySchurIterationTimeStat phases, ms
    double vinv_wvinv = 0, s_assembly = 0, factorize = 0, solves = 0, backsub = 0, refine = 0;
  };
  void init(BASection& s) {
    s_ = &s;
    check(nsn_schur_init(s.handle()), "nsn_schur_init");
  }
  double maxAbsHDiag() const {
    double v = 0;
    check(nsn_max_abs_h_diag(section().handle(), &v), "nsn_max_abs_h_diag");
    return v;
  }
  bool computeIncrement(BASection& s, double relLambda, double absLambda) {
    s_ = &s;
    const nsn_problem_info in = s.info();
    du_.assign(6 * (size_t)in.n_var_cameras, 0.0);
    dv_.assign(3 * (size_t)in.n_var_points, 0.0);
    const int rc = nsn_schur_step(s.handle(), relLambda, absLambda, du_.data(), dv_.data(), &sq_);
    check(rc, "nsn_schur_step");
    return rc == NSN_OK;
  }
  // incrementVector(): per variable camera (omega, upsilon), per variable point
  const std::vector<double>& incrementU() const { return du_; }
  const std::vector<double>& incrementV() const { return dv_; }
  double incrementVectorSquaredNorm() const { return sq_; }
  TimingStats timingStats() const {
    double ms[6] = {};
    check(nsn_phase_times(section().handle(), ms, 6), "nsn_phase_times");
    return TimingStats{ms[0], ms[1], ms[2], ms[3], ms[4], ms[5]};
  }

 private:
  const BASection& section() const {
    if (!s_) throw Error(NSN_ESTATE, "SolverCholeskySchur used before init");
    return *s_;
  }
  BASection* s_ = nullptr;
  std::vector<double> du_, dv_;
  double sq_ = 0.0;
};

// ---------------------------------------------------------------------------
// drivers (nelson/LevenbergMarquardt.h, nelson/GaussNewton.h).  The loop
// itself runs on the device (one graph launch per sub-iteration); settings,
// termination reasons and the per-sub-iteration chi2 log keep the
// reference's meaning.
// ---------------------------------------------------------------------------
struct LevenbergMarquardtSettings : nsn_lm_settings {
  LevenbergMarquardtSettings() { nsn_lm_default_settings(this); }
};
struct GaussNewtonSettings : nsn_gn_settings {
  GaussNewtonSettings() { nsn_gn_default_settings(this); }
};

struct Stats {  // LevenbergMarquardtStats / GaussNewtonStats
  std::vector<nsn_log_entry> log;
  std::string toString() const {
    std::ostringstream os;
    for (const auto& e : log) os << "[" << e.it << "," << e.subit << "] chi2 " << e.chi2 << " mu " << e.mu << "\n";
    return os.str();
  }
};

template <class Solver = SolverCholeskySchur>
class LevenbergMarquardt {
 public:
  LevenbergMarquardtSettings& settings() { return settings_; }
  const LevenbergMarquardtSettings& settings() const { return settings_; }
  TerminationReason solve(BASection& op) {
    const int cap = settings_.maxNumIt * settings_.maxNumSubIt + 8;
    stats_.log.assign(cap > 0 ? cap : 8, nsn_log_entry{});
    check(nsn_lm_solve(op.handle(), &settings_, &res_, stats_.log.data(), (int)stats_.log.size()), "nsn_lm_solve");
    stats_.log.resize(res_.n_log < (int)stats_.log.size() ? res_.n_log : stats_.log.size());
    op.download();
    return (TerminationReason)res_.reason;
  }
  const Stats& stats() const { return stats_; }
  int numIterations() const { return res_.iterations; }
  const nsn_lm_result& result() const { return res_; }

 private:
  LevenbergMarquardtSettings settings_;
  Stats stats_;
  nsn_lm_result res_{};
};

template <class Solver = SolverCholeskySchur>
class GaussNewton {
 public:
  GaussNewtonSettings& settings() { return settings_; }
  const GaussNewtonSettings& settings() const { return settings_; }
  TerminationReason solve(BASection& op) {
    stats_.log.assign(settings_.maxNumIt + 8, nsn_log_entry{});
    check(nsn_gn_solve(op.handle(), &settings_, &res_, stats_.log.data(), (int)stats_.log.size()), "nsn_gn_solve");
    stats_.log.resize(res_.n_log < (int)stats_.log.size() ? res_.n_log : stats_.log.size());
    op.download();
    return (TerminationReason)res_.reason;
  }
  const Stats& stats() const { return stats_; }
  int numIterations() const { return res_.iterations; }

 private:
  GaussNewtonSettings settings_;
  Stats stats_;
  nsn_lm_result res_{};
};

}  // namespace nelson_amd

#endif
