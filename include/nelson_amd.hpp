// nelson_amd.hpp -- C++ host interface of the MI355X Schur LM step, over the
// C ABI of nelson_amd.h (header only; link libnelson_amd.so).
//
// It keeps the names, call order and result conventions of the reference's
// API for this path:
//   DoubleSection<Derived, ParU, ParV, BU, BV>
//                              a user section with user edges, as
//                              nelson/DoubleSection.h:181-545: the user
//                              implements parameterU / parameterV (NodeId),
//                              oplus(inc), backupSolution, rollbackSolution;
//                              edges derive from EdgeUnaryU / EdgeUnaryV /
//                              EdgeBinaryUV / EdgeBinaryUU / EdgeBinaryVV /
//                              EdgeNaryU and implement update(bool) (calling
//                              setChi2) and the block updaters
//                              (nelson/EdgeBinary.h:121-158).  Edges are
//                              evaluated on the host; the Schur step runs on
//                              the GPU (BU <= 6, BV <= 3; with V-V edges the
//                              damped V is factorised together with S).
//   BASection                  the BA sample's section (nelson/samples/
//                              BASchurPerfomances.cpp:163-246) with its
//                              reprojection edges evaluated on the GPU
//   SolverCholeskySchur        init(H, b) / computeIncrement(H, b, rel, abs) /
//                              incrementVector / maxAbsHDiag / timingStats,
//                              nelson/SolverCholeskySchur.h:53-117
//   LevenbergMarquardt<S>      nelson/LevenbergMarquardt.h:90-132 (.hpp:35-146)
//   GaussNewton<S>             nelson/GaussNewton.h (.hpp:31-127)
// Lifecycle, as in the reference: set the parameters, parametersReady(),
// addEdge() per edge (the section takes ownership), optionally
// setUser2InternalIndexesU() / permuteAMD_SchurU(), structureReady(), then
// update(true) / solvers / drivers.  computeIncrement returns false (and a
// zero increment) when the factorisation fails, like the reference; misuse
// and device errors, which the reference only asserts, throw
// nelson_amd::Error.
#ifndef NELSON_AMD_HPP
#define NELSON_AMD_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
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
// node ids (nelson/EdgeInterface.h:35-55): variable parameters and fixed
// parameters are numbered separately; an int converts to a variable id
// ---------------------------------------------------------------------------
enum class NodeType { Variable, Fixed };

class NodeId {
 public:
  NodeId(int id = -1, NodeType t = NodeType::Variable) : id_(id), type_(t) {}
  int id() const { return id_; }
  NodeType type() const { return type_; }
  bool isVariable() const { return type_ == NodeType::Variable; }
  bool isFixed() const { return type_ == NodeType::Fixed; }
  static NodeId fixed(int i) { return NodeId(i, NodeType::Fixed); }

 private:
  int id_;
  NodeType type_;
};

// ---------------------------------------------------------------------------
// views of Hessian blocks and gradient segments handed to the edge updaters
// ---------------------------------------------------------------------------
class MatRef {  // rows x cols inside row-major storage (leading dimension ld), optionally transposed
 public:
  MatRef(double* p, int rows, int cols, int ld, bool trans = false)
      : p_(p), r_(rows), c_(cols), ld_(ld), t_(trans) {}
  int rows() const { return t_ ? c_ : r_; }
  int cols() const { return t_ ? r_ : c_; }
  double& operator()(int r, int c) { return t_ ? p_[c * ld_ + r] : p_[r * ld_ + c]; }
  double operator()(int r, int c) const { return t_ ? p_[c * ld_ + r] : p_[r * ld_ + c]; }
  // H += s * I (on the leading square part)
  MatRef& addIdentity(double s = 1.0) {
    for (int i = 0; i < std::min(rows(), cols()); i++) (*this)(i, i) += s;
    return *this;
  }

 private:
  double* p_;
  int r_, c_, ld_;
  bool t_;
};

class VecRef {
 public:
  VecRef(double* p, int n) : p_(p), n_(n) {}
  int size() const { return n_; }
  double& operator()(int i) { return p_[i]; }
  double& operator[](int i) { return p_[i]; }

 private:
  double* p_;
  int n_;
};

class ConstSeg {
 public:
  ConstSeg(const double* p, int n) : p_(p), n_(n) {}
  int size() const { return n_; }
  double operator()(int i) const { return p_[i]; }
  double operator[](int i) const { return p_[i]; }

 private:
  const double* p_;
  int n_;
};

// the increment (DoubleSectionHessianVectors): bU() holds BU values per
// variable U parameter in internal order, bV() BV values per V parameter
class HessianVectors {
 public:
  class Part {
   public:
    // block i (B_ values each, or the sizes of a section with per-parameter
    // sizes, off_[i] .. off_[i + 1])
    ConstSeg segment(int i) const {
      return off_.empty() ? ConstSeg(v_.data() + (size_t)B_ * i, B_) : ConstSeg(v_.data() + off_[i], off_[i + 1] - off_[i]);
    }
    int blockSize() const { return B_; }
    const std::vector<double>& values() const { return v_; }
    double operator[](size_t i) const { return v_[i]; }
    size_t size() const { return v_.size(); }

   private:
    friend class HessianVectors;
    friend class SolverCholeskySchur;
    int B_ = 0;
    std::vector<double> v_;
    std::vector<int> off_;
  };
  const Part& bU() const { return u_; }
  const Part& bV() const { return v_; }
  // a single section's increment (SingleSection::HessianVecType::segment)
  ConstSeg segment(int i) const { return u_.segment(i); }
  double squaredNorm() const {
    double s = 0, t = 0;
    for (double x : u_.v_) s += x * x;
    for (double x : v_.v_) t += x * x;
    return s + t;
  }
  const void* origin = nullptr;  // the section whose Hessian the increment solves

 private:
  friend class SolverCholeskySchur;
  Part u_, v_;
};

// What the solver reads of a section's Hessian: the device problem holding
// the blocks, and the block layout of the increment.
class SchurSystem {
 public:
  virtual ~SchurSystem() = default;
  virtual nsn_problem* problem() const = 0;
  virtual void prepare() const = 0;  // puts the current blocks on the device
  virtual double maxAbsHDiag() const = 0;
  virtual int blockU() const = 0;
  virtual int blockV() const = 0;
  virtual int numU() const = 0;  // variable parameters
  virtual int numV() const = 0;
  virtual const void* owner() const = 0;
  // true when the device system holds the V parameters as its first numV()
  // parameters and the U parameters after them (a V with off-diagonal
  // blocks, factorised together with S); false: separate U and V parts
  virtual bool vFirstInU() const { return false; }
  // the size of variable parameter u / q (at most blockU() / blockV())
  virtual int sizeU(int u) const {
    (void)u;
    return blockU();
  }
  virtual int sizeV(int q) const {
    (void)q;
    return blockV();
  }
  virtual bool variableSizes() const { return false; }
};

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

  // the section's Hessian as the solvers see it (DoubleSectionHessian's
  // chi2 / maxAbsValBVect / H / b): the blocks live on the device
  class Hessian : public SchurSystem {
   public:
    explicit Hessian(BASection* s) : s_(s) {}
    double chi2() const { return s_->chi2_; }
    double maxAbsValBVect() const {
      double v = 0;
      check(nsn_max_abs_b(s_->handle(), &v), "nsn_max_abs_b");
      return v;
    }
    const Hessian& H() const { return *this; }
    const Hessian& b() const { return *this; }
    nsn_problem* problem() const override { return s_->handle(); }
    void prepare() const override {}
    double maxAbsHDiag() const override {
      double v = 0;
      check(nsn_max_abs_h_diag(s_->handle(), &v), "nsn_max_abs_h_diag");
      return v;
    }
    int blockU() const override { return 6; }
    int blockV() const override { return 3; }
    int numU() const override { return s_->info().n_var_cameras; }
    int numV() const override { return s_->info().n_var_points; }
    const void* owner() const override { return s_; }

   private:
    BASection* s_;
  };

  BASection(double focal, double cx, double cy) : f_(focal), cx_(cx), cy_(cy), hess_(this) {}
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
  // the counterpart of permuteAMD_SchurU (nelson/DoubleSection.hpp:30-63) for
  // this section: the solver orders the cameras by reverse Cuthill-McKee on
  // the camera graph, so that S is stored and factorised over a narrow tile
  // envelope.  Unlike the reference, the increment, b and the Hessian blocks
  // stay in camera order at the interface; only the work changes.
  void permuteRCM_SchurU() {
    if (!ready_ || h_) throw Error(NSN_ESTATE, "permuteRCM_SchurU outside parametersReady .. structureReady");
    order_ = NSN_ORDER_RCM;
  }
  // nested dissection of the camera graph: independent parts (whose
  // diagonal chains the factorisation and the triangular solves run at the
  // same time) and their separators, aligned to the 64-row tiles of S; the
  // interface stays in camera order, as with permuteRCM_SchurU
  void permuteND_SchurU() {
    if (!ready_ || h_) throw Error(NSN_ESTATE, "permuteND_SchurU outside parametersReady .. structureReady");
    order_ = NSN_ORDER_ND;
  }
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
    d.camera_order = order_;
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
  const Hessian& hessian() const { return hess_; }
  // the BA sample's oplus with the solver's last increment, then the state
  // is read back into parameterU / parameterV
  void oplus() {
    check(nsn_oplus(handle()), "nsn_oplus");
    download();
  }
  // oplus(inc) of the reference's drivers: inc must be the increment that a
  // solver just computed on this section (it is applied from the device)
  void oplus(const HessianVectors& inc) {
    if (inc.origin != this) throw Error(NSN_ESTATE, "BASection::oplus: the increment comes from another section");
    oplus();
  }
  // DoubleSection::computeRhoChi2Change (nelson/DoubleSection.h:300-311)
  // with the solver's last increment and the current chi2
  double computeRhoChi2Change(double mu, const HessianVectors& inc, double oldChi2) const {
    if (inc.origin != this) throw Error(NSN_ESTATE, "computeRhoChi2Change: the increment comes from another section");
    double rho = 0;
    check(nsn_rho(handle(), mu, oldChi2, &rho), "nsn_rho");
    return rho;
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
  int order_ = NSN_ORDER_AS_GIVEN;
  std::vector<Camera> cams_;
  std::vector<std::array<double, 3>> pts_;
  std::vector<unsigned char> camFixed_, ptFixed_;
  std::vector<int> cam_, pt_;
  std::vector<double> obs_, w_;
  std::vector<int> priCam_, priPt_;
  std::vector<double> priCamP_, priCamW_, priPtP_, priPtW_;
  nsn_problem* h_ = nullptr;
  Hessian hess_;
};

// ---------------------------------------------------------------------------
// A user section with user edges (nelson/DoubleSection.h:181-545)
// ---------------------------------------------------------------------------
template <int BU, int BV>
class DoubleSectionHessian;

namespace detail {

// nelson/EdgeInterface.h: update(bool) computes the residual and calls
// setChi2; the section then runs the block updaters
class EdgeBase {
 public:
  virtual ~EdgeBase() = default;
  virtual void update(bool hessians) = 0;
  double chi2() const { return chi2_; }
  // section internals: the blocks this edge adds to
  virtual void applyUpdaters_() = 0;
  int uid_ = -1;          // W block / off-diagonal U block of a binary edge
  bool uidTrans_ = false; // the off-diagonal U block is stored as (j, i)
  std::vector<int> pairUid_;     // n-ary edges: the off-diagonal U block of each
  std::vector<char> pairTrans_;  // variable pair (a < b), in pair order

 protected:
  void setChi2(double v) { chi2_ = v; }

 private:
  double chi2_ = 0.0;
};

}  // namespace detail

namespace detail {
// the off-diagonal updaters may take the reference's transpose flag (the
// view handed over is already oriented, so it is always false) or not
template <class D, class = void>
struct H12TakesFlag : std::false_type {};
template <class D>
struct H12TakesFlag<D, std::void_t<decltype(std::declval<D&>().updateH12Block(std::declval<MatRef>(), false))>>
    : std::true_type {};
template <class D, class = void>
struct HijTakesFlag : std::false_type {};
template <class D>
struct HijTakesFlag<D, std::void_t<decltype(std::declval<D&>().updateHBlock(0, 0, std::declval<MatRef>(), false))>>
    : std::true_type {};
template <class D>
void updateH12(D* d, MatRef H) {
  if constexpr (H12TakesFlag<D>::value) d->updateH12Block(H, false);
  else d->updateH12Block(H);
}
}  // namespace detail

// an edge on one U parameter: the derived edge implements update(bool) and
// updateHBlock(MatRef H, VecRef b) (BU x BU, BU)
template <class Section, class Derived>
class EdgeUnaryU : public detail::EdgeBase {
 public:
  NodeId parId() const { return id_; }
  const auto& parameter() const { return section_->parameterU(id_); }
  Section& section() const { return *section_; }
  void bind_(Section* s, NodeId i) { section_ = s, id_ = i; }
  void applyUpdaters_() override {
    if (!id_.isVariable()) return;
    auto& H = section_->hessianStorage_();
    const int i = section_->internalU_(id_.id());
    static_cast<Derived*>(this)->updateHBlock(H.Ublock(i), H.bUseg(i));
  }

 private:
  Section* section_ = nullptr;
  NodeId id_;
};

// an edge on one V parameter: updateHBlock(MatRef H, VecRef b) (BV x BV, BV)
template <class Section, class Derived>
class EdgeUnaryV : public detail::EdgeBase {
 public:
  NodeId parId() const { return id_; }
  const auto& parameter() const { return section_->parameterV(id_); }
  Section& section() const { return *section_; }
  void bind_(Section* s, NodeId i) { section_ = s, id_ = i; }
  void applyUpdaters_() override {
    if (!id_.isVariable()) return;
    auto& H = section_->hessianStorage_();
    static_cast<Derived*>(this)->updateHBlock(H.Vblock(id_.id()), H.bVseg(id_.id()));
  }

 private:
  Section* section_ = nullptr;
  NodeId id_;
};

// an edge between a U and a V parameter: updateH11Block(H, b) (BU x BU, BU),
// updateH22Block(H, b) (BV x BV, BV), updateH12Block(H) (BU x BV, the W block)
template <class Section, class Derived>
class EdgeBinaryUV : public detail::EdgeBase {
 public:
  NodeId par_1_Id() const { return i_; }
  NodeId par_2_Id() const { return j_; }
  const auto& parameter_1() const { return section_->parameterU(i_); }
  const auto& parameter_2() const { return section_->parameterV(j_); }
  Section& section() const { return *section_; }
  void bind_(Section* s, NodeId i, NodeId j) { section_ = s, i_ = i, j_ = j; }
  void applyUpdaters_() override {
    auto& H = section_->hessianStorage_();
    Derived* d = static_cast<Derived*>(this);
    if (i_.isVariable()) {
      const int u = section_->internalU_(i_.id());
      d->updateH11Block(H.Ublock(u), H.bUseg(u));
    }
    if (j_.isVariable()) d->updateH22Block(H.Vblock(j_.id()), H.bVseg(j_.id()));
    if (i_.isVariable() && j_.isVariable()) detail::updateH12(d, H.Wblock(uid_));
  }

 private:
  Section* section_ = nullptr;
  NodeId i_, j_;
};

// an edge between two U parameters: updateH11Block / updateH22Block as
// above, updateH12Block(H) (BU x BU, the block of rows par_1, columns par_2)
template <class Section, class Derived>
class EdgeBinaryUU : public detail::EdgeBase {
 public:
  NodeId par_1_Id() const { return i_; }
  NodeId par_2_Id() const { return j_; }
  const auto& parameter_1() const { return section_->parameterU(i_); }
  const auto& parameter_2() const { return section_->parameterU(j_); }
  Section& section() const { return *section_; }
  void bind_(Section* s, NodeId i, NodeId j) { section_ = s, i_ = i, j_ = j; }
  void applyUpdaters_() override {
    auto& H = section_->hessianStorage_();
    Derived* d = static_cast<Derived*>(this);
    if (i_.isVariable()) {
      const int u = section_->internalU_(i_.id());
      d->updateH11Block(H.Ublock(u), H.bUseg(u));
    }
    if (j_.isVariable()) {
      const int u = section_->internalU_(j_.id());
      d->updateH22Block(H.Ublock(u), H.bUseg(u));
    }
    if (i_.isVariable() && j_.isVariable()) detail::updateH12(d, H.Uoffblock(uid_, uidTrans_));
  }

 private:
  Section* section_ = nullptr;
  NodeId i_, j_;
};

// an edge between two V parameters (nelson/DoubleSection.h:498-529,
// EdgeBinaryAdapterVV): updateH11Block / updateH22Block (BV x BV, BV),
// updateH12Block(H) (BV x BV, the block of rows par_1, columns par_2).  V is
// then not block diagonal: the solver factorises the damped V together with
// S (the V parameters eliminated first), which gives the Schur complement
// S = U - W V^-1 W^T as the trailing part of one factorisation.
template <class Section, class Derived>
class EdgeBinaryVV : public detail::EdgeBase {
 public:
  NodeId par_1_Id() const { return i_; }
  NodeId par_2_Id() const { return j_; }
  const auto& parameter_1() const { return section_->parameterV(i_); }
  const auto& parameter_2() const { return section_->parameterV(j_); }
  Section& section() const { return *section_; }
  void bind_(Section* s, NodeId i, NodeId j) { section_ = s, i_ = i, j_ = j; }
  void applyUpdaters_() override {
    auto& H = section_->hessianStorage_();
    Derived* d = static_cast<Derived*>(this);
    if (i_.isVariable()) d->updateH11Block(H.Vblock(i_.id()), H.bVseg(i_.id()));
    if (j_.isVariable()) d->updateH22Block(H.Vblock(j_.id()), H.bVseg(j_.id()));
    if (i_.isVariable() && j_.isVariable()) detail::updateH12(d, H.Voffblock(uid_, uidTrans_));
  }

 private:
  Section* section_ = nullptr;
  NodeId i_, j_;
};

// an edge on N parameters of the first section (nelson/EdgeNary.h:125-150;
// N = kDynamic: the count is set by addEdge): parameter(i) / parId(i) /
// numParams(); the derived edge implements update(bool) and
// updateHBlock(i, MatRef H, VecRef b) for each variable parameter i and
// updateHBlock(i, j, MatRef H[, bool transpose]) for each variable pair
// i < j (H: rows of parameter i, columns of parameter j)
constexpr int kDynamic = -1;

template <class Section, class Derived, int N>
class EdgeNaryU : public detail::EdgeBase {
 public:
  NodeId parId(int i) const { return ids_.at(i); }
  int numParams() const { return (int)ids_.size(); }
  const auto& parameter(int i) const { return section_->parameterU(ids_.at(i)); }
  Section& section() const { return *section_; }
  void bind_(Section* s, std::vector<NodeId> ids) { section_ = s, ids_ = std::move(ids); }
  void applyUpdaters_() override {
    auto& H = section_->hessianStorage_();
    Derived* d = static_cast<Derived*>(this);
    const int n = numParams();
    for (int a = 0; a < n; a++)
      if (ids_[a].isVariable()) {
        const int u = section_->internalU_(ids_[a].id());
        d->updateHBlock(a, H.Ublock(u), H.bUseg(u));
      }
    size_t k = 0;
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++)
        if (ids_[a].isVariable() && ids_[b].isVariable()) {
          MatRef Hab = H.Uoffblock(pairUid_[k], pairTrans_[k] != 0);
          if constexpr (detail::HijTakesFlag<Derived>::value) d->updateHBlock(a, b, Hab, false);
          else d->updateHBlock(a, b, Hab);
          k++;
        }
  }

 private:
  Section* section_ = nullptr;
  std::vector<NodeId> ids_;
};

// Host storage of the section's Hessian in internal order (U diagonal blocks
// BU x BU, off-diagonal U blocks and W blocks in row-major UID order, V
// diagonal blocks, bU, bV, chi2) and its device copy: prepare() embeds the
// blocks in the GPU's 6 / 3 layout (identity on the padded diagonal, zero
// gradient there, which leaves the increment of the real part unchanged and
// zero on the padding) and imports them.
template <int BU, int BV>
class DoubleSectionHessian : public SchurSystem {
  static_assert(BU >= 1 && BU <= 6 && BV >= 1 && BV <= 3, "block sizes up to 6 (U) and 3 (V)");

 public:
  struct Traits {  // DoubleSectionHessianTraits (nelson/DoubleSectionHessian.h)
    using Type = double;
    static constexpr int blockSizeU = BU, blockSizeV = BV;
  };
  ~DoubleSectionHessian() override {
    if (h_) nsn_problem_destroy(h_);
  }
  double chi2() const { return chi2_; }
  // DoubleSectionHessian::maxAbsValBVect (nelson/DoubleSectionHessian.hpp:60-63)
  double maxAbsValBVect() const {
    double m = 0;
    for (double x : bU_) m = std::max(m, std::fabs(x));
    for (double x : bV_) m = std::max(m, std::fabs(x));
    return m;
  }
  const DoubleSectionHessian& H() const { return *this; }
  const DoubleSectionHessian& b() const { return *this; }
  int numUoff() const { return (int)uoI_.size(); }
  int numW() const { return (int)wU_.size(); }
  // block access (internal order)
  MatRef Ublock(int i) { return MatRef(&U_[(size_t)BU * BU * i], su(i), su(i), BU); }
  MatRef Uoffblock(int uid, bool trans) {
    return MatRef(&Uo_[(size_t)BU * BU * uid], su(uoI_[uid]), su(uoJ_[uid]), BU, trans);
  }
  MatRef Vblock(int j) { return MatRef(&V_[(size_t)BV * BV * j], sv(j), sv(j), BV); }
  MatRef Wblock(int uid) { return MatRef(&W_[(size_t)BU * BV * uid], su(wU_[uid]), sv(wV_[uid]), BV); }
  MatRef Voffblock(int uid, bool trans) {
    return MatRef(&Vo_[(size_t)BV * BV * uid], sv(voI_[uid]), sv(voJ_[uid]), BV, trans);
  }
  int numVoff() const { return (int)voI_.size(); }
  VecRef bUseg(int i) { return VecRef(&bU_[(size_t)BU * i], su(i)); }
  VecRef bVseg(int j) { return VecRef(&bV_[(size_t)BV * j], sv(j)); }
  // per-parameter sizes (internal order; the block size when not given)
  int su(int u) const { return suI_.empty() ? BU : suI_[u]; }
  int sv(int q) const { return svI_.empty() ? BV : svI_[q]; }
  int sizeU(int u) const override { return su(u); }
  int sizeV(int q) const override { return sv(q); }
  bool variableSizes() const override { return !suI_.empty() || !svI_.empty(); }
  const std::vector<double>& bUvalues() const { return bU_; }
  const std::vector<double>& bVvalues() const { return bV_; }

  // SchurSystem
  nsn_problem* problem() const override {
    if (!h_) throw Error(NSN_ESTATE, "the section's Hessian is used before structureReady");
    return h_;
  }
  void prepare() const override {
    if (!dirty_) return;
    if (merged_) {
      prepareMerged();
      dirty_ = false;
      return;
    }
    const int NU = numU(), NV = numV(), nW = numW(), nO = numUoff();
    std::vector<double> U(36 * (size_t)NU, 0.0), Uo(36 * (size_t)nO, 0.0), V(9 * (size_t)NV, 0.0),
        W(18 * (size_t)nW, 0.0), bU(6 * (size_t)NU, 0.0), bV(3 * (size_t)NV, 0.0);
    for (int u = 0; u < NU; u++) {
      const int n = su(u);
      for (int a = 0; a < 6; a++)
        for (int c = 0; c < 6; c++)
          U[36 * (size_t)u + 6 * a + c] = (a < n && c < n) ? U_[(size_t)BU * BU * u + BU * a + c] : (a == c ? 1.0 : 0.0);
      for (int a = 0; a < n; a++) bU[6 * (size_t)u + a] = bU_[(size_t)BU * u + a];
    }
    for (int k = 0; k < nO; k++)
      for (int a = 0; a < BU; a++)
        for (int c = 0; c < BU; c++) Uo[36 * (size_t)k + 6 * a + c] = Uo_[(size_t)BU * BU * k + BU * a + c];
    for (int q = 0; q < NV; q++) {
      const int n = sv(q);
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++)
          V[9 * (size_t)q + 3 * a + c] = (a < n && c < n) ? V_[(size_t)BV * BV * q + BV * a + c] : (a == c ? 1.0 : 0.0);
      for (int a = 0; a < n; a++) bV[3 * (size_t)q + a] = bV_[(size_t)BV * q + a];
    }
    for (int k = 0; k < nW; k++)
      for (int a = 0; a < BU; a++)
        for (int c = 0; c < BV; c++) W[18 * (size_t)k + 3 * a + c] = W_[(size_t)BU * BV * k + BV * a + c];
    double none = 0.0;  // a section without V parameters or W blocks passes empty arrays
    auto ptr = [&](std::vector<double>& v) { return v.empty() ? &none : v.data(); };
    check(nsn_import_hessian_blocks(problem(), ptr(U), nO ? Uo.data() : nullptr, ptr(V), ptr(W), ptr(bU), ptr(bV),
                                    chi2_),
          "nsn_import_hessian_blocks");
    dirty_ = false;
  }
  // SolverCholeskySchur::maxAbsHDiag (nelson/SolverCholeskySchur.hpp:75-78):
  // the largest |diagonal| of V and of U
  double maxAbsHDiag() const override {
    double m = 0;
    for (int u = 0; u < numU(); u++)
      for (int a = 0; a < su(u); a++) m = std::max(m, std::fabs(U_[(size_t)BU * BU * u + (BU + 1) * a]));
    for (int q = 0; q < numV(); q++)
      for (int a = 0; a < sv(q); a++) m = std::max(m, std::fabs(V_[(size_t)BV * BV * q + (BV + 1) * a]));
    return m;
  }
  int blockU() const override { return BU; }
  int blockV() const override { return BV; }
  int numU() const override { return nU_; }
  int numV() const override { return nV_; }
  const void* owner() const override { return owner_; }
  bool vFirstInU() const override { return merged_; }

  // section internals
  void setup_(const void* owner, int nU, int nV, std::vector<std::pair<int, int>> wPairs,
              std::vector<std::pair<int, int>> uoPairs, std::vector<std::pair<int, int>> voPairs, Precision p,
              int device, std::vector<int> suI = {}, std::vector<int> svI = {}) {
    owner_ = owner;
    nU_ = nU;
    nV_ = nV;
    suI_ = std::move(suI);
    svI_ = std::move(svI);
    for (auto& w : wPairs) wU_.push_back(w.first), wV_.push_back(w.second);
    for (auto& o : uoPairs) uoI_.push_back(o.first), uoJ_.push_back(o.second);
    for (auto& o : voPairs) voI_.push_back(o.first), voJ_.push_back(o.second);
    U_.assign((size_t)BU * BU * nU, 0.0);
    Uo_.assign((size_t)BU * BU * uoI_.size(), 0.0);
    V_.assign((size_t)BV * BV * nV, 0.0);
    Vo_.assign((size_t)BV * BV * voI_.size(), 0.0);
    W_.assign((size_t)BU * BV * wU_.size(), 0.0);
    bU_.assign((size_t)BU * nU, 0.0);
    bV_.assign((size_t)BV * nV, 0.0);
    merged_ = !voI_.empty();
    nsn_block_desc d{};
    std::vector<int> mi, mj;
    if (merged_) {
      // V has off-diagonal blocks: one system of nV + nU parameters, the V
      // parameters first (device index q), then the U parameters (nV + u);
      // its off-diagonal blocks are the V-V blocks, the W blocks transposed
      // (rows V q, columns U u) and the U-U blocks, in row-major order
      for (size_t k = 0; k < voI_.size(); k++) mp_.push_back({voI_[k], voJ_[k], 0, (int)k});
      for (size_t k = 0; k < wU_.size(); k++) mp_.push_back({wV_[k], nV + wU_[k], 1, (int)k});
      for (size_t k = 0; k < uoI_.size(); k++) mp_.push_back({nV + uoI_[k], nV + uoJ_[k], 2, (int)k});
      std::sort(mp_.begin(), mp_.end(), [](const MergedBlock& x, const MergedBlock& y) {
        return x.i != y.i ? x.i < y.i : x.j < y.j;
      });
      for (auto& m : mp_) mi.push_back(m.i), mj.push_back(m.j);
      d.n_u = nV + nU;
      d.n_v = 0;
      d.n_w = 0;
      d.n_uoff = (int)mp_.size();
      d.uoff_i = mi.data();
      d.uoff_j = mj.data();
    } else {
      d.n_u = nU;
      d.n_v = nV;
      d.n_w = (int)wU_.size();
      d.w_u = wU_.data();
      d.w_v = wV_.data();
      d.n_uoff = (int)uoI_.size();
      d.uoff_i = uoI_.data();
      d.uoff_j = uoJ_.data();
    }
    d.precision = (int)p;
    d.device = device;
    check(nsn_problem_create_blocks(&d, &h_), "nsn_problem_create_blocks");
  }
  void clear_() {
    std::fill(U_.begin(), U_.end(), 0.0);
    std::fill(Uo_.begin(), Uo_.end(), 0.0);
    std::fill(V_.begin(), V_.end(), 0.0);
    std::fill(Vo_.begin(), Vo_.end(), 0.0);
    std::fill(W_.begin(), W_.end(), 0.0);
    std::fill(bU_.begin(), bU_.end(), 0.0);
    std::fill(bV_.begin(), bV_.end(), 0.0);
  }
  void setChi2_(double c) { chi2_ = c; }
  void touch_() { dirty_ = true; }

 private:
  struct MergedBlock {
    int i, j;  // device parameters, i < j
    int kind;  // 0: V-V block, 1: W block (transposed), 2: U-U block
    int k;     // its index in Vo_ / W_ / Uo_
  };
  // the merged system (V with off-diagonal blocks): every block embedded in
  // the device's 6 x 6 layout, V parameters first
  void prepareMerged() const {
    const int NU = numU(), NV = numV(), ND = NU + NV;
    std::vector<double> U(36 * (size_t)ND, 0.0), Uo(36 * mp_.size() + 1, 0.0), bU(6 * (size_t)ND, 0.0);
    for (int q = 0; q < NV; q++) {
      const int n = sv(q);
      for (int a = 0; a < 6; a++)
        for (int c = 0; c < 6; c++)
          U[36 * (size_t)q + 6 * a + c] = (a < n && c < n) ? V_[(size_t)BV * BV * q + BV * a + c] : (a == c ? 1.0 : 0.0);
      for (int a = 0; a < n; a++) bU[6 * (size_t)q + a] = bV_[(size_t)BV * q + a];
    }
    for (int u = 0; u < NU; u++) {
      const int n = su(u);
      double* dst = &U[36 * (size_t)(NV + u)];
      for (int a = 0; a < 6; a++)
        for (int c = 0; c < 6; c++)
          dst[6 * a + c] = (a < n && c < n) ? U_[(size_t)BU * BU * u + BU * a + c] : (a == c ? 1.0 : 0.0);
      for (int a = 0; a < n; a++) bU[6 * (size_t)(NV + u) + a] = bU_[(size_t)BU * u + a];
    }
    for (size_t m = 0; m < mp_.size(); m++) {
      double* dst = &Uo[36 * m];
      const int k = mp_[m].k;
      if (mp_[m].kind == 0) {
        for (int a = 0; a < BV; a++)
          for (int c = 0; c < BV; c++) dst[6 * a + c] = Vo_[(size_t)BV * BV * k + BV * a + c];
      } else if (mp_[m].kind == 1) {
        for (int a = 0; a < BV; a++)
          for (int c = 0; c < BU; c++) dst[6 * a + c] = W_[(size_t)BU * BV * k + BV * c + a];
      } else {
        for (int a = 0; a < BU; a++)
          for (int c = 0; c < BU; c++) dst[6 * a + c] = Uo_[(size_t)BU * BU * k + BU * a + c];
      }
    }
    double none = 0.0;
    check(nsn_import_hessian_blocks(problem(), U.data(), mp_.empty() ? nullptr : Uo.data(), &none, &none,
                                    bU.data(), &none, chi2_),
          "nsn_import_hessian_blocks");
  }

  const void* owner_ = nullptr;
  int nU_ = 0, nV_ = 0;
  std::vector<int> wU_, wV_, uoI_, uoJ_, voI_, voJ_;
  std::vector<double> U_, Uo_, V_, Vo_, W_, bU_, bV_;
  bool merged_ = false;
  std::vector<MergedBlock> mp_;
  std::vector<int> suI_, svI_;  // per-parameter sizes, internal order (empty: the block size)
  double chi2_ = 0.0;
  mutable bool dirty_ = true;
  nsn_problem* h_ = nullptr;
};

template <class Derived, class ParU, class ParV, int BU, int BV>
class DoubleSection {
 public:
  using ParameterUType = ParU;
  using ParameterVType = ParV;
  using HessianT = DoubleSectionHessian<BU, BV>;
  using Hessian = HessianT;  // P::Hessian::Traits names the solver's argument
  using HessianVectorsT = HessianVectors;
  static constexpr int blockSizeU = BU, blockSizeV = BV;
  template <class E>
  using EdgeUnaryU = nelson_amd::EdgeUnaryU<Derived, E>;
  template <class E>
  using EdgeUnaryV = nelson_amd::EdgeUnaryV<Derived, E>;
  template <class E>
  using EdgeBinaryUV = nelson_amd::EdgeBinaryUV<Derived, E>;
  template <class E>
  using EdgeBinaryUU = nelson_amd::EdgeBinaryUU<Derived, E>;
  template <class E>
  using EdgeBinaryVV = nelson_amd::EdgeBinaryVV<Derived, E>;
  template <class E, int N = kDynamic>
  using EdgeNaryU = nelson_amd::EdgeNaryU<Derived, E, N>;

  DoubleSection() = default;
  DoubleSection(int numU, int numV) : nU_(numU), nV_(numV) {}
  DoubleSection(const DoubleSection&) = delete;
  DoubleSection& operator=(const DoubleSection&) = delete;
  virtual ~DoubleSection() = default;

  // numbers of VARIABLE parameters (nelson/DoubleSection.h:40-166)
  virtual int numParametersU() const { return nU_; }
  virtual int numParametersV() const { return nV_; }
  // per-parameter sizes, user order (the reference's mat::Variable block
  // sizes, BaseNumSizeParametersU<mat::Variable, ...>::parameterUSize,
  // nelson/DoubleSection.h:76-110): each at most BU / BV; empty (the
  // default): every parameter has the block size BU / BV
  virtual const std::vector<int>& parameterUSize() const { return noSizes_; }
  virtual const std::vector<int>& parameterVSize() const { return noSizes_; }

  void parametersReady() {
    if (structured_) throw Error(NSN_ESTATE, "parametersReady after structureReady");
    ready_ = true;
    u2i_.resize(numParametersU());
    std::iota(u2i_.begin(), u2i_.end(), 0);
  }

  // addEdge takes ownership of the edge (nelson/DoubleSection.hpp:338-365)
  template <class E>
  void addEdge(NodeId i, nelson_amd::EdgeUnaryU<Derived, E>* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    checkAdd(i, true);
    e->bind_(&self(), i);
    edges_.push_back(std::move(own));
  }
  template <class E>
  void addEdge(NodeId i, nelson_amd::EdgeUnaryV<Derived, E>* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    checkAdd(i, false);
    e->bind_(&self(), i);
    edges_.push_back(std::move(own));
  }
  template <class E>
  void addEdge(NodeId i, NodeId j, nelson_amd::EdgeBinaryUV<Derived, E>* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    checkAdd(i, true);
    checkAdd(j, false);
    e->bind_(&self(), i, j);
    if (i.isVariable() && j.isVariable()) uv_.push_back({edges_.size(), {i.id(), j.id()}});
    edges_.push_back(std::move(own));
  }
  template <class E>
  void addEdge(NodeId i, NodeId j, nelson_amd::EdgeBinaryUU<Derived, E>* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    checkAdd(i, true);
    checkAdd(j, true);
    if (i.isVariable() && j.isVariable() && i.id() == j.id())
      throw Error(NSN_EINVAL, "addEdge: a binary U edge joins a parameter to itself");
    e->bind_(&self(), i, j);
    if (i.isVariable() && j.isVariable()) uu_.push_back({edges_.size(), {i.id(), j.id()}});
    edges_.push_back(std::move(own));
  }
  template <class E>
  void addEdge(NodeId i, NodeId j, nelson_amd::EdgeBinaryVV<Derived, E>* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    checkAdd(i, false);
    checkAdd(j, false);
    if (i.isVariable() && j.isVariable() && i.id() == j.id())
      throw Error(NSN_EINVAL, "addEdge: a binary V edge joins a parameter to itself");
    e->bind_(&self(), i, j);
    if (i.isVariable() && j.isVariable()) vv_.push_back({edges_.size(), {i.id(), j.id()}});
    edges_.push_back(std::move(own));
  }

  // an n-ary edge on parameters of the first section
  template <class E, int N>
  void addEdge(const std::array<NodeId, N>& ids, nelson_amd::EdgeNaryU<Derived, E, N>* e) {
    addNary(std::vector<NodeId>(ids.begin(), ids.end()), e);
  }
  template <class E>
  void addEdge(const std::vector<NodeId>& ids, nelson_amd::EdgeNaryU<Derived, E, kDynamic>* e) {
    addNary(ids, e);
  }

  // the internal order of the U parameters (bU and the increment's U part
  // follow it; nelson/DoubleSection.hpp:30-63): u2i[user index] = internal
  void setUser2InternalIndexesU(const std::vector<int>& u2i) {
    if (!ready_ || structured_) throw Error(NSN_ESTATE, "setUser2InternalIndexesU outside parametersReady .. structureReady");
    std::vector<int> s(u2i);
    std::sort(s.begin(), s.end());
    for (int k = 0; k < (int)s.size(); k++)
      if (s[k] != k || s.size() != u2i_.size()) throw Error(NSN_EINVAL, "setUser2InternalIndexesU: not a permutation");
    u2i_ = u2i;
  }
  // permuteAMD_SchurU (nelson/DoubleSection.hpp:30-63): a fill-reducing
  // order of the U parameters on the pattern of U + W W^T, by greedy minimum
  // degree with exact degrees (the role Eigen's AMD plays in the reference).
  // S is factorised dense here, so the order only fixes the internal order.
  void permuteAMD_SchurU() {
    if (!ready_ || structured_) throw Error(NSN_ESTATE, "permuteAMD_SchurU outside parametersReady .. structureReady");
    const int n = numParametersU();
    std::vector<std::set<int>> adj(n);
    for (auto* lst : {&uu_, &nary_})
      for (auto& e : *lst) adj[e.second.first].insert(e.second.second), adj[e.second.second].insert(e.second.first);
    std::map<int, std::vector<int>> byV;
    for (auto& e : uv_) byV[e.second.second].push_back(e.second.first);
    for (auto& kv : byV)
      for (int a : kv.second)
        for (int b : kv.second)
          if (a != b) adj[a].insert(b);
    std::vector<char> done(n, 0);
    for (int k = 0; k < n; k++) {
      int best = -1;
      for (int v = 0; v < n; v++)
        if (!done[v] && (best < 0 || adj[v].size() < adj[best].size())) best = v;
      u2i_[best] = k;
      done[best] = 1;
      std::vector<int> nb(adj[best].begin(), adj[best].end());
      for (int a : nb) {
        adj[a].erase(best);
        for (int b : nb)
          if (a != b) adj[a].insert(b);
      }
      adj[best].clear();
    }
  }
  const std::vector<int>& user2internalIndexesU() const { return u2i_; }

  // structureReady (nelson/DoubleSection.hpp:83-246): the block structure in
  // UID order (row-major, internal indices) and the device problem
  void structureReady(Precision p = Precision::FP64, int device = 0) {
    if (!ready_) throw Error(NSN_ESTATE, "structureReady before parametersReady");
    if (structured_) throw Error(NSN_ESTATE, "structureReady called twice");
    std::vector<std::pair<int, int>> w, o;
    for (auto& e : uv_) w.push_back({u2i_[e.second.first], e.second.second});
    for (auto* lst : {&uu_, &nary_})
      for (auto& e : *lst) {
        const int a = u2i_[e.second.first], b = u2i_[e.second.second];
        o.push_back({std::min(a, b), std::max(a, b)});
      }
    auto uniq = [](std::vector<std::pair<int, int>>& v) {
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
    };
    std::vector<std::pair<int, int>> vo;
    for (auto& e : vv_) vo.push_back({std::min(e.second.first, e.second.second), std::max(e.second.first, e.second.second)});
    uniq(w);
    uniq(o);
    uniq(vo);
    for (auto& e : vv_) {
      const int a = e.second.first, b = e.second.second;
      edges_[e.first]->uid_ =
          (int)(std::lower_bound(vo.begin(), vo.end(), std::make_pair(std::min(a, b), std::max(a, b))) - vo.begin());
      edges_[e.first]->uidTrans_ = a > b;
    }
    for (auto& e : uv_)
      edges_[e.first]->uid_ =
          (int)(std::lower_bound(w.begin(), w.end(), std::make_pair(u2i_[e.second.first], e.second.second)) - w.begin());
    for (auto& e : uu_) {
      const int a = u2i_[e.second.first], b = u2i_[e.second.second];
      edges_[e.first]->uid_ =
          (int)(std::lower_bound(o.begin(), o.end(), std::make_pair(std::min(a, b), std::max(a, b))) - o.begin());
      edges_[e.first]->uidTrans_ = a > b;
    }
    for (auto& e : nary_) {
      const int a = u2i_[e.second.first], b = u2i_[e.second.second];
      auto& ed = *edges_[e.first];
      ed.pairUid_.push_back(
          (int)(std::lower_bound(o.begin(), o.end(), std::make_pair(std::min(a, b), std::max(a, b))) - o.begin()));
      ed.pairTrans_.push_back(a > b ? 1 : 0);
    }
    std::vector<int> suI, svI;
    {
      const std::vector<int>& su = parameterUSize();
      const std::vector<int>& sv = parameterVSize();
      if (!su.empty()) {
        if ((int)su.size() != numParametersU()) throw Error(NSN_EINVAL, "parameterUSize: one size per U parameter");
        suI.assign(su.size(), 0);
        for (size_t i = 0; i < su.size(); i++) {
          if (su[i] < 1 || su[i] > BU) throw Error(NSN_EINVAL, "parameterUSize: sizes from 1 to the block size");
          suI[u2i_[i]] = su[i];
        }
      }
      if (!sv.empty()) {
        if ((int)sv.size() != numParametersV()) throw Error(NSN_EINVAL, "parameterVSize: one size per V parameter");
        for (int x : sv)
          if (x < 1 || x > BV) throw Error(NSN_EINVAL, "parameterVSize: sizes from 1 to the block size");
        svI = sv;
      }
    }
    H_.setup_(this, numParametersU(), numParametersV(), w, o, vo, p, device, suI, svI);
    structured_ = true;
  }

  // DoubleSection::update(bool) (nelson/DoubleSection.hpp:452-467): every
  // edge in insertion order (chi2 = sum), then with `hessians` the cleared
  // blocks receive each edge's updaters, in reverse insertion order as the
  // reference's per-block lists run
  void update(bool hessians) {
    if (!structured_) throw Error(NSN_ESTATE, "update before structureReady");
    if (hessians) H_.clear_();
    double chi2 = 0.0;
    for (auto& e : edges_) {
      e->update(hessians);
      chi2 += e->chi2();
    }
    H_.setChi2_(chi2);
    if (hessians) {
      for (auto it = edges_.rbegin(); it != edges_.rend(); ++it) (*it)->applyUpdaters_();
      H_.touch_();
    }
  }
  const HessianT& hessian() const { return H_; }

  // nelson/DoubleSection.h:300-311, with b of the last update(true)
  double computeRhoChi2Change(double mu, const HessianVectors& inc, double oldChi2) const {
    const auto& bu = H_.bUvalues();  // BU / BV values per parameter, its size used
    const auto& bv = H_.bVvalues();
    double du = 0, dv = 0;
    for (int u = 0; u < H_.numU(); u++) {
      const ConstSeg x = inc.bU().segment(u);
      for (int a = 0; a < x.size(); a++) du += x(a) * (mu * x(a) - bu[(size_t)BU * u + a]);
    }
    for (int q = 0; q < H_.numV(); q++) {
      const ConstSeg x = inc.bV().segment(q);
      for (int a = 0; a < x.size(); a++) dv += x(a) * (mu * x(a) - bv[(size_t)BV * q + a]);
    }
    return (oldChi2 - H_.chi2()) / (du + dv);
  }

  // section internals used by the edges
  HessianT& hessianStorage_() { return H_; }
  int internalU_(int userId) const { return u2i_[userId]; }

 protected:
  int sizeU_() const { return nU_; }

 private:
  Derived& self() { return static_cast<Derived&>(*this); }
  template <class Edge>
  void addNary(const std::vector<NodeId>& ids, Edge* e) {
    std::unique_ptr<detail::EdgeBase> own(e);
    std::vector<int> vars;
    for (const NodeId& i : ids) {
      checkAdd(i, true);
      if (i.isVariable()) vars.push_back(i.id());
    }
    std::vector<int> sorted(vars);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      throw Error(NSN_EINVAL, "addEdge: an n-ary edge lists a parameter twice");
    e->bind_(&self(), ids);
    for (size_t a = 0; a < vars.size(); a++)
      for (size_t b = a + 1; b < vars.size(); b++) nary_.push_back({edges_.size(), {vars[a], vars[b]}});
    edges_.push_back(std::move(own));
  }
  void checkAdd(NodeId i, bool isU) {
    if (!ready_) throw Error(NSN_ESTATE, "addEdge before parametersReady");
    if (structured_) throw Error(NSN_ESTATE, "addEdge after structureReady");
    const int n = isU ? numParametersU() : numParametersV();
    if (i.isVariable() && (i.id() < 0 || i.id() >= n))
      throw Error(NSN_EINVAL, "addEdge: variable parameter " + std::to_string(i.id()) + " out of range");
  }
  int nU_ = 0, nV_ = 0;
  bool ready_ = false, structured_ = false;
  std::vector<int> u2i_;
  std::vector<std::unique_ptr<detail::EdgeBase>> edges_;
  std::vector<std::pair<size_t, std::pair<int, int>>> uv_, uu_;  // (edge, (user U id, V / U id))
  std::vector<std::pair<size_t, std::pair<int, int>>> vv_;       // (edge, (V id, V id))
  std::vector<std::pair<size_t, std::pair<int, int>>> nary_;    // (edge, variable pair), in pair order
  HessianT H_;
  const std::vector<int> noSizes_;
};

// ---------------------------------------------------------------------------
// A single section (nelson/SingleSection.hpp:63-368: pose graphs, scan
// alignment): one parameter set, unary and binary edges between its
// parameters.  It is a DoubleSection without a second section, so H = U and
// the GPU factorises it directly (S = U, no V blocks, no W).
// ---------------------------------------------------------------------------
namespace detail {
struct NoParameter {};
}  // namespace detail

template <class Derived, class Par, int B>
class SingleSection : public DoubleSection<Derived, Par, detail::NoParameter, B, 1> {
  using Base = DoubleSection<Derived, Par, detail::NoParameter, B, 1>;

 public:
  using ParameterType = Par;
  using HessianVecType = HessianVectors;
  // edges: update(bool) with setChi2; unary: updateHBlock(MatRef H, VecRef
  // b); binary: updateH11Block(H, b), updateH22Block(H, b),
  // updateH12Block(MatRef H) with H the block of rows parameter_1, columns
  // parameter_2 (already oriented: no transpose flag)
  template <class E>
  using EdgeUnary = nelson_amd::EdgeUnaryU<Derived, E>;
  template <class E>
  using EdgeBinary = nelson_amd::EdgeBinaryUU<Derived, E>;
  template <class E, int N = kDynamic>
  using EdgeNary = nelson_amd::EdgeNaryU<Derived, E, N>;

  SingleSection() = default;
  explicit SingleSection(int numParameters) : Base(numParameters, 0) {}

  // the user implements parameter(NodeId) (both constnesses)
  virtual int numParameters() const { return this->sizeU_(); }
  // per-parameter sizes (mat::Variable, nelson/SingleSection.h:79-110): one
  // size per parameter, each at most B; empty: every parameter has size B
  virtual const std::vector<int>& parameterSize() const { return noSizes_; }
  const std::vector<int>& parameterUSize() const override { return parameterSize(); }
  int numParametersU() const override { return static_cast<const Derived&>(*this).numParameters(); }
  int numParametersV() const override { return 0; }
  Par& parameterU(NodeId i) { return static_cast<Derived&>(*this).parameter(i); }
  const Par& parameterU(NodeId i) const { return static_cast<const Derived&>(*this).parameter(i); }

  // nelson/SingleSection.hpp names of the ordering calls
  void setUser2InternalIndexes(const std::vector<int>& u2i) { Base::setUser2InternalIndexesU(u2i); }
  void permuteAMD() { Base::permuteAMD_SchurU(); }
  const std::vector<int>& user2internalIndexes() const { return Base::user2internalIndexesU(); }

 private:
  const std::vector<int> noSizes_;
};

// ---------------------------------------------------------------------------
// SolverCholeskySchur (nelson/SolverCholeskySchur.h:53-117) on the GPU
// ---------------------------------------------------------------------------
// time points of the last computeIncrement, milliseconds after
// t0_startIteration, with the reference's names
// (SolverCholeskySchurIterationTimeStat, nelson/SolverCholeskySchur.h:25-38):
// the W and U refreshes and the solve set-up are empty here (the blocks are
// read in place, the symbolic work is done at structureReady), bS and S are
// assembled together, V~ and V together; refinement (mixed precision only)
// follows t10.
struct SolverCholeskySchurIterationTimeStat {
  double t1_VInvbVComputed = 0, t2_WRefreshed = 0, t3_URefreshed = 0, t4_bSComputed = 0, t5_SComputed = 0,
         t6_SSolveInit = 0, t7_SFactorized = 0, t8_bUComputed = 0, t9_bVtildeComputed = 0, t10_bVComputed = 0,
         refinementDone = 0;
  std::string toString() const {
    std::ostringstream os;
    os << "VInvbV " << t1_VInvbVComputed << " ms, S " << t5_SComputed - t3_URefreshed << " ms, factorize "
       << t7_SFactorized - t6_SSolveInit << " ms, bU " << t8_bUComputed - t7_SFactorized << " ms, bV "
       << t10_bVComputed - t8_bUComputed << " ms, refinement " << refinementDone - t10_bVComputed << " ms";
    return os.str();
  }
};

class SolverCholeskySchur {
 public:
  using Type = double;
  // init(H, b) (.hpp:44-62): binds the section's Hessian; the structure was
  // analysed at structureReady
  void init(const SchurSystem& H, const SchurSystem& b) {
    (void)b;
    sys_ = &H;
    H.prepare();
    check(nsn_schur_init(H.problem()), "nsn_schur_init");
  }
  double maxAbsHDiag() const { return sys().maxAbsHDiag(); }
  // computeIncrement(H, b, rel, abs) (.hpp:91-157): false (and a zero
  // increment) when the factorisation fails
  bool computeIncrement(const SchurSystem& H, const SchurSystem& b, double relLambda, double absLambda) {
    (void)b;
    sys_ = &H;
    H.prepare();
    const int NU = H.numU(), NV = H.numV(), BU = H.blockU(), BV = H.blockV();
    // with V factorised together with S the device's U part holds the V
    // parameters first (6 values each), then the U parameters
    const bool vu = H.vFirstInU();
    const int offU = vu ? NV : 0;
    std::vector<double> du(6 * (size_t)(NU + offU) + 1), dv(3 * (size_t)(vu ? 0 : NV) + 1);
    double sq = 0;
    const int rc = nsn_schur_step(H.problem(), relLambda, absLambda, du.data(), dv.data(), &sq);
    check(rc, "nsn_schur_step");
    inc_.origin = H.owner();
    inc_.u_.B_ = BU;
    inc_.v_.B_ = BV;
    // per-parameter sizes: block u at off_[u] (compact, as the reference's
    // vector of a mat::Variable section); otherwise B values per block
    const bool var = H.variableSizes();
    inc_.u_.off_.assign(var ? NU + 1 : 0, 0);
    inc_.v_.off_.assign(var ? NV + 1 : 0, 0);
    if (var) {
      for (int u = 0; u < NU; u++) inc_.u_.off_[u + 1] = inc_.u_.off_[u] + H.sizeU(u);
      for (int q = 0; q < NV; q++) inc_.v_.off_[q + 1] = inc_.v_.off_[q] + H.sizeV(q);
    }
    inc_.u_.v_.resize(var ? (size_t)inc_.u_.off_[NU] : (size_t)BU * NU);
    inc_.v_.v_.resize(var ? (size_t)inc_.v_.off_[NV] : (size_t)BV * NV);
    for (int u = 0; u < NU; u++) {
      const size_t o = var ? (size_t)inc_.u_.off_[u] : (size_t)BU * u;
      for (int a = 0; a < H.sizeU(u); a++) inc_.u_.v_[o + a] = du[6 * (size_t)(offU + u) + a];
    }
    for (int q = 0; q < NV; q++) {
      const size_t o = var ? (size_t)inc_.v_.off_[q] : (size_t)BV * q;
      for (int a = 0; a < H.sizeV(q); a++) inc_.v_.v_[o + a] = vu ? du[6 * (size_t)q + a] : dv[3 * (size_t)q + a];
    }
    // the padding's increment is exactly zero, so the norm is the device's
    sq_ = sq;
    return rc == NSN_OK;
  }
  const HessianVectors& incrementVector() const { return inc_; }
  double incrementVectorSquaredNorm() const { return sq_; }
  SolverCholeskySchurIterationTimeStat timingStats() const {
    double ms[6] = {};
    check(nsn_phase_times(sys().problem(), ms, 6), "nsn_phase_times");
    SolverCholeskySchurIterationTimeStat t;
    t.t1_VInvbVComputed = ms[0];
    t.t2_WRefreshed = t.t3_URefreshed = ms[0];
    t.t4_bSComputed = t.t5_SComputed = t.t6_SSolveInit = ms[0] + ms[1];
    t.t7_SFactorized = t.t6_SSolveInit + ms[2];
    t.t8_bUComputed = t.t7_SFactorized + ms[3];
    t.t9_bVtildeComputed = t.t10_bVComputed = t.t8_bUComputed + ms[4];
    t.refinementDone = t.t10_bVComputed + ms[5];
    return t;
  }

 private:
  const SchurSystem& sys() const {
    if (!sys_) throw Error(NSN_ESTATE, "SolverCholeskySchur used before init");
    return *sys_;
  }
  const SchurSystem* sys_ = nullptr;
  HessianVectors inc_;
  double sq_ = 0.0;
};

// ---------------------------------------------------------------------------
// drivers (nelson/LevenbergMarquardt.h, nelson/GaussNewton.h).  On a
// BASection the loop runs on the device (one graph launch per
// sub-iteration); on a user section the loop runs on the host, calling the
// section and the solver in the reference's order.  Settings, termination
// reasons and the per-sub-iteration (it, subit, chi2, mu) log keep the
// reference's meaning.
// ---------------------------------------------------------------------------
// The reference's solver selection (nelson/SolverTraits.h:12-55,
// SolverTraitsBase.h:12-40, OrderingTraits.h:9-11, MatrixWrapperTraits.h:11-12):
// every solver kind, wrapper and ordering names the GPU Schur solver here
// (dense and envelope storage of S are chosen from its structure, the
// camera order by the section), so code written against
// SolverTraits<kind>::Solver<HessianTraits, ...> compiles unchanged.
constexpr int solverCholeskyDense = 1;
constexpr int solverCholeskySparse = 2;
constexpr int solverCholeskySchur = 3;
constexpr int solverCholeskySchurDiagBlockInverse = 4;
constexpr int solverCholeskySchurDiagBlockInverseWWtMult = 5;
constexpr int choleskyNaturalOrdering = 0;
constexpr int choleskyAMDOrdering = 1;
constexpr int choleskyCOLAMDOrdering = 2;
constexpr int matrixWrapperDense = 1;
constexpr int matrixWrapperSparse = 2;
template <int kind>
struct SolverTraits {
  template <class HessianTraits, int... options>
  using Solver = SolverCholeskySchur;
};

struct LevenbergMarquardtSettings : nsn_lm_settings {
  LevenbergMarquardtSettings() { nsn_lm_default_settings(this); }
};
struct GaussNewtonSettings : nsn_gn_settings {
  GaussNewtonSettings() { nsn_gn_default_settings(this); }
};

struct Stats {  // LevenbergMarquardtStats / GaussNewtonStats
  std::vector<nsn_log_entry> log;
  void add(int it, int subit, double chi2, double mu) { log.push_back(nsn_log_entry{it, subit, chi2, mu}); }
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

  // the BA section: the loop on the device
  TerminationReason solve(BASection& op) {
    const int cap = settings_.maxNumIt * settings_.maxNumSubIt + 8;
    stats_.log.assign(cap > 0 ? cap : 8, nsn_log_entry{});
    check(nsn_lm_solve(op.handle(), &settings_, &res_, stats_.log.data(), (int)stats_.log.size()), "nsn_lm_solve");
    stats_.log.resize(res_.n_log < (int)stats_.log.size() ? res_.n_log : stats_.log.size());
    op.download();
    return (TerminationReason)res_.reason;
  }

  // a user section: LevenbergMarquardt::solve (nelson/LevenbergMarquardt.hpp:35-146)
  template <class Section>
  TerminationReason solve(Section& op) {
    const auto& s = settings_;
    stats_.log.clear();
    res_ = nsn_lm_result{};
    op.update(true);
    stats_.add(-1, -1, op.hessian().chi2(), 0.0);
    auto finish = [&](TerminationReason r, int it) {
      res_.reason = (int)r;
      res_.iterations = it;
      res_.n_log = (int)stats_.log.size();
      res_.final_chi2 = op.hessian().chi2();
      return r;
    };
    if (op.hessian().maxAbsValBVect() <= s.epsBVector) return finish(TerminationReason::SuccessAtStart, 0);
    op.backupSolution();
    double oldChi2 = op.hessian().chi2();
    solver_.init(op.hessian().H(), op.hessian().b());
    double mu = s.tau * solver_.maxAbsHDiag();
    double curV = s.initV;
    for (int iter = 0; iter < s.maxNumIt; iter++) {
      int subiter = 0;
      while (true) {
        const bool ok = solver_.computeIncrement(op.hessian().H(), op.hessian().b(), 0.0, mu);
        res_.sub_iterations++;
        if (ok && iter >= s.minNumIt && solver_.incrementVectorSquaredNorm() <= s.epsIncSquare)
          return finish(TerminationReason::SuccessSmallIncrement, iter);
        double rho = -1;
        if (ok) {
          op.oplus(solver_.incrementVector());
          op.update(false);
          stats_.add(iter, subiter, op.hessian().chi2(), mu);
          rho = op.computeRhoChi2Change(mu, solver_.incrementVector(), oldChi2);
        } else {
          stats_.add(iter, subiter, -1.0, mu);
        }
        if (ok && rho > 0) {
          if (iter > s.minNumIt && op.hessian().maxAbsValBVect() <= s.epsBVector)
            return finish(TerminationReason::SuccessEpsB, iter);
          if (iter > s.minNumIt && op.hessian().chi2() < s.epsChi2)
            return finish(TerminationReason::SuccessSmallResiduals, iter);
          const double fc = 2 * rho - 1;
          mu *= std::max(s.minScaleFactor, std::min(s.maxScaleFactor, 1 - fc * fc * fc));
          curV = s.initV;
        } else {
          op.rollbackSolution();
          mu *= curV;
          curV *= s.penaltyV;
        }
        if (++subiter >= s.maxNumSubIt) return finish(TerminationReason::MaxSubIterationReached, iter);
        if (ok && rho > 0) break;
      }
      op.update(true);
      op.backupSolution();
      oldChi2 = op.hessian().chi2();
    }
    return finish(TerminationReason::MaxIterationReached, s.maxNumIt);
  }
  const Stats& stats() const { return stats_; }
  int numIterations() const { return res_.iterations; }
  const nsn_lm_result& result() const { return res_; }

 private:
  LevenbergMarquardtSettings settings_;
  Stats stats_;
  nsn_lm_result res_{};
  Solver solver_;
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

  // a user section: GaussNewton::solve (nelson/GaussNewton.hpp:31-127)
  template <class Section>
  TerminationReason solve(Section& op) {
    const auto& s = settings_;
    stats_.log.clear();
    res_ = nsn_lm_result{};
    op.update(true);
    stats_.add(-1, -1, op.hessian().chi2(), 0.0);
    auto finish = [&](TerminationReason r, int it) {
      res_.reason = (int)r;
      res_.iterations = it;
      res_.n_log = (int)stats_.log.size();
      res_.final_chi2 = op.hessian().chi2();
      return r;
    };
    if (op.hessian().maxAbsValBVect() <= s.epsBVector) return finish(TerminationReason::SuccessAtStart, 0);
    solver_.init(op.hessian().H(), op.hessian().b());
    for (int iter = 0; iter < s.maxNumIt; iter++) {
      const bool ok = solver_.computeIncrement(op.hessian().H(), op.hessian().b(), s.relLambda, s.absLambda);
      res_.sub_iterations++;
      if (!ok) return finish(TerminationReason::SolverFailure, iter);
      if (iter >= s.minNumIt && solver_.incrementVectorSquaredNorm() <= s.epsIncSquare)
        return finish(TerminationReason::SuccessSmallIncrement, iter);
      op.oplus(solver_.incrementVector());
      op.update(true);
      stats_.add(iter, 0, op.hessian().chi2(), 0.0);
      if (iter > s.minNumIt && op.hessian().maxAbsValBVect() <= s.epsBVector)
        return finish(TerminationReason::SuccessEpsB, iter);
      if (iter > s.minNumIt && op.hessian().chi2() < s.epsChi2)
        return finish(TerminationReason::SuccessSmallResiduals, iter);
    }
    return finish(TerminationReason::MaxIterationReached, s.maxNumIt);
  }
  const Stats& stats() const { return stats_; }
  int numIterations() const { return res_.iterations; }

 private:
  GaussNewtonSettings settings_;
  Stats stats_;
  nsn_lm_result res_{};
  Solver solver_;
};

}  // namespace nelson_amd

#endif
