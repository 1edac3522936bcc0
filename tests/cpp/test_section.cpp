// A user section with user edges through the C++ interface, written the way
// the reference's tests write theirs (nelson/tests/testDoubleSectionCommon.h:
// 2-D points in the first section, 3-D points in the second; unary edges on
// both, U-U difference edges, U-V edges on the first two coordinates; the
// measurements are exact for the ground truth, so the optimum has chi2 = 0).
// With VV = 0 the second section is block diagonal (no V-V edges), the
// configuration of the reference's Schur tests
// (nelson/tests/testDoubleSectionSchurCommon.hpp); with VV = 1 every pair of
// 3-D points has a V-V difference edge, as in the reference's DoubleSection
// tests with a non-block-diagonal V (testDoubleSectionCommon.hpp:31-38).
//
// Reads the problem from the file given as argv[1]:
//   N2 N3 UU WALL VV       (UU: 1 = U-U edges on every pair; WALL: 1 = a U-V
//                           edge on every pair, 0 = on (i, i) only; VV: 1 =
//                           V-V edges on every pair)
//   N2 lines gt2, N3 lines gt3, N2 lines p2, N3 lines p3 (the start)
// and prints one JSON object per solve: Gauss-Newton and Levenberg-Marquardt
// with the default settings, LM with the stopping tests off, LM with a
// reversed and with a minimum-degree internal order of the U parameters,
// and one Schur step.  The GPU test compares them with the oracle.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "nelson_amd.hpp"

using namespace nelson_amd;

struct Point2d {
  double p[2];
};
struct Point3d {
  double p[3];
};

template <class Section>
class EdgeUnaryPoint2d : public EdgeUnaryU<Section, EdgeUnaryPoint2d<Section>> {
  double meas_[2], err_[2] = {0, 0};

 public:
  explicit EdgeUnaryPoint2d(const double* m) : meas_{m[0], m[1]} {}
  void update(bool) override {
    const auto& par = this->parameter();
    for (int k = 0; k < 2; k++) err_[k] = par.p[k] - meas_[k];
    this->setChi2(err_[0] * err_[0] + err_[1] * err_[1]);
  }
  void updateHBlock(MatRef H, VecRef b) {
    H.addIdentity();
    for (int k = 0; k < 2; k++) b(k) += err_[k];
  }
};

template <class Section>
class EdgeUnaryPoint3d : public EdgeUnaryV<Section, EdgeUnaryPoint3d<Section>> {
  double meas_[3], err_[3] = {0, 0, 0};

 public:
  explicit EdgeUnaryPoint3d(const double* m) : meas_{m[0], m[1], m[2]} {}
  void update(bool) override {
    const auto& par = this->parameter();
    double c = 0;
    for (int k = 0; k < 3; k++) err_[k] = par.p[k] - meas_[k], c += err_[k] * err_[k];
    this->setChi2(c);
  }
  void updateHBlock(MatRef H, VecRef b) {
    H.addIdentity();
    for (int k = 0; k < 3; k++) b(k) += err_[k];
  }
};

// (p_i - p_j) - meas between two 2-D points
template <class Section>
class EdgeBinaryPoint2d : public EdgeBinaryUU<Section, EdgeBinaryPoint2d<Section>> {
  double meas_[2], err_[2] = {0, 0};

 public:
  explicit EdgeBinaryPoint2d(const double* m) : meas_{m[0], m[1]} {}
  void update(bool) override {
    const auto& a = this->parameter_1();
    const auto& b = this->parameter_2();
    for (int k = 0; k < 2; k++) err_[k] = (a.p[k] - b.p[k]) - meas_[k];
    this->setChi2(err_[0] * err_[0] + err_[1] * err_[1]);
  }
  void updateH11Block(MatRef H, VecRef v) {
    H.addIdentity();
    for (int k = 0; k < 2; k++) v(k) += err_[k];
  }
  void updateH22Block(MatRef H, VecRef v) {
    H.addIdentity();
    for (int k = 0; k < 2; k++) v(k) -= err_[k];
  }
  void updateH12Block(MatRef H) { H.addIdentity(-1.0); }
};

// (p_i - p_j) - meas between two 3-D points
template <class Section>
class EdgeBinaryPoint3d : public EdgeBinaryVV<Section, EdgeBinaryPoint3d<Section>> {
  double meas_[3], err_[3] = {0, 0, 0};

 public:
  explicit EdgeBinaryPoint3d(const double* m) : meas_{m[0], m[1], m[2]} {}
  void update(bool) override {
    const auto& a = this->parameter_1();
    const auto& b = this->parameter_2();
    double c = 0;
    for (int k = 0; k < 3; k++) err_[k] = (a.p[k] - b.p[k]) - meas_[k], c += err_[k] * err_[k];
    this->setChi2(c);
  }
  void updateH11Block(MatRef H, VecRef v) {
    H.addIdentity();
    for (int k = 0; k < 3; k++) v(k) += err_[k];
  }
  void updateH22Block(MatRef H, VecRef v) {
    H.addIdentity();
    for (int k = 0; k < 3; k++) v(k) -= err_[k];
  }
  void updateH12Block(MatRef H) { H.addIdentity(-1.0); }
};

// (p2_i - p3_j.head(2)) - meas between a 2-D and a 3-D point
template <class Section>
class EdgeBinaryPoint2d3d : public EdgeBinaryUV<Section, EdgeBinaryPoint2d3d<Section>> {
  double meas_[2], err_[2] = {0, 0};

 public:
  explicit EdgeBinaryPoint2d3d(const double* m) : meas_{m[0], m[1]} {}
  void update(bool) override {
    const auto& a = this->parameter_1();
    const auto& b = this->parameter_2();
    for (int k = 0; k < 2; k++) err_[k] = (a.p[k] - b.p[k]) - meas_[k];
    this->setChi2(err_[0] * err_[0] + err_[1] * err_[1]);
  }
  void updateH11Block(MatRef H, VecRef v) {
    H.addIdentity();
    for (int k = 0; k < 2; k++) v(k) += err_[k];
  }
  void updateH22Block(MatRef H, VecRef v) {  // diag(1, 1, 0), (err, 0)
    H(0, 0) += 1.0;
    H(1, 1) += 1.0;
    for (int k = 0; k < 2; k++) v(k) -= err_[k];
  }
  void updateH12Block(MatRef H) {
    H(0, 0) -= 1.0;
    H(1, 1) -= 1.0;
  }
};

class Points2d3d : public DoubleSection<Points2d3d, Point2d, Point3d, 2, 3> {
 public:
  Points2d3d(const std::vector<Point2d>& p2, const std::vector<Point3d>& p3)
      : DoubleSection((int)p2.size(), (int)p3.size()), p2_(p2), p3_(p3) {}
  Point2d& parameterU(NodeId i) { return i.isVariable() ? p2_.at(i.id()) : fixed2_; }
  const Point2d& parameterU(NodeId i) const { return i.isVariable() ? p2_.at(i.id()) : fixed2_; }
  Point3d& parameterV(NodeId i) { return i.isVariable() ? p3_.at(i.id()) : fixed3_; }
  const Point3d& parameterV(NodeId i) const { return i.isVariable() ? p3_.at(i.id()) : fixed3_; }
  // the increment's U part is in internal order
  void oplus(const HessianVectors& inc) {
    for (size_t i = 0; i < p2_.size(); i++) {
      const ConstSeg d = inc.bU().segment(user2internalIndexesU()[i]);
      for (int k = 0; k < 2; k++) p2_[i].p[k] += d[k];
    }
    for (size_t j = 0; j < p3_.size(); j++) {
      const ConstSeg d = inc.bV().segment((int)j);
      for (int k = 0; k < 3; k++) p3_[j].p[k] += d[k];
    }
  }
  void backupSolution() { b2_ = p2_, b3_ = p3_; }
  void rollbackSolution() { p2_ = b2_, p3_ = b3_; }

 private:
  std::vector<Point2d> p2_, b2_;
  std::vector<Point3d> p3_, b3_;
  Point2d fixed2_{{0, 0}};
  Point3d fixed3_{{0, 0, 0}};
};

struct ToyData {
  int N2 = 0, N3 = 0, uu = 1, wall = 1, vv = 0;
  std::vector<Point2d> gt2, p2;
  std::vector<Point3d> gt3, p3;
};

static ToyData read(const char* path) {
  std::ifstream f(path);
  ToyData d;
  f >> d.N2 >> d.N3 >> d.uu >> d.wall >> d.vv;
  d.gt2.resize(d.N2), d.p2.resize(d.N2), d.gt3.resize(d.N3), d.p3.resize(d.N3);
  for (auto& q : d.gt2) f >> q.p[0] >> q.p[1];
  for (auto& q : d.gt3) f >> q.p[0] >> q.p[1] >> q.p[2];
  for (auto& q : d.p2) f >> q.p[0] >> q.p[1];
  for (auto& q : d.p3) f >> q.p[0] >> q.p[1] >> q.p[2];
  if (!f) throw Error(NSN_EINVAL, std::string("cannot read ") + path);
  return d;
}

// testDoubleSectionCommon.hpp:7-60: the edges in the reference's order
static void build(Points2d3d& pss, const ToyData& d, int order) {
  pss.parametersReady();
  for (int i = 0; i < d.N2; i++) pss.addEdge(i, new EdgeUnaryPoint2d<Points2d3d>(d.gt2[i].p));
  for (int j = 0; j < d.N3; j++) pss.addEdge(j, new EdgeUnaryPoint3d<Points2d3d>(d.gt3[j].p));
  if (d.uu)
    for (int i = 0; i < d.N2; i++)
      for (int j = i + 1; j < d.N2; j++) {
        const double m[2] = {d.gt2[i].p[0] - d.gt2[j].p[0], d.gt2[i].p[1] - d.gt2[j].p[1]};
        pss.addEdge(i, j, new EdgeBinaryPoint2d<Points2d3d>(m));
      }
  if (d.vv)
    for (int i = 0; i < d.N3; i++)
      for (int j = i + 1; j < d.N3; j++) {
        const double m[3] = {d.gt3[i].p[0] - d.gt3[j].p[0], d.gt3[i].p[1] - d.gt3[j].p[1],
                             d.gt3[i].p[2] - d.gt3[j].p[2]};
        pss.addEdge(i, j, new EdgeBinaryPoint3d<Points2d3d>(m));
      }
  for (int i = 0; i < d.N2; i++)
    for (int j = 0; j < d.N3; j++) {
      if (!d.wall && i != j) continue;
      const double m[2] = {d.gt2[i].p[0] - d.gt3[j].p[0], d.gt2[i].p[1] - d.gt3[j].p[1]};
      pss.addEdge(i, j, new EdgeBinaryPoint2d3d<Points2d3d>(m));
    }
  if (order == 1) {
    std::vector<int> rev(d.N2);
    for (int i = 0; i < d.N2; i++) rev[i] = d.N2 - 1 - i;
    pss.setUser2InternalIndexesU(rev);
  } else if (order == 2) {
    pss.permuteAMD_SchurU();
  }
  pss.structureReady();
}

static void print_log(const char* name, TerminationReason tc, int its, double chi2, const Stats& st,
                      const Points2d3d& pss, int N2, int N3) {
  std::printf("{\"solve\": \"%s\", \"reason\": \"%s\", \"iterations\": %d, \"final_chi2\": %.17g, \"log\": [", name,
              toString(tc).c_str(), its, chi2);
  for (size_t i = 0; i < st.log.size(); i++)
    std::printf("%s[%d, %d, %.17g, %.17g]", i ? ", " : "", st.log[i].it, st.log[i].subit, st.log[i].chi2,
                st.log[i].mu);
  std::printf("], \"u2i\": [");
  for (int i = 0; i < N2; i++) std::printf("%s%d", i ? ", " : "", pss.user2internalIndexesU()[i]);
  std::printf("], \"p2\": [");
  for (int i = 0; i < N2; i++) std::printf("%s%.17g, %.17g", i ? ", " : "", pss.parameterU(i).p[0], pss.parameterU(i).p[1]);
  std::printf("], \"p3\": [");
  for (int j = 0; j < N3; j++)
    std::printf("%s%.17g, %.17g, %.17g", j ? ", " : "", pss.parameterV(j).p[0], pss.parameterV(j).p[1],
                pss.parameterV(j).p[2]);
  std::printf("]}\n");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s problem.txt\n", argv[0]);
    return 2;
  }
  try {
    const ToyData d = read(argv[1]);
    {
      Points2d3d pss(d.p2, d.p3);
      build(pss, d, 0);
      GaussNewton<SolverCholeskySchur> gn;
      const TerminationReason tc = gn.solve(pss);
      print_log("gn", tc, gn.numIterations(), pss.hessian().chi2(), gn.stats(), pss, d.N2, d.N3);
    }
    {
      Points2d3d pss(d.p2, d.p3);
      build(pss, d, 0);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      const TerminationReason tc = lm.solve(pss);
      print_log("lm", tc, lm.numIterations(), pss.hessian().chi2(), lm.stats(), pss, d.N2, d.N3);
    }
    {
      Points2d3d pss(d.p2, d.p3);
      build(pss, d, 0);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      lm.settings().maxNumIt = 3;
      lm.settings().epsIncSquare = -1.0;  // never: an exactly zero increment would stop the run
      lm.settings().epsChi2 = 0.0;
      lm.settings().epsBVector = 0.0;
      const TerminationReason tc = lm.solve(pss);
      print_log("lm_nostop", tc, lm.numIterations(), pss.hessian().chi2(), lm.stats(), pss, d.N2, d.N3);
    }
    for (int order : {1, 2}) {
      Points2d3d pss(d.p2, d.p3);
      build(pss, d, order);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      const TerminationReason tc = lm.solve(pss);
      print_log(order == 1 ? "lm_reversed" : "lm_min_degree", tc, lm.numIterations(), pss.hessian().chi2(),
                lm.stats(), pss, d.N2, d.N3);
    }
    {
      Points2d3d pss(d.p2, d.p3);
      build(pss, d, 1);
      pss.update(true);
      SolverCholeskySchur solver;
      solver.init(pss.hessian().H(), pss.hessian().b());
      const bool ok = solver.computeIncrement(pss.hessian().H(), pss.hessian().b(), 0.05, 0.3);
      const HessianVectors& inc = solver.incrementVector();
      std::printf("{\"solve\": \"step\", \"ok\": %s, \"sq_norm\": %.17g, \"max_abs_h_diag\": %.17g, "
                  "\"max_abs_b\": %.17g, \"chi2\": %.17g, \"u2i\": [",
                  ok ? "true" : "false", solver.incrementVectorSquaredNorm(), solver.maxAbsHDiag(),
                  pss.hessian().maxAbsValBVect(), pss.hessian().chi2());
      for (int i = 0; i < d.N2; i++) std::printf("%s%d", i ? ", " : "", pss.user2internalIndexesU()[i]);
      std::printf("], \"du\": [");
      for (size_t k = 0; k < inc.bU().size(); k++) std::printf("%s%.17g", k ? ", " : "", inc.bU()[k]);
      std::printf("], \"dv\": [");
      for (size_t k = 0; k < inc.bV().size(); k++) std::printf("%s%.17g", k ? ", " : "", inc.bV()[k]);
      std::printf("]}\n");
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "nelson_amd error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
