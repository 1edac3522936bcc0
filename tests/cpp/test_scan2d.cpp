// A single section through the C++ interface, written the way the
// reference's 2-D scan alignment test writes it
// (nelson/tests/testGaussNewton2DScansAlign.cpp): SE(2) poses, one fixed;
// binary point-to-line edges between two scans; Gauss-Newton (pure and with
// relLambda = 0.01) and Levenberg-Marquardt, in the natural, the
// minimum-degree (permuteAMD) and a reversed internal order.
//
// Reads argv[1]:
//   NP NPTS NE                  poses (pose 0 is the fixed one), points per scan, edges
//   NP lines "theta tx ty"      the start (pose 0: its fixed value)
//   NP x NPTS lines "x y nx ny" scan points and normals in the scan's frame
//   NE lines "i j"              an edge from scan i (model) to scan j (target)
// and prints one JSON object per solve.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "nelson_amd.hpp"

using namespace nelson_amd;

// pose x_w = R(theta) x + t; the update is the left one of the reference's
// SE2Pose::oplus, pose <- exp(delta) pose, delta = (rho_x, rho_y, phi)
struct SE2Pose {
  double th = 0, tx = 0, ty = 0;
  void oplus(const ConstSeg& d) {
    const double rx = d[0], ry = d[1], ph = d[2];
    double a, b;  // V(phi) = [[a, -b], [b, a]]
    if (std::fabs(ph) < 1e-8) {
      a = 1.0 - ph * ph / 6.0;
      b = ph / 2.0;
    } else {
      a = std::sin(ph) / ph;
      b = (1.0 - std::cos(ph)) / ph;
    }
    const double c = std::cos(ph), s = std::sin(ph);
    const double nx = c * tx - s * ty + a * rx - b * ry;
    const double ny = s * tx + c * ty + b * rx + a * ry;
    th += ph;
    tx = nx;
    ty = ny;
  }
};

struct Scan {
  std::vector<double> x, y, nx, ny;
};

template <class Section>
class PointLineEdge : public EdgeBinaryUU<Section, PointLineEdge<Section>> {
  const Scan *model_, *target_;
  double Hs_[9] = {}, bs_[3] = {};

 public:
  PointLineEdge(const Scan* model, const Scan* target) : model_(model), target_(target) {}
  void update(bool hessians) override {
    const SE2Pose& m = this->parameter_1();
    const SE2Pose& t = this->parameter_2();
    const double cm = std::cos(m.th), sm = std::sin(m.th), ct = std::cos(t.th), st = std::sin(t.th);
    double chi2 = 0;
    double N0 = 0, N1 = 0, NxNy = 0, NxC = 0, NyC = 0, C2 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (size_t k = 0; k < model_->x.size(); k++) {
      const double mx = cm * model_->x[k] - sm * model_->y[k] + m.tx, my = sm * model_->x[k] + cm * model_->y[k] + m.ty;
      const double nx = cm * model_->nx[k] - sm * model_->ny[k], ny = sm * model_->nx[k] + cm * model_->ny[k];
      const double px = ct * target_->x[k] - st * target_->y[k] + t.tx, py = st * target_->x[k] + ct * target_->y[k] + t.ty;
      const double e = (mx - px) * nx + (my - py) * ny;
      chi2 += e * e;
      if (hessians) {
        const double C = px * ny - py * nx;
        N0 += nx * nx, N1 += ny * ny, NxNy += nx * ny, NxC += nx * C, NyC += ny * C, C2 += C * C;
        b0 += nx * e, b1 += ny * e, b2 += C * e;
      }
    }
    this->setChi2(chi2);
    if (hessians) {
      const double H[9] = {N0, NxNy, NxC, NxNy, N1, NyC, NxC, NyC, C2};
      for (int q = 0; q < 9; q++) Hs_[q] = H[q];
      bs_[0] = b0, bs_[1] = b1, bs_[2] = b2;
    }
  }
  void updateH11Block(MatRef H, VecRef b) {
    for (int r = 0; r < 3; r++) {
      for (int c = 0; c < 3; c++) H(r, c) += Hs_[3 * r + c];
      b(r) += bs_[r];
    }
  }
  void updateH12Block(MatRef H) {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) H(r, c) -= Hs_[3 * r + c];
  }
  void updateH22Block(MatRef H, VecRef b) {
    for (int r = 0; r < 3; r++) {
      for (int c = 0; c < 3; c++) H(r, c) += Hs_[3 * r + c];
      b(r) -= bs_[r];
    }
  }
};

// the same edge as an n-ary edge on two poses
// (nelson/tests/testGaussNewton2DScansAlignNary.cpp): parameter(0) is the
// model, parameter(1) the target; N = 2 or the count set by addEdge
template <class Section, int N>
class PointLineEdgeNary : public EdgeNaryU<Section, PointLineEdgeNary<Section, N>, N> {
  PointLineEdge<Section> e_;  // the residual and its blocks

 public:
  PointLineEdgeNary(const Scan* model, const Scan* target) : e_(model, target) {}
  void update(bool hessians) override;
  void updateHBlock(int i, MatRef H, VecRef b) {
    if (i == 0) e_.updateH11Block(H, b);
    else e_.updateH22Block(H, b);
  }
  void updateHBlock(int i, int j, MatRef H, bool transpose) {
    (void)i, (void)j, (void)transpose;
    e_.updateH12Block(H);
  }
};

class SE2PoseSection : public SingleSection<SE2PoseSection, SE2Pose, 3> {
 public:
  SE2PoseSection(const std::vector<SE2Pose>& poses, const SE2Pose& fixed)
      : SingleSection((int)poses.size()), poses_(poses), fixed_(fixed) {}
  SE2Pose& parameter(NodeId i) { return i.isVariable() ? poses_.at(i.id()) : fixed_; }
  const SE2Pose& parameter(NodeId i) const { return i.isVariable() ? poses_.at(i.id()) : fixed_; }
  void oplus(const HessianVecType& inc) {
    for (int i = 0; i < numParameters(); i++) poses_[i].oplus(inc.segment(user2internalIndexes()[i]));
  }
  void backupSolution() { bck_ = poses_; }
  void rollbackSolution() { poses_ = bck_; }

 private:
  std::vector<SE2Pose> poses_, bck_;
  SE2Pose fixed_;
};

// the n-ary edge evaluates through a binary edge object that reads the
// section's parameters by the n-ary edge's ids
template <class Section, int N>
void PointLineEdgeNary<Section, N>::update(bool hessians) {
  e_.bind_(&this->section(), this->parId(0), this->parId(1));
  e_.update(hessians);
  this->setChi2(e_.chi2());
}

struct Data {
  int NP = 0, NPTS = 0, NE = 0;
  std::vector<SE2Pose> start;
  std::vector<Scan> scans;
  std::vector<std::pair<int, int>> edges;
};

static Data read(const char* path) {
  std::ifstream f(path);
  Data d;
  f >> d.NP >> d.NPTS >> d.NE;
  d.start.resize(d.NP);
  for (auto& p : d.start) f >> p.th >> p.tx >> p.ty;
  d.scans.resize(d.NP);
  for (auto& s : d.scans) {
    s.x.resize(d.NPTS), s.y.resize(d.NPTS), s.nx.resize(d.NPTS), s.ny.resize(d.NPTS);
    for (int k = 0; k < d.NPTS; k++) f >> s.x[k] >> s.y[k] >> s.nx[k] >> s.ny[k];
  }
  d.edges.resize(d.NE);
  for (auto& e : d.edges) f >> e.first >> e.second;
  if (!f) throw Error(NSN_EINVAL, std::string("cannot read ") + path);
  return d;
}

// kind 0: binary edges, 1: n-ary edges with N = 2, 2: n-ary edges with the
// count given at addEdge
static void build(SE2PoseSection& pss, const Data& d, int order, int kind = 0) {
  pss.parametersReady();
  auto id = [](int i) { return i == 0 ? NodeId::fixed(0) : NodeId(i - 1); };
  for (const auto& e : d.edges) {
    const Scan *m = &d.scans[e.first], *t = &d.scans[e.second];
    if (kind == 0) pss.addEdge(id(e.first), id(e.second), new PointLineEdge<SE2PoseSection>(m, t));
    else if (kind == 1) pss.addEdge({id(e.first), id(e.second)}, new PointLineEdgeNary<SE2PoseSection, 2>(m, t));
    else pss.addEdge(std::vector<NodeId>{id(e.first), id(e.second)},
                     new PointLineEdgeNary<SE2PoseSection, kDynamic>(m, t));
  }
  if (order == 1) {
    pss.permuteAMD();
  } else if (order == 2) {
    std::vector<int> rev(d.NP - 1);
    for (int i = 0; i < d.NP - 1; i++) rev[i] = d.NP - 2 - i;
    pss.setUser2InternalIndexes(rev);
  }
  pss.structureReady();
}

static void print(const char* name, TerminationReason tc, int its, const Stats& st, SE2PoseSection& pss, int NP) {
  pss.update(true);
  std::printf("{\"solve\": \"%s\", \"reason\": \"%s\", \"iterations\": %d, \"chi2_after\": %.17g, \"log\": [", name,
              toString(tc).c_str(), its, pss.hessian().chi2());
  for (size_t i = 0; i < st.log.size(); i++)
    std::printf("%s[%d, %d, %.17g, %.17g]", i ? ", " : "", st.log[i].it, st.log[i].subit, st.log[i].chi2,
                st.log[i].mu);
  std::printf("], \"poses\": [");
  for (int i = 0; i < NP - 1; i++) {
    const SE2Pose& p = pss.parameter(i);
    std::printf("%s%.17g, %.17g, %.17g", i ? ", " : "", p.th, p.tx, p.ty);
  }
  std::printf("]}\n");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s problem.txt\n", argv[0]);
    return 2;
  }
  try {
    const Data d = read(argv[1]);
    const std::vector<SE2Pose> vars(d.start.begin() + 1, d.start.end());
    const char* order_name[3] = {"", "_amd", "_reversed"};
    for (int order = 0; order < 3; order++) {
      for (int damped = 0; damped < 2; damped++) {
        SE2PoseSection pss(vars, d.start[0]);
        build(pss, d, order);
        GaussNewton<SolverCholeskySchur> gn;
        if (damped) gn.settings().relLambda = 0.01;
        const TerminationReason tc = gn.solve(pss);
        const std::string name = std::string(damped ? "gn_damped" : "gn") + order_name[order];
        print(name.c_str(), tc, gn.numIterations(), gn.stats(), pss, d.NP);
      }
      SE2PoseSection pss(vars, d.start[0]);
      build(pss, d, order);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      const TerminationReason tc = lm.solve(pss);
      const std::string name = std::string("lm") + order_name[order];
      print(name.c_str(), tc, lm.numIterations(), lm.stats(), pss, d.NP);
    }
    for (int kind = 1; kind < 3; kind++)
      for (int order = 0; order < 2; order++) {
        SE2PoseSection pss(vars, d.start[0]);
        build(pss, d, order, kind);
        LevenbergMarquardt<SolverCholeskySchur> lm;
        const TerminationReason tc = lm.solve(pss);
        const std::string name = std::string(kind == 1 ? "lm_nary" : "lm_nary_dynamic") + order_name[order];
        print(name.c_str(), tc, lm.numIterations(), lm.stats(), pss, d.NP);
        SE2PoseSection pg(vars, d.start[0]);
        build(pg, d, order, kind);
        GaussNewton<SolverCholeskySchur> gn;
        const TerminationReason tg = gn.solve(pg);
        const std::string gname = std::string(kind == 1 ? "gn_nary" : "gn_nary_dynamic") + order_name[order];
        print(gname.c_str(), tg, gn.numIterations(), gn.stats(), pg, d.NP);
      }
  } catch (const Error& e) {
    std::fprintf(stderr, "nelson_amd error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
