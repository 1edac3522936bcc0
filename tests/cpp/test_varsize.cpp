// A user section whose parameters have their own sizes (the reference's
// mat::Variable block sizes: parameterUSize() / parameterVSize() return one
// size per parameter, nelson/DoubleSection.h:76-110, as its tests
// testDoubleSectionSchurDiagBlockV*.cpp do): U parameters of 1 to 3
// coordinates, V parameters of 1 to 3, unary edges x - gt on every
// parameter, difference edges on the leading coordinates two parameters
// share (U-U between neighbours, U-V on a pattern of pairs, V-V between
// neighbours when asked).  The measurements are exact for the ground truth,
// so the optimum has chi2 = 0.
//
// Reads argv[1]:
//   NU NV VV
//   NU sizes, NV sizes
//   NU lines gt (size values), NV lines gt, NU lines start, NV lines start
// and prints one JSON object per solve: Gauss-Newton, Levenberg-Marquardt,
// and one damped Schur step (rel 0.05, abs 0.3) at the start.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <vector>

#include "nelson_amd.hpp"

using namespace nelson_amd;

struct Par {
  double p[3] = {0, 0, 0};
};

struct Data {
  int NU = 0, NV = 0, vv = 0;
  std::vector<int> su, sv;
  std::vector<Par> gtU, gtV, pU, pV;
};

template <class Section>
class UnaryU : public EdgeUnaryU<Section, UnaryU<Section>> {
  Par m_;
  int n_;
  double e_[3] = {0, 0, 0};

 public:
  UnaryU(const Par& m, int n) : m_(m), n_(n) {}
  void update(bool) override {
    double c = 0;
    for (int k = 0; k < n_; k++) e_[k] = this->parameter().p[k] - m_.p[k], c += e_[k] * e_[k];
    this->setChi2(c);
  }
  void updateHBlock(MatRef H, VecRef b) {
    H.addIdentity();
    for (int k = 0; k < n_; k++) b(k) += e_[k];
  }
};

template <class Section>
class UnaryV : public EdgeUnaryV<Section, UnaryV<Section>> {
  Par m_;
  int n_;
  double e_[3] = {0, 0, 0};

 public:
  UnaryV(const Par& m, int n) : m_(m), n_(n) {}
  void update(bool) override {
    double c = 0;
    for (int k = 0; k < n_; k++) e_[k] = this->parameter().p[k] - m_.p[k], c += e_[k] * e_[k];
    this->setChi2(c);
  }
  void updateHBlock(MatRef H, VecRef b) {
    H.addIdentity();
    for (int k = 0; k < n_; k++) b(k) += e_[k];
  }
};

// (a - b) - meas on the first m coordinates of two parameters; the block
// views have the parameters' own sizes, so only their leading m x m parts
// are touched
template <class Base>
class Diff : public Base {
  double meas_[3];
  int m_;
  double e_[3] = {0, 0, 0};

 public:
  Diff(const double* meas, int m) : meas_{meas[0], meas[1], meas[2]}, m_(m) {}
  void update(bool) override {
    double c = 0;
    for (int k = 0; k < m_; k++)
      e_[k] = (this->parameter_1().p[k] - this->parameter_2().p[k]) - meas_[k], c += e_[k] * e_[k];
    this->setChi2(c);
  }
  void updateH11Block(MatRef H, VecRef v) {
    for (int k = 0; k < m_; k++) H(k, k) += 1.0, v(k) += e_[k];
  }
  void updateH22Block(MatRef H, VecRef v) {
    for (int k = 0; k < m_; k++) H(k, k) += 1.0, v(k) -= e_[k];
  }
  void updateH12Block(MatRef H) {
    for (int k = 0; k < m_; k++) H(k, k) -= 1.0;
  }
};

class VarSection : public DoubleSection<VarSection, Par, Par, 3, 3> {
 public:
  explicit VarSection(const Data& d) : DoubleSection(d.NU, d.NV), d_(d), u_(d.pU), v_(d.pV) {}
  const std::vector<int>& parameterUSize() const override { return d_.su; }
  const std::vector<int>& parameterVSize() const override { return d_.sv; }
  Par& parameterU(NodeId i) { return i.isVariable() ? u_.at(i.id()) : fixed_; }
  const Par& parameterU(NodeId i) const { return i.isVariable() ? u_.at(i.id()) : fixed_; }
  Par& parameterV(NodeId i) { return i.isVariable() ? v_.at(i.id()) : fixed_; }
  const Par& parameterV(NodeId i) const { return i.isVariable() ? v_.at(i.id()) : fixed_; }
  void oplus(const HessianVectors& inc) {
    for (int i = 0; i < d_.NU; i++) {
      const ConstSeg x = inc.bU().segment(user2internalIndexesU()[i]);
      for (int k = 0; k < x.size(); k++) u_[i].p[k] += x(k);
    }
    for (int j = 0; j < d_.NV; j++) {
      const ConstSeg x = inc.bV().segment(j);
      for (int k = 0; k < x.size(); k++) v_[j].p[k] += x(k);
    }
  }
  void backupSolution() { bu_ = u_, bv_ = v_; }
  void rollbackSolution() { u_ = bu_, v_ = bv_; }
  const std::vector<Par>& u() const { return u_; }
  const std::vector<Par>& v() const { return v_; }

 private:
  const Data& d_;
  std::vector<Par> u_, v_, bu_, bv_;
  Par fixed_;
};

class EdgeUU : public Diff<EdgeBinaryUU<VarSection, EdgeUU>> {
 public:
  using Diff::Diff;
};
class EdgeUV : public Diff<EdgeBinaryUV<VarSection, EdgeUV>> {
 public:
  using Diff::Diff;
};
class EdgeVV : public Diff<EdgeBinaryVV<VarSection, EdgeVV>> {
 public:
  using Diff::Diff;
};

// the U parameters alone as a single section (nelson/SingleSection.h with
// mat::Variable sizes: parameterSize()), unary and U-U edges
class VarSingle : public SingleSection<VarSingle, Par, 3> {
 public:
  explicit VarSingle(const Data& d) : SingleSection(d.NU), d_(d), u_(d.pU) {}
  const std::vector<int>& parameterSize() const override { return d_.su; }
  Par& parameter(NodeId i) { return i.isVariable() ? u_.at(i.id()) : fixed_; }
  const Par& parameter(NodeId i) const { return i.isVariable() ? u_.at(i.id()) : fixed_; }
  void oplus(const HessianVectors& inc) {
    for (int i = 0; i < d_.NU; i++) {
      const ConstSeg x = inc.segment(user2internalIndexes()[i]);
      for (int k = 0; k < x.size(); k++) u_[i].p[k] += x(k);
    }
  }
  void backupSolution() { b_ = u_; }
  void rollbackSolution() { u_ = b_; }
  const std::vector<Par>& u() const { return u_; }

 private:
  const Data& d_;
  std::vector<Par> u_, b_;
  Par fixed_;
};

class SingleUU : public Diff<EdgeBinaryUU<VarSingle, SingleUU>> {
 public:
  using Diff::Diff;
};

static Data read(const char* path) {
  std::ifstream f(path);
  Data d;
  f >> d.NU >> d.NV >> d.vv;
  d.su.resize(d.NU), d.sv.resize(d.NV);
  for (int& x : d.su) f >> x;
  for (int& x : d.sv) f >> x;
  auto rd = [&](std::vector<Par>& v, const std::vector<int>& sz) {
    v.resize(sz.size());
    for (size_t i = 0; i < sz.size(); i++)
      for (int k = 0; k < sz[i]; k++) f >> v[i].p[k];
  };
  rd(d.gtU, d.su), rd(d.gtV, d.sv), rd(d.pU, d.su), rd(d.pV, d.sv);
  if (!f) throw Error(NSN_EINVAL, std::string("cannot read ") + path);
  return d;
}

static void build(VarSection& s, const Data& d) {
  s.parametersReady();
  for (int i = 0; i < d.NU; i++) s.addEdge(i, new UnaryU<VarSection>(d.gtU[i], d.su[i]));
  for (int j = 0; j < d.NV; j++) s.addEdge(j, new UnaryV<VarSection>(d.gtV[j], d.sv[j]));
  auto meas = [](const Par& a, const Par& b, int m, double* out) {
    for (int k = 0; k < 3; k++) out[k] = k < m ? a.p[k] - b.p[k] : 0.0;
  };
  double m3[3];
  for (int i = 0; i + 1 < d.NU; i++) {
    const int m = std::min(d.su[i], d.su[i + 1]);
    meas(d.gtU[i], d.gtU[i + 1], m, m3);
    s.addEdge(i, i + 1, new EdgeUU(m3, m));
  }
  if (d.vv)
    for (int j = 0; j + 1 < d.NV; j++) {
      const int m = std::min(d.sv[j], d.sv[j + 1]);
      meas(d.gtV[j], d.gtV[j + 1], m, m3);
      s.addEdge(j, j + 1, new EdgeVV(m3, m));
    }
  for (int i = 0; i < d.NU; i++)
    for (int j = 0; j < d.NV; j++) {
      if ((i + j) % 2) continue;
      const int m = std::min(d.su[i], d.sv[j]);
      meas(d.gtU[i], d.gtV[j], m, m3);
      s.addEdge(i, j, new EdgeUV(m3, m));
    }
  s.structureReady();
}

static void print_solve(const char* name, TerminationReason tc, const Stats& st, const VarSection& s, const Data& d) {
  std::printf("{\"solve\": \"%s\", \"reason\": \"%s\", \"final_chi2\": %.17g, \"log\": [", name, toString(tc).c_str(),
              s.hessian().chi2());
  for (size_t i = 0; i < st.log.size(); i++)
    std::printf("%s[%d, %d, %.17g, %.17g]", i ? ", " : "", st.log[i].it, st.log[i].subit, st.log[i].chi2,
                st.log[i].mu);
  std::printf("], \"x\": [");
  bool first = true;
  for (int i = 0; i < d.NU; i++)
    for (int k = 0; k < d.su[i]; k++) std::printf("%s%.17g", first ? "" : ", ", s.u()[i].p[k]), first = false;
  for (int j = 0; j < d.NV; j++)
    for (int k = 0; k < d.sv[j]; k++) std::printf("%s%.17g", first ? "" : ", ", s.v()[j].p[k]), first = false;
  std::printf("]}\n");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s problem.txt\n", argv[0]);
    return 2;
  }
  try {
    const Data d = read(argv[1]);
    {
      VarSection s(d);
      build(s, d);
      GaussNewton<SolverCholeskySchur> gn;
      const TerminationReason tc = gn.solve(s);
      print_solve("gn", tc, gn.stats(), s, d);
    }
    {
      VarSection s(d);
      build(s, d);
      // the solver named as the reference's tests name it
      // (nelson/tests/testDoubleSectionCommon.hpp:61)
      using Solver = SolverTraits<solverCholeskySchur>::Solver<VarSection::Hessian::Traits, matrixWrapperDense,
                                                                matrixWrapperSparse, solverCholeskyDense,
                                                                choleskyAMDOrdering, choleskyAMDOrdering>;
      LevenbergMarquardt<Solver> lm;
      const TerminationReason tc = lm.solve(s);
      print_solve("lm", tc, lm.stats(), s, d);
    }
    {
      VarSection s(d);
      build(s, d);
      s.update(true);
      SolverCholeskySchur solver;
      solver.init(s.hessian().H(), s.hessian().b());
      const bool ok = solver.computeIncrement(s.hessian().H(), s.hessian().b(), 0.05, 0.3);
      const HessianVectors& inc = solver.incrementVector();
      std::printf("{\"solve\": \"step\", \"ok\": %s, \"chi2\": %.17g, \"sq_norm\": %.17g, \"dx\": [", ok ? "true" : "false",
                  s.hessian().chi2(), solver.incrementVectorSquaredNorm());
      bool first = true;
      for (int i = 0; i < d.NU; i++) {
        const ConstSeg x = inc.bU().segment(s.user2internalIndexesU()[i]);
        for (int k = 0; k < x.size(); k++) std::printf("%s%.17g", first ? "" : ", ", x(k)), first = false;
      }
      for (int j = 0; j < d.NV; j++) {
        const ConstSeg x = inc.bV().segment(j);
        for (int k = 0; k < x.size(); k++) std::printf("%s%.17g", first ? "" : ", ", x(k)), first = false;
      }
      std::printf("], \"u_size\": %zu, \"v_size\": %zu}\n", inc.bU().size(), inc.bV().size());
    }
    {
      // the single section: LM from the start, then its final state
      VarSingle s(d);
      s.parametersReady();
      for (int i = 0; i < d.NU; i++) s.addEdge(i, new UnaryU<VarSingle>(d.gtU[i], d.su[i]));
      double m3[3];
      for (int i = 0; i + 1 < d.NU; i++) {
        const int m = std::min(d.su[i], d.su[i + 1]);
        for (int k = 0; k < 3; k++) m3[k] = k < m ? d.gtU[i].p[k] - d.gtU[i + 1].p[k] : 0.0;
        s.addEdge(i, i + 1, new SingleUU(m3, m));
      }
      s.structureReady();
      LevenbergMarquardt<SolverCholeskySchur> lm;
      const TerminationReason tc = lm.solve(s);
      std::printf("{\"solve\": \"single_lm\", \"reason\": \"%s\", \"final_chi2\": %.17g, \"x\": [",
                  toString(tc).c_str(), s.hessian().chi2());
      bool first = true;
      for (int i = 0; i < d.NU; i++)
        for (int k = 0; k < d.su[i]; k++) std::printf("%s%.17g", first ? "" : ", ", s.u()[i].p[k]), first = false;
      std::printf("]}\n");
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "nelson_amd error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
