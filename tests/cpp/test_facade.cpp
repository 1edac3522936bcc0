// The C++ interface (nelson_amd.hpp) used the way the reference's tests use
// DoubleSection / SolverCholeskySchur / LevenbergMarquardt / GaussNewton
// (nelson/tests/testDoubleSectionCommon.hpp:55-130): Config 1's BA problem
// (10 cameras, 500 points, 2000 observations) built through parameterU /
// parameterV / addEdge / structureReady, one Schur step, an LM solve and a
// GN solve, then a step and an LM solve with the cameras ordered by
// permuteRCM_SchurU, and a step of a 300-camera ring problem with the cameras
// in nested-dissection order (permuteND_SchurU: independent diagonal chains).
// Prints one JSON object per solve on stdout; the GPU test compares them with
// the oracle.
#include <cmath>
#include <cstdio>
#include <vector>

#include "nelson_amd.hpp"

using namespace nelson_amd;

static void build(BASection& pss, const std::vector<double>& cams, const std::vector<double>& pts,
                  const std::vector<int>& ci, const std::vector<int>& pi, const std::vector<double>& obs, int C,
                  int P, bool priors = false, bool rcm = false, bool nd = false) {
  pss.setNumParameters(C, P);
  for (int i = 0; i < C; i++)
    for (int k = 0; k < 3; k++) {
      pss.parameterU(i).w[k] = cams[6 * i + k];
      pss.parameterU(i).t[k] = cams[6 * i + 3 + k];
    }
  for (int j = 0; j < P; j++)
    for (int k = 0; k < 3; k++) pss.parameterV(j)[k] = pts[3 * j + k];
  pss.setFixedU(0);  // gauge of the BA sample (BASchurPerfomances.cpp:245-246)
  pss.setFixedV(0);
  pss.parametersReady();
  if (rcm) pss.permuteRCM_SchurU();
  if (nd) pss.permuteND_SchurU();
  for (size_t e = 0; e < ci.size(); e++) pss.addEdge(ci[e], pi[e], obs[2 * e], obs[2 * e + 1]);
  if (priors) {
    for (int c : {0, 2, 5, 7}) {
      BASection::Camera pr;
      std::array<double, 6> w;
      for (int k = 0; k < 6; k++) {
        const double off = 0.01 * (k % 3 + 1) * (k < 3 ? 1.0 : 5.0);
        (k < 3 ? pr.w[k] : pr.t[k - 3]) = cams[6 * c + k] + off;
        w[k] = 0.5 + 0.25 * ((k + c) % 4);
      }
      pss.addEdgeUnaryU(c, pr, w);
    }
    for (int q : {0, 3, 17, 250, 499}) {
      std::array<double, 3> pr, w;
      for (int k = 0; k < 3; k++) {
        pr[k] = pts[3 * q + k] + 0.01 * (k % 3 + 1);
        w[k] = 0.5 + 0.25 * ((k + q) % 4);
      }
      pss.addEdgeUnaryV(q, pr, w);
    }
  }
  pss.structureReady();
}

static void print_log(const char* name, TerminationReason tc, int its, double chi2, const Stats& st) {
  std::printf("{\"solve\": \"%s\", \"reason\": \"%s\", \"iterations\": %d, \"final_chi2\": %.17g, \"log\": [", name,
              toString(tc).c_str(), its, chi2);
  for (size_t i = 0; i < st.log.size(); i++)
    std::printf("%s[%d, %d, %.17g, %.17g]", i ? ", " : "", st.log[i].it, st.log[i].subit, st.log[i].chi2,
                st.log[i].mu);
  std::printf("]}\n");
}

int main() {
  nsn_gen_params g{};
  g.n_cameras = 10; g.n_points = 500; g.k_min = 4; g.k_max = 4; g.layout = 0; g.arc = 0; g.seed = 1;
  g.obs_sigma = 1.0; g.outlier_frac = 0.0; g.outlier_sigma = 0.0; g.rot_sigma = 0.01; g.trans_sigma = 0.05;
  g.point_sigma = 0.1; g.focal = 500.0; g.cx = 320.0; g.cy = 240.0;
  const int C = g.n_cameras, P = g.n_points, cap = P * g.k_max;
  std::vector<double> cgt(6 * C), pgt(3 * P), c0(6 * C), p0(3 * P), obs(2 * cap);
  std::vector<int> ci(cap), pi(cap);
  int ne = 0;
  check(nsn_generate_ba(&g, cap, cgt.data(), pgt.data(), c0.data(), p0.data(), ci.data(), pi.data(), obs.data(), &ne),
        "nsn_generate_ba");
  ci.resize(ne);
  pi.resize(ne);
  obs.resize(2 * ne);
  try {
    {
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P);
      pss.update(true);
      SolverCholeskySchur solver;
      solver.init(pss.hessian().H(), pss.hessian().b());
      const bool ok = solver.computeIncrement(pss.hessian().H(), pss.hessian().b(), 0.0, 62.0);
      std::printf("{\"solve\": \"step\", \"ok\": %s, \"sq_norm\": %.17g, \"max_abs_h_diag\": %.17g, \"du0\": %.17g}\n",
                  ok ? "true" : "false", solver.incrementVectorSquaredNorm(), solver.maxAbsHDiag(),
                  solver.incrementVector().bU()[0]);
    }
    {
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      lm.settings().maxNumIt = 20;
      const TerminationReason tc = lm.solve(pss);
      print_log("lm", tc, lm.numIterations(), pss.update(false), lm.stats());
    }
    {
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P);
      GaussNewton<SolverCholeskySchur> gn;
      gn.settings().maxNumIt = 5;
      gn.settings().epsIncSquare = 0.0;
      const TerminationReason tc = gn.solve(pss);
      print_log("gn", tc, gn.numIterations(), pss.chi2(), gn.stats());
    }
    {
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P, true);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      lm.settings().maxNumIt = 20;
      const TerminationReason tc = lm.solve(pss);
      print_log("lm_priors", tc, lm.numIterations(), pss.update(false), lm.stats());
    }
    {
      // the cameras ordered by reverse Cuthill-McKee inside the solver: the
      // increment stays in camera order, so the same values as "step"
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P, false, true);
      pss.update(true);
      SolverCholeskySchur solver;
      solver.init(pss.hessian().H(), pss.hessian().b());
      const bool ok = solver.computeIncrement(pss.hessian().H(), pss.hessian().b(), 0.0, 62.0);
      std::printf("{\"solve\": \"step_rcm\", \"ok\": %s, \"sq_norm\": %.17g, \"du0\": %.17g, \"du\": [",
                  ok ? "true" : "false", solver.incrementVectorSquaredNorm(), solver.incrementVector().bU()[0]);
      for (int i = 0; i < C - 1; i++)
        for (int k = 0; k < 6; k++)
          std::printf("%s%.17g", i + k ? ", " : "", solver.incrementVector().bU().segment(i)[k]);
      std::printf("]}\n");
    }
    {
      BASection pss(g.focal, g.cx, g.cy);
      build(pss, c0, p0, ci, pi, obs, C, P, false, true);
      LevenbergMarquardt<SolverCholeskySchur> lm;
      lm.settings().maxNumIt = 20;
      const TerminationReason tc = lm.solve(pss);
      print_log("lm_rcm", tc, lm.numIterations(), pss.update(false), lm.stats());
    }
    {
      // a 300-camera ring (each point seen by 4-8 cameras of a 40-camera arc)
      // in nested-dissection order: one step, the increment in camera order
      nsn_gen_params r = g;
      r.n_cameras = 300; r.n_points = 30000; r.k_min = 4; r.k_max = 8; r.layout = 1; r.arc = 40; r.seed = 31;
      const int RC = r.n_cameras, RP = r.n_points, rcap = RP * r.k_max;
      std::vector<double> rcgt(6 * RC), rpgt(3 * RP), rc0(6 * RC), rp0(3 * RP), robs(2 * rcap);
      std::vector<int> rci(rcap), rpi(rcap);
      int rne = 0;
      check(nsn_generate_ba(&r, rcap, rcgt.data(), rpgt.data(), rc0.data(), rp0.data(), rci.data(), rpi.data(),
                            robs.data(), &rne),
            "nsn_generate_ba");
      rci.resize(rne);
      rpi.resize(rne);
      robs.resize(2 * rne);
      BASection pss(r.focal, r.cx, r.cy);
      build(pss, rc0, rp0, rci, rpi, robs, RC, RP, false, false, true);
      pss.update(true);
      SolverCholeskySchur solver;
      solver.init(pss.hessian().H(), pss.hessian().b());
      const double mu = 1e-4 * solver.maxAbsHDiag();
      const bool ok = solver.computeIncrement(pss.hessian().H(), pss.hessian().b(), 0.0, mu);
      std::printf("{\"solve\": \"step_nd\", \"ok\": %s, \"mu\": %.17g, \"du\": [", ok ? "true" : "false", mu);
      for (int i = 0; i < RC - 1; i++)
        for (int k = 0; k < 6; k++)
          std::printf("%s%.17g", i + k ? ", " : "", solver.incrementVector().bU().segment(i)[k]);
      std::printf("]}\n");
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "nelson_amd error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
