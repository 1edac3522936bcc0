// Host structure stage (see structure.h).
#include "structure.h"

#include <algorithm>
#include <numeric>

#include "internal.h"

namespace nsn {

int host_structure(const nsn_problem_desc* d, HostStructure& s) {
  const int C = d->n_cameras, P = d->n_points, E = d->n_edges;
  s.C = C;
  s.P = P;
  s.E = E;
  s.cam_var.assign(C, -1);
  s.pt_var.assign(P, -1);
  s.var_cam_user.clear();
  s.var_pt_user.clear();
  for (int c = 0; c < C; c++)
    if (!(d->cam_fixed && d->cam_fixed[c])) {
      s.cam_var[c] = (int)s.var_cam_user.size();
      s.var_cam_user.push_back(c);
    }
  for (int p = 0; p < P; p++)
    if (!(d->pt_fixed && d->pt_fixed[p])) {
      s.pt_var[p] = (int)s.var_pt_user.size();
      s.var_pt_user.push_back(p);
    }
  s.NU = (int)s.var_cam_user.size();
  s.NV = (int)s.var_pt_user.size();
  const int NU = s.NU, NV = s.NV;
  for (int e = 0; e < E; e++) {
    if (d->cam_idx[e] < 0 || d->cam_idx[e] >= C || d->pt_idx[e] < 0 || d->pt_idx[e] >= P)
      return fail(NSN_EINVAL, "edge %d references camera %d / point %d out of range", e, d->cam_idx[e], d->pt_idx[e]);
  }
  // point-major order: (point, camera, edge index)
  s.pe.resize(E);
  std::iota(s.pe.begin(), s.pe.end(), 0);
  std::stable_sort(s.pe.begin(), s.pe.end(), [&](int a, int b) {
    if (d->pt_idx[a] != d->pt_idx[b]) return d->pt_idx[a] < d->pt_idx[b];
    return d->cam_idx[a] < d->cam_idx[b];
  });
  s.pt_ptr.assign(P + 1, 0);
  s.pe_cam.assign(E, 0);
  s.pe_wblk.assign(E, -1);
  s.pe_obs.assign(2 * (size_t)E, 0.0);
  s.pe_w.assign(E, 1.0);
  for (int e = 0; e < E; e++) s.pt_ptr[d->pt_idx[e] + 1]++;
  for (int p = 0; p < P; p++) s.pt_ptr[p + 1] += s.pt_ptr[p];
  s.wptr.assign(NV + 1, 0);
  s.w_cam.clear();
  s.w_pt.clear();
  for (int k = 0; k < E; k++) {
    const int e = s.pe[k];
    s.pe_cam[k] = d->cam_idx[e];
    s.pe_obs[2 * (size_t)k] = d->obs[2 * (size_t)e];
    s.pe_obs[2 * (size_t)k + 1] = d->obs[2 * (size_t)e + 1];
    s.pe_w[k] = d->weight ? d->weight[e] : 1.0;
    const int cv = s.cam_var[d->cam_idx[e]], pv = s.pt_var[d->pt_idx[e]];
    if (cv >= 0 && pv >= 0) {
      // repeated observations of one (camera, point) pair are adjacent and
      // share one W block
      const bool same = !s.w_cam.empty() && s.w_pt.back() == pv && s.w_cam.back() == cv;
      if (!same) {
        s.w_cam.push_back(cv);
        s.w_pt.push_back(pv);
        s.wptr[pv + 1]++;
      }
      s.pe_wblk[k] = (int)s.w_cam.size() - 1;
    }
  }
  for (int v = 0; v < NV; v++) s.wptr[v + 1] += s.wptr[v];
  s.nW = (int)s.w_cam.size();
  const int nW = s.nW;
  // camera-major edges of the variable cameras: (camera, point, edge index)
  s.ce.clear();
  s.ce.reserve(E);
  for (int e = 0; e < E; e++)
    if (s.cam_var[d->cam_idx[e]] >= 0) s.ce.push_back(e);
  std::stable_sort(s.ce.begin(), s.ce.end(), [&](int a, int b) {
    if (d->cam_idx[a] != d->cam_idx[b]) return d->cam_idx[a] < d->cam_idx[b];
    return d->pt_idx[a] < d->pt_idx[b];
  });
  s.Ec = (int)s.ce.size();
  s.cam_ptr.assign(NU + 1, 0);
  s.ce_pt.assign(s.Ec, 0);
  s.ce_obs.assign(2 * (size_t)s.Ec, 0.0);
  s.ce_w.assign(s.Ec, 1.0);
  for (int k = 0; k < s.Ec; k++) {
    const int e = s.ce[k];
    s.cam_ptr[s.cam_var[d->cam_idx[e]] + 1]++;
    s.ce_pt[k] = d->pt_idx[e];
    s.ce_obs[2 * (size_t)k] = d->obs[2 * (size_t)e];
    s.ce_obs[2 * (size_t)k + 1] = d->obs[2 * (size_t)e + 1];
    s.ce_w[k] = d->weight ? d->weight[e] : 1.0;
  }
  for (int u = 0; u < NU; u++) s.cam_ptr[u + 1] += s.cam_ptr[u];
  // W blocks grouped by camera (ascending point within a camera)
  s.cw_ptr.assign(NU + 1, 0);
  s.cw_idx.assign(nW, 0);
  for (int k = 0; k < nW; k++) s.cw_ptr[s.w_cam[k] + 1]++;
  for (int u = 0; u < NU; u++) s.cw_ptr[u + 1] += s.cw_ptr[u];
  {
    std::vector<int> fill(s.cw_ptr.begin(), s.cw_ptr.end() - 1);
    for (int k = 0; k < nW; k++) s.cw_idx[fill[s.w_cam[k]]++] = k;
  }
  // reference order of W blocks: row-major (camera, point)
  s.w_ref = s.cw_idx;
  // S structure: pairs of blocks sharing a point, grouped by camera pair
  // (MatrixWWtMultiplier::prepare, nelson/MatrixWWtMultiplier.hpp:18-120)
  struct PairRec {
    int ci, cj, ka, kb;
  };
  std::vector<PairRec> pr;
  for (int v = 0; v < NV; v++)
    for (int a = s.wptr[v]; a < s.wptr[v + 1]; a++)
      for (int b = a + 1; b < s.wptr[v + 1]; b++) pr.push_back({s.w_cam[a], s.w_cam[b], a, b});
  std::stable_sort(pr.begin(), pr.end(),
                   [](const PairRec& x, const PairRec& y) { return x.ci != y.ci ? x.ci < y.ci : x.cj < y.cj; });
  s.npairs = (long long)pr.size();
  s.sb_i.clear();
  s.sb_j.clear();
  s.sb_ptr.clear();
  s.sb_pairs.assign(2 * pr.size(), 0);
  size_t q = 0;
  for (int u = 0; u < NU; u++) {
    s.sb_i.push_back(u);
    s.sb_j.push_back(u);
    s.sb_ptr.push_back((int)q);
    while (q < pr.size() && pr[q].ci == u) {
      const int cj = pr[q].cj;
      s.sb_i.push_back(u);
      s.sb_j.push_back(cj);
      s.sb_ptr.push_back((int)q);
      while (q < pr.size() && pr[q].ci == u && pr[q].cj == cj) {
        s.sb_pairs[2 * q] = pr[q].ka;
        s.sb_pairs[2 * q + 1] = pr[q].kb;
        q++;
      }
    }
  }
  s.sb_ptr.push_back((int)q);
  s.nsb = (int)s.sb_i.size();
  return NSN_OK;
}

}  // namespace nsn
