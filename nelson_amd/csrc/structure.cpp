// Host structure stage (see structure.h).
#include "structure.h"

#include <algorithm>
#include <functional>
#include <numeric>

#include "internal.h"

namespace nsn {

std::vector<int> rcm_order(const std::vector<std::vector<int>>& adj) {
  const int n = (int)adj.size();
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<int> level(n, -1);
  auto deg = [&](int v) { return (int)adj[v].size(); };
  // breadth-first levels from s inside its component (over unplaced nodes);
  // returns the last level's node of least degree and the eccentricity
  auto far_node = [&](int st, int& ecc) {
    std::vector<int> q{st};
    std::vector<int> touched{st};
    level[st] = 0;
    for (size_t h = 0; h < q.size(); h++)
      for (int w : adj[q[h]])
        if (!seen[w] && level[w] < 0) {
          level[w] = level[q[h]] + 1;
          q.push_back(w);
          touched.push_back(w);
        }
    ecc = level[q.back()];
    int best = q.back();
    for (int v : q)
      if (level[v] == ecc && deg(v) < deg(best)) best = v;
    for (int v : touched) level[v] = -1;
    return best;
  };
  std::vector<int> by_deg(n);
  for (int i = 0; i < n; i++) by_deg[i] = i;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg(a) < deg(b); });
  for (int cand : by_deg) {
    if (seen[cand]) continue;
    // pseudo-peripheral start: move to a far node while the eccentricity grows
    int st = cand, ecc = 0;
    int far = far_node(st, ecc);
    for (int rep = 0; rep < 4; rep++) {
      int e2 = 0;
      const int f2 = far_node(far, e2);
      if (e2 <= ecc) break;
      st = far;
      ecc = e2;
      far = f2;
    }
    const size_t base = order.size();
    order.push_back(st);
    seen[st] = 1;
    for (size_t h = base; h < order.size(); h++) {
      std::vector<int> nb;
      for (int w : adj[order[h]])
        if (!seen[w]) nb.push_back(w);
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg(a) < deg(b); });
      for (int w : nb) {
        seen[w] = 1;
        order.push_back(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

std::vector<int> nd_order(const std::vector<std::vector<int>>& adj, int align, int depth) {
  const int n = (int)adj.size();
  std::vector<int> out;
  out.reserve(n);
  std::vector<int> loc(n, -1);  // node -> position in the current part's order
  if (align < 1) align = 1;
  // reverse Cuthill-McKee order of the subgraph induced by `nodes`
  auto sub_rcm = [&](const std::vector<int>& nodes) {
    const int N = (int)nodes.size();
    for (int a = 0; a < N; a++) loc[nodes[a]] = a;
    std::vector<std::vector<int>> sa(N);
    for (int a = 0; a < N; a++)
      for (int w : adj[nodes[a]])
        if (loc[w] >= 0) sa[a].push_back(loc[w]);
    for (int v : nodes) loc[v] = -1;
    std::vector<int> o = rcm_order(sa);
    for (int& v : o) v = nodes[v];
    return o;
  };
  std::vector<char> inpart(n, 0);
  std::function<void(const std::vector<int>&, int)> part = [&](const std::vector<int>& nodes, int lev) {
    const int N = (int)nodes.size();
    std::vector<int> ord = sub_rcm(nodes);
    {
      // orientation: the part's nodes adjacent to nodes outside it (the
      // separators, which come later) go to the end of its order, so that
      // the separator rows of S start late in the part (the envelope of a
      // row runs from its first non-zero block to the diagonal)
      for (int v : nodes) inpart[v] = 1;
      double pos = 0.0;
      int nbd = 0;
      for (int a = 0; a < N; a++)
        for (int w : adj[ord[a]])
          if (!inpart[w]) {
            pos += a;
            nbd++;
            break;
          }
      for (int v : nodes) inpart[v] = 0;
      if (nbd > 0 && pos / nbd < 0.5 * (N - 1)) std::reverse(ord.begin(), ord.end());
    }
    if (lev >= depth || N < 4 * align) {
      out.insert(out.end(), ord.begin(), ord.end());
      return;
    }
    // sweep the prefixes A = ord[0, m): S = the nodes outside A with a
    // neighbour in A (cnt[v] = neighbours of v in A), B = the rest
    for (int a = 0; a < N; a++) loc[ord[a]] = a;
    std::vector<int> cnt(N, 0);
    std::vector<char> inA(N, 0);
    int sizeS = 0, best_m = -1;
    long long best = N;
    for (int m = 1; m <= N; m++) {
      const int u = m - 1;
      if (cnt[u] > 0) sizeS--;
      inA[u] = 1;
      for (int w : adj[ord[u]]) {
        const int lw = loc[w];
        if (lw < 0 || inA[lw]) continue;
        if (cnt[lw]++ == 0) sizeS++;
      }
      if (m % align) continue;
      const int nb = N - m - sizeS;
      if (nb < align) break;
      const long long cost = std::max(m, nb) + sizeS;
      if (cost < best) {
        best = cost;
        best_m = m;
      }
    }
    std::vector<int> A, B, S;
    if (best_m > 0 && best * 100 <= 85LL * N) {
      std::vector<char> inS(N, 0);
      for (int a = 0; a < best_m; a++)
        for (int w : adj[ord[a]]) {
          const int lw = loc[w];
          if (lw >= best_m) inS[lw] = 1;
        }
      for (int a = 0; a < N; a++) (a < best_m ? A : inS[a] ? S : B).push_back(ord[a]);
    }
    for (int v : nodes) loc[v] = -1;
    if (A.empty()) {
      out.insert(out.end(), ord.begin(), ord.end());
      return;
    }
    part(A, lev + 1);
    part(B, lev + 1);
    out.insert(out.end(), S.begin(), S.end());
  };
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  part(all, 0);
  return out;
}

int number_variables(const nsn_problem_desc* d, HostStructure& s) {
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
  const int NU = s.NU;
  for (int e = 0; e < E; e++) {
    if (d->cam_idx[e] < 0 || d->cam_idx[e] >= C || d->pt_idx[e] < 0 || d->pt_idx[e] >= P)
      return fail(NSN_EINVAL, "edge %d references camera %d / point %d out of range", e, d->cam_idx[e], d->pt_idx[e]);
  }
  s.var_cam_api.resize(NU);
  for (int u = 0; u < NU; u++) s.var_cam_api[u] = u;
  if (d->camera_order != NSN_ORDER_AS_GIVEN && d->camera_order != NSN_ORDER_RCM && d->camera_order != NSN_ORDER_ND)
    return fail(NSN_EINVAL, "unknown camera_order %d", d->camera_order);
  return NSN_OK;
}

bool wants_camera_order(const nsn_problem_desc* d, const HostStructure& s) {
  return (d->camera_order == NSN_ORDER_RCM || d->camera_order == NSN_ORDER_ND) && s.NU > 2;
}

void apply_camera_order(const nsn_problem_desc* d, HostStructure& s, const std::vector<std::vector<int>>& adj) {
  // solver position -> interface index; nested dissection in parts of
  // multiples of 32 cameras (192 rows: three 64-row tiles)
  const std::vector<int> ord = d->camera_order == NSN_ORDER_ND ? nd_order(adj, kNdAlign, kNdDepth) : rcm_order(adj);
  const std::vector<int> api_user = s.var_cam_user;
  for (int v = 0; v < s.NU; v++) {
    s.var_cam_api[v] = ord[v];
    s.var_cam_user[v] = api_user[ord[v]];
    s.cam_var[api_user[ord[v]]] = v;
  }
}

std::vector<std::vector<int>> camera_graph_from_blocks(const HostStructure& s) {
  // the off-diagonal S blocks (i, j), i < j, in row-major order: adj[x]
  // receives its lower neighbours in ascending order before its upper ones
  std::vector<std::vector<int>> adj(s.NU);
  for (int b = 0; b < s.nsb; b++)
    if (s.sb_i[b] != s.sb_j[b]) {
      adj[s.sb_i[b]].push_back(s.sb_j[b]);
      adj[s.sb_j[b]].push_back(s.sb_i[b]);
    }
  return adj;
}

void w_ref_order(HostStructure& s) {
  // reference order of W blocks: row-major (camera, point), cameras in
  // interface order
  std::vector<int> api_var(s.NU);
  for (int u = 0; u < s.NU; u++) api_var[s.var_cam_api[u]] = u;
  s.w_ref.clear();
  s.w_ref.reserve(s.nW);
  for (int a = 0; a < s.NU; a++) {
    const int u = api_var[a];
    s.w_ref.insert(s.w_ref.end(), s.cw_idx.begin() + s.cw_ptr[u], s.cw_idx.begin() + s.cw_ptr[u + 1]);
  }
}

int host_structure(const nsn_problem_desc* d, HostStructure& s, bool allow_order) {
  const int rc = number_variables(d, s);
  if (rc) return rc;
  const int P = d->n_points, E = d->n_edges, NU = s.NU, NV = s.NV;
  if (allow_order && wants_camera_order(d, s)) {
    // the camera graph: variable cameras sharing a variable point
    std::vector<std::vector<int>> cams_of(P), adj(NU);
    for (int e = 0; e < E; e++) {
      const int cv = s.cam_var[d->cam_idx[e]];
      if (cv >= 0 && s.pt_var[d->pt_idx[e]] >= 0) cams_of[d->pt_idx[e]].push_back(cv);
    }
    for (auto& l : cams_of) {
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
      for (size_t a = 0; a < l.size(); a++)
        for (size_t b = a + 1; b < l.size(); b++) {
          adj[l[a]].push_back(l[b]);
          adj[l[b]].push_back(l[a]);
        }
      std::vector<int>().swap(l);
    }
    for (auto& l : adj) {
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
    }
    apply_camera_order(d, s, adj);
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
  s.w_dup = false;
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
      if (same) s.w_dup = true;
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
  // grouped by variable camera in solver order
  std::stable_sort(s.ce.begin(), s.ce.end(), [&](int a, int b) {
    const int ca = s.cam_var[d->cam_idx[a]], cb = s.cam_var[d->cam_idx[b]];
    if (ca != cb) return ca < cb;
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
  w_ref_order(s);
  // S structure: pairs of blocks sharing a point, grouped by camera pair
  // (MatrixWWtMultiplier::prepare, nelson/MatrixWWtMultiplier.hpp:18-120)
  struct PairRec {
    int ci, cj, ka, kb;
  };
  std::vector<PairRec> pr;
  for (int v = 0; v < NV; v++)
    for (int a = s.wptr[v]; a < s.wptr[v + 1]; a++)
      for (int b = a + 1; b < s.wptr[v + 1]; b++) {
        // the lower camera (solver order) first: the block of S it gives is
        // (ci, cj), ci < cj, the sum of Z_a Z_b^T
        if (s.w_cam[a] < s.w_cam[b]) pr.push_back({s.w_cam[a], s.w_cam[b], a, b});
        else pr.push_back({s.w_cam[b], s.w_cam[a], b, a});
      }
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
