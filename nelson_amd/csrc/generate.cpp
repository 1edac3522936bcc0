// Synthetic bundle-adjustment generator (host only).
//
// Plays the role of the sample's BA_Problem constructor + addNoise
// (nelson/samples/BASchurPerfomances.cpp:40-100), with a portable RNG so the
// same seed gives the same problem on every machine: SplitMix64 -> 53-bit
// uniforms -> Box-Muller normals (cosine branch only).  Draw order: camera
// poses, points, observations (per point: k, cameras, noise, outlier), then
// initial-state perturbations (cameras, then points).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nelson_amd.h"
#include "internal.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = 1.0 - uniform();
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  int uniform_int(int lo, int hi) {  // inclusive
    int v = lo + (int)(uniform() * (double)(hi - lo + 1));
    return v > hi ? hi : v;
  }
};

// world -> camera pose looking from position c along direction d (unit)
void look_at(const double c[3], const double d[3], double R[9], double t[3]) {
  const double up[3] = {0.0, 1.0, 0.0};
  double x[3] = {d[1] * up[2] - d[2] * up[1], d[2] * up[0] - d[0] * up[2], d[0] * up[1] - d[1] * up[0]};
  double nx = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (double& v : x) v /= nx;
  double y[3] = {d[1] * x[2] - d[2] * x[1], d[2] * x[0] - d[0] * x[2], d[0] * x[1] - d[1] * x[0]};
  for (int j = 0; j < 3; j++) { R[0 * 3 + j] = x[j]; R[1 * 3 + j] = y[j]; R[2 * 3 + j] = d[j]; }
  for (int i = 0; i < 3; i++) t[i] = -(R[i * 3 + 0] * c[0] + R[i * 3 + 1] * c[1] + R[i * 3 + 2] * c[2]);
}

}  // namespace

extern "C" int nsn_generate_ba(const nsn_gen_params* p, int edge_cap, double* cams_gt, double* pts_gt,
                               double* cams0, double* pts0, int* cam_idx, int* pt_idx, double* obs,
                               int* n_edges) {
  if (!p || !cams_gt || !pts_gt || !cams0 || !pts0 || !cam_idx || !pt_idx || !obs || !n_edges)
    return nsn::fail(NSN_EINVAL, "nsn_generate_ba: null argument");
  const int C = p->n_cameras, P = p->n_points;
  if (C < 1 || P < 1 || p->k_min < 1 || p->k_max < p->k_min)
    return nsn::fail(NSN_EINVAL, "nsn_generate_ba: bad sizes");
  SplitMix64 rng(p->seed);
  std::vector<double> R(9 * (size_t)C), t(3 * (size_t)C);
  const double ring_r = (p->layout == 0) ? 7.5 : 10.0;
  for (int c = 0; c < C; c++) {
    double phi = 2.0 * M_PI * (double)c / (double)C;
    double pos[3] = {ring_r * std::cos(phi), 0.0, ring_r * std::sin(phi)};
    double d[3] = {-std::cos(phi), 0.0, -std::sin(phi)};
    look_at(pos, d, &R[9 * c], &t[3 * c]);
    nsn::rot_log(&R[9 * c], cams_gt + 6 * c);
    for (int i = 0; i < 3; i++) cams_gt[6 * c + 3 + i] = t[3 * c + i];
  }
  // points; layout 1 also records each point's arc start
  std::vector<int> arc_start((size_t)P, 0);
  for (int q = 0; q < P; q++) {
    double off[3];
    for (double& v : off) v = -2.0 + 4.0 * rng.uniform();
    if (p->layout == 0) {
      for (int i = 0; i < 3; i++) pts_gt[3 * q + i] = off[i];
    } else {
      // layout 2: the camera window does not wrap round the ring
      const int amax = (p->layout == 2) ? std::max(0, C - std::max(p->arc, 1)) : C - 1;
      int a = rng.uniform_int(0, amax);
      arc_start[q] = a;
      int m = (a + p->arc / 2) % C;
      double phi = 2.0 * M_PI * (double)m / (double)C;
      double centre[3] = {(ring_r - 7.5) * std::cos(phi), 0.0, (ring_r - 7.5) * std::sin(phi)};
      for (int i = 0; i < 3; i++) pts_gt[3 * q + i] = centre[i] + off[i];
    }
  }
  // observations
  int E = 0;
  std::vector<int> pool;
  for (int q = 0; q < P; q++) {
    int n_pool = (p->layout == 0 || p->arc <= 0 || p->arc > C) ? C : p->arc;
    pool.resize((size_t)n_pool);
    for (int i = 0; i < n_pool; i++) pool[i] = (p->layout == 0 || n_pool == C) ? i : (arc_start[q] + i) % C;
    int k = rng.uniform_int(p->k_min, p->k_max);
    if (k > n_pool) k = n_pool;
    for (int i = 0; i < k; i++) {  // partial Fisher-Yates
      int j = rng.uniform_int(i, n_pool - 1);
      int tmp = pool[i]; pool[i] = pool[j]; pool[j] = tmp;
    }
    for (int i = 0; i < k; i++) {
      int c = pool[i];
      if (E >= edge_cap) return nsn::fail(NSN_EINVAL, "nsn_generate_ba: edge_cap too small");
      const double* Rc = &R[9 * c];
      const double* X = pts_gt + 3 * q;
      double Xc[3];
      for (int r = 0; r < 3; r++)
        Xc[r] = Rc[r * 3 + 0] * X[0] + Rc[r * 3 + 1] * X[1] + Rc[r * 3 + 2] * X[2] + t[3 * c + r];
      double u = p->focal * Xc[0] / Xc[2] + p->cx;
      double v = p->focal * Xc[1] / Xc[2] + p->cy;
      double nu = p->obs_sigma * rng.normal();
      double nv = p->obs_sigma * rng.normal();
      if (p->outlier_frac > 0 && rng.uniform() < p->outlier_frac) {
        nu = p->outlier_sigma * rng.normal();
        nv = p->outlier_sigma * rng.normal();
      }
      cam_idx[E] = c;
      pt_idx[E] = q;
      obs[2 * E] = u + nu;
      obs[2 * E + 1] = v + nv;
      E++;
    }
  }
  *n_edges = E;
  // initial state: perturb all but camera 0 / point 0
  std::memcpy(cams0, cams_gt, sizeof(double) * 6 * (size_t)C);
  std::memcpy(pts0, pts_gt, sizeof(double) * 3 * (size_t)P);
  for (int c = 1; c < C; c++) {
    for (int i = 0; i < 3; i++) cams0[6 * c + i] += p->rot_sigma * rng.normal();
    for (int i = 0; i < 3; i++) cams0[6 * c + 3 + i] += p->trans_sigma * rng.normal();
  }
  for (int q = 1; q < P; q++)
    for (int i = 0; i < 3; i++) pts0[3 * q + i] += p->point_sigma * rng.normal();
  return NSN_OK;
}

// Synthetic BA on a given visibility graph: the geometry of the reference's BA
// sample (nelson/samples/BASchurPerfomances.cpp:40-75): points uniform in the
// cube [-point_radius, point_radius]^3, cameras at cam_radius from the origin
// in a uniformly random direction, looking at the origin with a random roll.
// Observations are the projections plus N(0, obs_sigma) pixel noise; the
// initial state perturbs every camera but 0 and every point but 0 (the
// sample's addNoise, :77-92, with normal instead of uniform draws).  Draw
// order: points, cameras (direction, then roll vector), observation noise in
// edge order, camera perturbations, point perturbations.
extern "C" int nsn_generate_on_graph(const nsn_graph_gen_params* p, int n_cameras, int n_points, int n_edges,
                                     const int* cam_idx, const int* pt_idx, double* cams_gt, double* pts_gt,
                                     double* cams0, double* pts0, double* obs) {
  if (!p || !cam_idx || !pt_idx || !cams_gt || !pts_gt || !cams0 || !pts0 || !obs)
    return nsn::fail(NSN_EINVAL, "nsn_generate_on_graph: null argument");
  const int C = n_cameras, P = n_points, E = n_edges;
  if (C < 1 || P < 1 || E < 0) return nsn::fail(NSN_EINVAL, "nsn_generate_on_graph: bad sizes");
  if (!(p->point_radius > 0) || !(p->cam_radius > std::sqrt(3.0) * p->point_radius))
    return nsn::fail(NSN_EINVAL,
                     "nsn_generate_on_graph: cam_radius must exceed sqrt(3) point_radius (every point in front)");
  for (int e = 0; e < E; e++)
    if (cam_idx[e] < 0 || cam_idx[e] >= C || pt_idx[e] < 0 || pt_idx[e] >= P)
      return nsn::fail(NSN_EINVAL, "nsn_generate_on_graph: edge %d (%d, %d) out of range", e, cam_idx[e], pt_idx[e]);
  SplitMix64 rng(p->seed);
  for (int q = 0; q < P; q++)
    for (int i = 0; i < 3; i++) pts_gt[3 * q + i] = p->point_radius * (2.0 * rng.uniform() - 1.0);
  std::vector<double> R(9 * (size_t)C), t(3 * (size_t)C);
  auto unit = [&rng](double v[3]) {
    double n2 = 0.0;
    while (n2 < 1e-12) {
      for (int i = 0; i < 3; i++) v[i] = 2.0 * rng.uniform() - 1.0;
      n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    const double n = std::sqrt(n2);
    for (int i = 0; i < 3; i++) v[i] /= n;
  };
  for (int c = 0; c < C; c++) {
    double dir[3], roll[3];
    unit(dir);
    double z[3] = {-dir[0], -dir[1], -dir[2]};
    double dot = 1.0;
    while (1.0 - std::fabs(dot) < 1e-6) {
      unit(roll);
      dot = roll[0] * z[0] + roll[1] * z[1] + roll[2] * z[2];
    }
    double y[3] = {z[1] * roll[2] - z[2] * roll[1], z[2] * roll[0] - z[0] * roll[2], z[0] * roll[1] - z[1] * roll[0]};
    double ny = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    for (double& v : y) v /= ny;
    double x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
    double* Rc = &R[9 * (size_t)c];
    for (int j = 0; j < 3; j++) { Rc[0 * 3 + j] = x[j]; Rc[1 * 3 + j] = y[j]; Rc[2 * 3 + j] = z[j]; }
    const double centre[3] = {p->cam_radius * dir[0], p->cam_radius * dir[1], p->cam_radius * dir[2]};
    for (int i = 0; i < 3; i++)
      t[3 * c + i] = -(Rc[i * 3 + 0] * centre[0] + Rc[i * 3 + 1] * centre[1] + Rc[i * 3 + 2] * centre[2]);
    nsn::rot_log(Rc, cams_gt + 6 * (size_t)c);
    for (int i = 0; i < 3; i++) cams_gt[6 * (size_t)c + 3 + i] = t[3 * c + i];
  }
  for (int e = 0; e < E; e++) {
    const int c = cam_idx[e], q = pt_idx[e];
    const double* Rc = &R[9 * (size_t)c];
    const double* X = pts_gt + 3 * (size_t)q;
    double Xc[3];
    for (int r = 0; r < 3; r++)
      Xc[r] = Rc[r * 3 + 0] * X[0] + Rc[r * 3 + 1] * X[1] + Rc[r * 3 + 2] * X[2] + t[3 * c + r];
    obs[2 * (size_t)e] = p->focal * Xc[0] / Xc[2] + p->cx + p->obs_sigma * rng.normal();
    obs[2 * (size_t)e + 1] = p->focal * Xc[1] / Xc[2] + p->cy + p->obs_sigma * rng.normal();
  }
  std::memcpy(cams0, cams_gt, sizeof(double) * 6 * (size_t)C);
  std::memcpy(pts0, pts_gt, sizeof(double) * 3 * (size_t)P);
  for (int c = 1; c < C; c++) {
    for (int i = 0; i < 3; i++) cams0[6 * (size_t)c + i] += p->rot_sigma * rng.normal();
    for (int i = 0; i < 3; i++) cams0[6 * (size_t)c + 3 + i] += p->trans_sigma * rng.normal();
  }
  for (int q = 1; q < P; q++)
    for (int i = 0; i < 3; i++) pts0[3 * (size_t)q + i] += p->point_sigma * rng.normal();
  return NSN_OK;
}
