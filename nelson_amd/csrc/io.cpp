// Visibility-graph files: the text format of the reference's Problem::load
// (nelson/samples/Problems.cpp:8-30), "nCameras nPoints nEdges" followed by
// nEdges "camera point" pairs, whitespace separated.  The reference reads the
// numbers without checks; this reader rejects a short or malformed file,
// negative counts and pairs out of range, naming the offending entry.
#include <cerrno>
#include <cstdio>
#include <cstring>

#include "internal.h"
#include "nelson_amd.h"

namespace {

struct File {
  FILE* f;
  explicit File(const char* path) : f(std::fopen(path, "r")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool read_int(FILE* f, long long* v) { return std::fscanf(f, "%lld", v) == 1; }

}  // namespace

extern "C" int nsn_visibility_load(const char* path, int capacity, int* n_cameras, int* n_points, int* n_edges,
                                   int* cam_idx, int* pt_idx) {
  if (!path || !n_cameras || !n_points || !n_edges) return nsn::fail(NSN_EINVAL, "nsn_visibility_load: null argument");
  File file(path);
  if (!file.f) return nsn::fail(NSN_EINVAL, "nsn_visibility_load: cannot open %s: %s", path, std::strerror(errno));
  long long nc, np, ne;
  if (!read_int(file.f, &nc) || !read_int(file.f, &np) || !read_int(file.f, &ne))
    return nsn::fail(NSN_EINVAL, "nsn_visibility_load: %s: missing header (nCameras nPoints nEdges)", path);
  if (nc < 1 || np < 1 || ne < 0 || nc > 0x7fffffff || np > 0x7fffffff || ne > 0x7fffffff)
    return nsn::fail(NSN_EINVAL, "nsn_visibility_load: %s: bad header %lld %lld %lld", path, nc, np, ne);
  *n_cameras = (int)nc;
  *n_points = (int)np;
  *n_edges = (int)ne;
  if (!cam_idx && !pt_idx) return NSN_OK;  // sizes only
  if (!cam_idx || !pt_idx) return nsn::fail(NSN_EINVAL, "nsn_visibility_load: null argument");
  if (capacity < ne)
    return nsn::fail(NSN_EINVAL, "nsn_visibility_load: %s holds %lld pairs, capacity %d", path, ne, capacity);
  for (long long e = 0; e < ne; e++) {
    long long c, q;
    if (!read_int(file.f, &c) || !read_int(file.f, &q))
      return nsn::fail(NSN_EINVAL, "nsn_visibility_load: %s: pair %lld of %lld missing or malformed", path, e, ne);
    if (c < 0 || c >= nc || q < 0 || q >= np)
      return nsn::fail(NSN_EINVAL, "nsn_visibility_load: %s: pair %lld (%lld, %lld) out of range", path, e, c, q);
    cam_idx[e] = (int)c;
    pt_idx[e] = (int)q;
  }
  return NSN_OK;
}

// Bundle Adjustment in the Large (BAL) problem files: a header "nCameras
// nPoints nObservations", nObservations lines "camera point x y", then 9
// numbers per camera (angle-axis, translation, f, k1, k2) and 3 per point,
// whitespace separated.  This is the data format of the ladybug problem whose
// visibility graph the reference embeds (nelson/samples/Problems.cpp:38).
extern "C" int nsn_bal_load(const char* path, int capacity, int* n_cameras, int* n_points, int* n_obs,
                            int* cam_idx, int* pt_idx, double* obs, double* cams, double* intrinsics,
                            double* pts) {
  if (!path || !n_cameras || !n_points || !n_obs) return nsn::fail(NSN_EINVAL, "nsn_bal_load: null argument");
  File file(path);
  if (!file.f) return nsn::fail(NSN_EINVAL, "nsn_bal_load: cannot open %s: %s", path, std::strerror(errno));
  long long nc, np, ne;
  if (!read_int(file.f, &nc) || !read_int(file.f, &np) || !read_int(file.f, &ne))
    return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: missing header (nCameras nPoints nObservations)", path);
  if (nc < 1 || np < 1 || ne < 0 || nc > 0x7fffffff / 9 || np > 0x7fffffff / 3 || ne > 0x7fffffff / 2)
    return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: bad header %lld %lld %lld", path, nc, np, ne);
  *n_cameras = (int)nc;
  *n_points = (int)np;
  *n_obs = (int)ne;
  const bool want = cam_idx || pt_idx || obs || cams || intrinsics || pts;
  if (!want) return NSN_OK;  // sizes only
  if (!cam_idx || !pt_idx || !obs || !cams || !intrinsics || !pts)
    return nsn::fail(NSN_EINVAL, "nsn_bal_load: null argument");
  if (capacity < ne)
    return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s holds %lld observations, capacity %d", path, ne, capacity);
  for (long long e = 0; e < ne; e++) {
    long long c, q;
    double x, y;
    if (!read_int(file.f, &c) || !read_int(file.f, &q) || std::fscanf(file.f, "%lf %lf", &x, &y) != 2)
      return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: observation %lld of %lld missing or malformed", path, e, ne);
    if (c < 0 || c >= nc || q < 0 || q >= np)
      return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: observation %lld (%lld, %lld) out of range", path, e, c, q);
    cam_idx[e] = (int)c;
    pt_idx[e] = (int)q;
    obs[2 * e] = x;
    obs[2 * e + 1] = y;
  }
  for (long long c = 0; c < nc; c++) {
    double v[9];
    for (int k = 0; k < 9; k++)
      if (std::fscanf(file.f, "%lf", &v[k]) != 1)
        return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: camera %lld parameter %d missing or malformed", path, c, k);
    for (int k = 0; k < 6; k++) cams[6 * c + k] = v[k];
    for (int k = 0; k < 3; k++) intrinsics[3 * c + k] = v[6 + k];
  }
  for (long long q = 0; q < np; q++)
    for (int k = 0; k < 3; k++)
      if (std::fscanf(file.f, "%lf", &pts[3 * q + k]) != 1)
        return nsn::fail(NSN_EINVAL, "nsn_bal_load: %s: point %lld coordinate %d missing or malformed", path, q, k);
  return NSN_OK;
}
