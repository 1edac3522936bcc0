// Error reporting and host rotation helpers.
#include <cmath>
#include <cstring>
#include <string>

#include "internal.h"
#include "nelson_amd.h"

namespace nsn {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void clear_error() { g_err.clear(); }

// Rodrigues' formula R = I + A [w]x + B [w]x^2 with A = sin(th)/th,
// B = (1 - cos th)/th^2 (written 2 sin^2(th/2)/th^2), Taylor series below 1e-3.
void rot_exp(const double w[3], double R[9]) {
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = std::sqrt(th2), A, B;
  if (th < 1e-3) {
    A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
  } else {
    double h = std::sin(0.5 * th);
    A = std::sin(th) / th;
    B = 2.0 * h * h / th2;
  }
  const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double k2 = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j] + K[i * 3 + 2] * K[2 * 3 + j];
      R[i * 3 + j] = (i == j ? 1.0 : 0.0) + A * K[i * 3 + j] + B * k2;
    }
}

// Rotation matrix -> angle-axis through a unit quaternion (largest-component
// extraction, Shepperd's method), then w = 2 atan2(|v|, qw) v/|v|.
void rot_log(const double R[9], double w[3]) {
  double tr = R[0] + R[4] + R[8];
  double q[4];  // w, x, y, z
  if (tr > R[0] && tr > R[4] && tr > R[8]) {
    double s = std::sqrt(1.0 + tr) * 2.0;
    q[0] = 0.25 * s;
    q[1] = (R[7] - R[5]) / s;
    q[2] = (R[2] - R[6]) / s;
    q[3] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
    q[0] = (R[7] - R[5]) / s;
    q[1] = 0.25 * s;
    q[2] = (R[1] + R[3]) / s;
    q[3] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
    q[0] = (R[2] - R[6]) / s;
    q[1] = (R[1] + R[3]) / s;
    q[2] = 0.25 * s;
    q[3] = (R[5] + R[7]) / s;
  } else {
    double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
    q[0] = (R[3] - R[1]) / s;
    q[1] = (R[2] + R[6]) / s;
    q[2] = (R[5] + R[7]) / s;
    q[3] = 0.25 * s;
  }
  if (q[0] < 0) for (double& v : q) v = -v;
  double nv = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (nv < 1e-300) { w[0] = w[1] = w[2] = 0.0; return; }
  double ang = 2.0 * std::atan2(nv, q[0]);
  for (int i = 0; i < 3; i++) w[i] = ang * q[i + 1] / nv;
}

}  // namespace nsn

extern "C" const char* nsn_last_error(void) { return nsn::g_err.c_str(); }

extern "C" int nsn_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return NSN_OK;
}

extern "C" void nsn_lm_default_settings(nsn_lm_settings* s) {
  if (!s) return;
  s->maxNumIt = 100; s->maxNumSubIt = 10; s->minNumIt = 0;
  s->epsBVector = 1e-15; s->epsIncSquare = 1e-15 * 1e-15; s->epsChi2 = 1e-15;
  s->tau = 1e-3; s->penaltyV = 2.0; s->initV = 2.0;
  s->maxScaleFactor = 2.0 / 3.0; s->minScaleFactor = 1.0 / 3.0;
}

extern "C" void nsn_gn_default_settings(nsn_gn_settings* s) {
  if (!s) return;
  s->maxNumIt = 100; s->minNumIt = 0;
  s->epsBVector = 1e-15; s->epsIncSquare = 1e-15 * 1e-15; s->epsChi2 = 1e-15;
  s->absLambda = 0; s->relLambda = 0;
}
