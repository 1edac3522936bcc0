// Entry points without a reference counterpart that are not built yet:
// they report NSN_EINVAL with a message instead of doing anything.
#include "internal.h"
#include "nelson_amd.h"

using nsn::fail;

extern "C" int nsn_batch_create(const nsn_problem_desc*, int, nsn_batch** out) {
  if (out) *out = nullptr;
  return fail(NSN_EINVAL, "nsn_batch_create: not available in this build");
}
extern "C" int nsn_batch_destroy(nsn_batch*) { return NSN_OK; }
extern "C" int nsn_batch_set_state(nsn_batch*, const double*, const double*) {
  return fail(NSN_EINVAL, "nsn_batch_set_state: not available in this build");
}
extern "C" int nsn_batch_get_state(nsn_batch*, double*, double*) {
  return fail(NSN_EINVAL, "nsn_batch_get_state: not available in this build");
}
extern "C" int nsn_batch_lm_solve(nsn_batch*, const nsn_lm_settings*, double*, double*) {
  return fail(NSN_EINVAL, "nsn_batch_lm_solve: not available in this build");
}
