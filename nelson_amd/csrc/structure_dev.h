// Structure stage on the device (HIP sources only).
#pragma once

#include <hip/hip_runtime.h>

#include "structure.h"

namespace nsn {

// host_structure with its edge-proportional part (edge orders, W blocks,
// camera-major lists, the (W, W) pairs of S grouped by block) built on the
// current HIP device with radix sorts, on stream st; equal to
// host_structure's result field by field.  The camera order is taken on the
// camera graph read off the S blocks of a first pass in the given order.
int device_structure(const nsn_problem_desc* d, HostStructure& s, bool allow_order, hipStream_t st);

// fields of a and b that differ (0: equal), naming the first in *what
int structure_diff(const HostStructure& a, const HostStructure& b, const char** what);

}  // namespace nsn
