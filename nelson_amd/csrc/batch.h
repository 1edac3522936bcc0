// Device view of one window of a batch of independent BA problems.
#pragma once

#include "kernels.h"

namespace nsn {

struct BWin {
  BAView v;                                   // the window's structure and state
  int n;                                      // order of S (6 NU), <= 256
  int nP;                                     // n padded to a multiple of 64
  int nsb;                                    // S blocks (i <= j)
  const int *sb_i, *sb_j, *sb_ptr, *sb_pairs;  // and their (W, W) pair lists
  void *U, *bU, *V, *bV, *W, *Vinv, *Y, *tW;  // Hessian blocks (element type of the batch)
  void* S;                                    // nP x nP, column-major, lower
  void* Linv;                                 // nP / 64 inverses of the diagonal factor tiles
  double *xU, *xV, *cams_bak, *pts_bak;       // xU: nP entries
};

// One persistent launch: `grid` workgroups of 256 threads take windows from
// *counter (zeroed by the caller) and run each window's LM solve with the
// settings in its control block.  nb: the factorisation variant, 8 or 16
// (panels of that width) or 64 (64x64 tiles).
template <typename T>
int launch_batch_lm(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int grid, hipStream_t st);

}  // namespace nsn
