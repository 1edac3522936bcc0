// Device view of one window of a batch of independent BA problems.
#pragma once

#include "kernels.h"

namespace nsn {

constexpr int kSChunk = 32;  // terms per chunk of the S assembly
constexpr int kPart = 42;    // partial sums per chunk: the 6x6 block, then 6 of bS

struct BWin {
  BAView v;                                   // the window's structure and state
  int n;                                      // order of S (6 NU), <= 256
  int nP;                                     // n padded to a multiple of 64
  int nsb;                                    // S blocks (i <= j)
  const int *sb_i, *sb_j;                     // their camera pair
  // the terms Z_a Z_b^T of every S block ((a, a) over the camera's W blocks
  // for a diagonal block, the (W, W) pairs sharing a point otherwise), block
  // after block (terms: int2 (a, b)), cut into chunks of at most kSChunk
  // terms; the chunks are numbered block after block (block b owns chunks
  // blk_ch_ptr[b] .. blk_ch_ptr[b + 1]) and listed longest first in ch
  // (int4: first term, end term, chunk number, 1 for a diagonal block);
  // sb_diag[u]: the diagonal block of camera u
  int nch;
  const int *terms, *ch, *blk_ch_ptr, *sb_diag;
  void *U, *bU, *V, *bV, *W, *Vinv, *Y, *tW;  // Hessian blocks (element type of the batch); Y = W L^-T
  void* part;                                 // kPart partial sums per chunk
  void* S;                                    // nP x nP, column-major, lower
  void* Linv;                                 // nP / 64 inverses of the diagonal factor tiles
  double *xU, *xV, *cams_bak, *pts_bak;       // xU: nP entries
  // iterative refinement (mixed precision): the FP64 residual (rU 6 NU,
  // rV 3 NV) and gradient (bU64, bV64), its reduction to the camera system
  // (rS) and the camera correction (zU), nP entries each
  double *rU, *rV, *bU64, *bV64, *rS, *zU;
  int ir_steps;    // at most this many refinement steps per increment
  double ir_tol;   // a step stops after its residual once max |r| <= ir_tol * max |b| (0: never)
};

// One persistent launch: `grid` workgroups of 256 threads take windows from
// counter[0] (zeroed by the caller) in the order counter[1 .. nwin] (largest
// windows first, so that the last windows taken are short) and run each window's LM solve with the
// settings in its control block.  nb: the factorisation variant, 8 or 16
// (panels of that width) or 64 (64x64 tiles).  refine: the instance that
// refines every increment (mixed precision only; -2 for FP64).
template <typename T>
int launch_batch_lm(const BWin* wins, int nwin, DevCtl* ctls, int* counter, int nb, int grid, bool refine,
                    hipStream_t st);

// zeroes the window counter of the persistent launch (a kernel, so that the
// launch can be captured in a graph without a memset node)
void launch_batch_reset(int* counter, hipStream_t st);

}  // namespace nsn
