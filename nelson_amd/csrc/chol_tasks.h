// The worker task queue of the persistent tile Cholesky (chol_kernels.hip),
// built on the host.  Plain C++ (no HIP types), so that host tests can check
// its order and simulate the scheduler.
#pragma once

namespace nsn {

// the worker task list of the factorisation, in queue order (kinds: 0 tile
// (i >= j+2), 1 pre-update of (j+1, j), 2 pre-update of (j+1, j+1), 3 the
// sub-diagonal tile (j+1, j)).  The diagonal chain at step j waits for the
// two pre-updates of column j, whose last terms need the tiles (j+1, j-1)
// and (j, j-1); each of these tasks accumulates about j tile products in one
// workgroup.  So the pre-updates of column j and the tile (j+2, j) are
// queued `ahead` columns early (with the tasks of column j - ahead): they
// then have all but their last terms done when those terms' tiles are
// published.  Only such hoisted tasks can wait for tasks not yet taken, and
// at most 3 * ahead + 3 of them do so at a time, so the queue cannot
// deadlock while the launch has more workers than that (chol_lookahead).
// first (may be null: dense): the first stored tile column of each tile
// row (an envelope with first[i] <= i - 2); tiles outside it are zero in S
// and in the factor and get no task, and the updates of a tile (i, j) run
// over k >= max(first[i], first[j]) only.
inline void chol_task_list(int NT, int ahead, int* out, int* ntasks, const int* first = nullptr) {
  int t = 0;
  auto add = [&](int kind, int i, int j) {
    if (out) {
      out[3 * t] = kind;
      out[3 * t + 1] = i;
      out[3 * t + 2] = j;
    }
    t++;
  };
  auto add_early = [&](int j) {
    if (j + 1 < NT) {
      add(1, j + 1, j);
      if (j >= 1) add(2, j + 1, j);
    }
    if (j + 2 < NT) add(0, j + 2, j);
  };
  for (int j = 0; j <= ahead && j < NT; j++) add_early(j);
  for (int j = 0; j < NT; j++) {
    if (j >= 1 && j + ahead < NT) add_early(j + ahead);
    if (j + 1 < NT) add(3, j + 1, j);
    for (int i = j + 3; i < NT; i++)
      if (!first || first[i] <= j) add(0, i, j);
  }
  *ntasks = t;
}
// the look-ahead of chol_task_list for a launch of `workers` worker
// workgroups: 12 columns, fewer when the launch is small (more than
// 3 * ahead + 3 workers are needed), none below 8 workers
inline int chol_lookahead(int workers) {
  int a = (workers - 4) / 4;
  return a < 0 ? 0 : (a > 12 ? 12 : a);
}

}  // namespace nsn
