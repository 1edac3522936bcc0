// The worker task queue and the diagonal chains of the persistent tile
// Cholesky (chol_kernels.hip), built on the host.  Plain C++ (no HIP types),
// so that host tests can check the order and simulate the scheduler.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace nsn {

// first (may be null: dense) holds the first stored tile column of each tile
// row: tiles left of it are zero in S and in the factor.
inline int chol_first(const int* first, int i) { return first ? first[i] : 0; }

// The diagonal chains.  Column j whose tile row has no tile left of the
// diagonal (first[j] == j) depends on no earlier column: its diagonal step
// can start at once, and a new chain starts there.  Chains are run by the
// chain workgroups, at most max_groups of them: group g runs the steps
// starts[g] .. starts[g+1]-1 in order (a group may hold several chains, run
// one after the other).  starts has groups + 1 entries, starts[0] = 0.
inline std::vector<int> chol_chain_groups(int NT, const int* first, int max_groups) {
  std::vector<int> st;
  for (int j = 0; j < NT; j++)
    if (j == 0 || chol_first(first, j) == j) st.push_back(j);
  if (st.empty()) st.push_back(0);
  if (max_groups < 1) max_groups = 1;
  if ((int)st.size() > max_groups) {
    // merge neighbouring chains into max_groups groups of similar length
    std::vector<int> merged{0};
    for (int g = 1; g < max_groups; g++) {
      const long long target = (long long)NT * g / max_groups;
      int best = -1;
      for (int s : st)
        if (s > merged.back() && (best < 0 || std::llabs((long long)s - target) < std::llabs((long long)best - target)))
          best = s;
      if (best > merged.back()) merged.push_back(best);
    }
    st = merged;
  }
  st.push_back(NT);
  return st;
}

// The queue position key of every column: the earliest round at which its
// diagonal step can run, counting one round per step of its chain group and
// after every column its tile row reaches (first[j] <= k < j).  A dense S has
// key(j) = j.  Column keys order the queue: every column a task reads has a
// smaller key than the task's own column.
inline std::vector<int> chol_keys(int NT, const int* first, const std::vector<int>& starts) {
  std::vector<int> key(NT, 0);
  size_t g = 0;
  for (int j = 0; j < NT; j++) {
    while (g + 1 < starts.size() && starts[g + 1] <= j) g++;
    int m = j > starts[g] ? key[j - 1] : -1;  // the previous step of the same group
    for (int k = chol_first(first, j); k < j; k++) m = std::max(m, key[k]);
    key[j] = m + 1;
  }
  return key;
}

// the worker task list of the factorisation, in queue order (kinds: 0 tile
// (i >= j+2), 1 pre-update of (j+1, j), 2 pre-update of (j+1, j+1), 3 the
// sub-diagonal tile (j+1, j)).  The diagonal step j+1 waits for the two
// pre-updates of column j, whose last terms need the tiles (j+1, j-1) and
// (j, j-1); each of these tasks accumulates about j tile products in one
// workgroup.  So the pre-updates of column j and the tile (j+2, j) are
// queued `ahead` keys early (before the tasks of the columns with key
// key(j) - ahead): they then have all but their last terms done when those
// terms' tiles are published.  Only such hoisted tasks can wait for tasks
// not yet taken, at most 3 * ahead + 3 of them per chain group at a time, so
// the queue cannot deadlock while the launch has more workers than that
// (chol_lookahead).  Tiles outside the envelope get no task, and the updates
// of a tile (i, j) run over k >= max(first[i], first[j]) only.  A column j
// that starts a chain (first[j] == j) has no pre-updates (there is no
// column before it in its tile row); the pre-update of (j+1, j+1) exists
// only when its row reaches a column before j.
inline void chol_task_list(int NT, int ahead, int* out, int* ntasks, const int* first = nullptr,
                           int max_groups = 1) {
  const std::vector<int> starts = chol_chain_groups(NT, first, max_groups);
  const std::vector<int> key = chol_keys(NT, first, starts);
  struct Q {
    int key, cls, j, ord, kind, i, jj;
  };
  std::vector<Q> q;
  auto F = [&](int i) { return chol_first(first, i); };
  for (int j = 0; j < NT; j++) {
    const int ke = std::max(0, key[j] - ahead);
    int o = 0;
    if (j + 1 < NT && F(j + 1) <= j) {
      q.push_back({ke, 0, j, o++, 1, j + 1, j});
      if (F(j + 1) < j) {
        // the pre-update of (j+1, j+1) reads the tiles (j+1, k) of every
        // column k its row reaches, which may come after column j (a
        // separator row reaches into the other subtrees): it is queued
        // after all of them, less the look-ahead
        int nat = key[j];
        for (int k = F(j + 1); k < j; k++) nat = std::max(nat, key[k] + 1);
        q.push_back({std::max(0, nat - ahead), 0, j, o++, 2, j + 1, j});
      }
    }
    if (j + 2 < NT && F(j + 2) <= j) q.push_back({ke, 0, j, o++, 0, j + 2, j});
    if (j + 1 < NT && F(j + 1) <= j) q.push_back({key[j], 1, j, o++, 3, j + 1, j});
    for (int i = j + 3; i < NT; i++)
      if (F(i) <= j) q.push_back({key[j], 1, j, o++, 0, i, j});
  }
  std::stable_sort(q.begin(), q.end(), [](const Q& a, const Q& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.j != b.j) return a.j < b.j;
    return a.ord < b.ord;
  });
  if (out)
    for (size_t t = 0; t < q.size(); t++) {
      out[3 * t] = q[t].kind;
      out[3 * t + 1] = q[t].i;
      out[3 * t + 2] = q[t].jj;
    }
  *ntasks = (int)q.size();
}

// the look-ahead of chol_task_list for a launch of `workers` worker
// workgroups beside `groups` chain workgroups: 12 columns, fewer when the
// launch is small (more than groups * (3 * ahead + 3) workers are needed),
// none below 8 workers per group
inline int chol_lookahead(int workers, int groups = 1) {
  if (groups < 1) groups = 1;
  int a = (workers / groups - 4) / 4;
  return a < 0 ? 0 : (a > 12 ? 12 : a);
}

}  // namespace nsn
