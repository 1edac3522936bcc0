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

// Task kinds of the worker queue (5 ints per task: kind, i, j, a, b):
//   CT_TILE (i, j), i >= j+2: C_ij -= sum_k L_ik L_jk^T, then L_ij = C_ij L_jj^-T
//   CT_PRE_SUB (j+1, j): the pre-update of C_{j+1,j} with every k < j
//   CT_PRE_DIAG (j+1, j): the pre-update of C_{j+1,j+1} with every k < j
//   CT_SUB (j+1, j): L_{j+1,j} = C_{j+1,j} L_jj^-T from the pre-update
//   CT_PART (r0, r1, slot, c): -sum of L_{r0,k} L_{r1,k}^T over the chunk c
//     of k (kCholPartLen columns from c * kCholPartLen, not before the
//     first column both rows reach) into the scratch tile `slot`
// A sum of more than kCholSplitMin terms (a separator row of a
// nested-dissection order reaching over a whole subtree: one workgroup would
// take about a millisecond over it) is cut into chunks aligned to multiples
// of kCholPartLen: every chunk but the last is a CT_PART task, queued after
// the other tasks of the key at which its tiles can be complete, and the
// sum's own task (a = first slot, b = number of chunks) adds the chunks in
// order, then its last chunk.  Measured on a 4000-camera path in that order
// (sums of up to 370 terms): factorisation 2.53 -> 2.13 ms; shorter sums
// (the 1000-camera ring's reach 93 terms) lose more to the extra tile
// round trips than the parallel chunks gain.
enum : int { CT_TILE = 0, CT_PRE_SUB = 1, CT_PRE_DIAG = 2, CT_SUB = 3, CT_PART = 4 };
constexpr int kCholTaskInts = 5;
constexpr int kCholPartLen = 16;
constexpr int kCholSplitMin = 128;

struct CholQueue {
  std::vector<int> task;  // kCholTaskInts per task, in queue order
  int ntasks = 0;
  int nparts = 0;  // scratch tiles of the CT_PART tasks
};

// the worker task queue of the factorisation.  The diagonal step j+1 waits
// for the two pre-updates of column j, whose last terms need the tiles
// (j+1, j-1) and (j, j-1); each of these tasks accumulates about j tile
// products in one workgroup.  So the pre-updates of column j and the tile
// (j+2, j) are queued `ahead` keys early (before the tasks of the columns
// with key key(j) - ahead): they then have all but their last terms done
// when those terms' tiles are published.  Only such hoisted tasks can wait
// for tasks not yet taken, at most 3 * ahead + 3 of them per chain group at
// a time, so the queue cannot deadlock while the launch has more workers
// than that (chol_lookahead); chol_pick_lookahead checks the actual queue
// instead (chol_queue_completes).  Tiles outside the envelope get no task,
// and the updates of a tile (i, j) run over k >= max(first[i], first[j])
// only.  A column j that starts a chain (first[j] == j) has no pre-updates
// (there is no column before it in its tile row); the pre-update of
// (j+1, j+1) exists only when its row reaches a column before j.
inline CholQueue chol_queue(int NT, int ahead, const int* first = nullptr, int max_groups = 1,
                            int split_min = kCholSplitMin) {
  const std::vector<int> starts = chol_chain_groups(NT, first, max_groups);
  const std::vector<int> key = chol_keys(NT, first, starts);
  struct Q {
    int key, cls, j, ord;
    int rec[kCholTaskInts];
  };
  std::vector<Q> q;
  CholQueue out;
  auto F = [&](int i) { return chol_first(first, i); };
  auto L = kCholPartLen;
  // a sum task over the rows (r0, r1), k < kend, queued at (ke, cls): its
  // chunks become CT_PART tasks at their own keys
  auto add_sum = [&](int kind, int i, int j, int r0, int r1, int kend, int ke, int cls, int& o) {
    const int k0 = std::max(F(r0), F(r1));
    int a = 0, b = 0;
    if (kend - k0 > split_min) {
      const int c0 = k0 / L, cl = (kend - 1) / L;
      a = out.nparts;
      b = cl - c0;
      for (int c = c0; c < cl; c++) {
        // queued after the other tasks of the key at which its tiles can be
        // complete, but never after its own sum's task (then just before it)
        int nat = 0;
        for (int k = std::max(k0, c * L); k < (c + 1) * L; k++) nat = std::max(nat, key[k] + 1);
        if (nat < ke) q.push_back({nat, 2, j, o++, {CT_PART, r0, r1, out.nparts++, c}});
        else q.push_back({ke, cls, j, o++, {CT_PART, r0, r1, out.nparts++, c}});
      }
    }
    q.push_back({ke, cls, j, o++, {kind, i, j, a, b}});
  };
  for (int j = 0; j < NT; j++) {
    const int ke = std::max(0, key[j] - ahead);
    int o = 0;
    if (j + 1 < NT && F(j + 1) <= j) {
      add_sum(CT_PRE_SUB, j + 1, j, j + 1, j, j, ke, 0, o);
      if (F(j + 1) < j) {
        // the pre-update of (j+1, j+1) reads the tiles (j+1, k) of every
        // column k its row reaches, which may come after column j (a
        // separator row reaches into the other subtrees): it is queued
        // after all of them, less the look-ahead
        int nat = key[j];
        for (int k = F(j + 1); k < j; k++) nat = std::max(nat, key[k] + 1);
        add_sum(CT_PRE_DIAG, j + 1, j, j + 1, j + 1, j, std::max(0, nat - ahead), 0, o);
      }
    }
    if (j + 2 < NT && F(j + 2) <= j) add_sum(CT_TILE, j + 2, j, j + 2, j, j, ke, 0, o);
    if (j + 1 < NT && F(j + 1) <= j) q.push_back({key[j], 1, j, o++, {CT_SUB, j + 1, j, 0, 0}});
    for (int i = j + 3; i < NT; i++)
      if (F(i) <= j) add_sum(CT_TILE, i, j, i, j, j, key[j], 1, o);
  }
  std::stable_sort(q.begin(), q.end(), [](const Q& a, const Q& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.j != b.j) return a.j < b.j;
    return a.ord < b.ord;
  });
  out.ntasks = (int)q.size();
  out.task.resize((size_t)kCholTaskInts * q.size());
  for (size_t t = 0; t < q.size(); t++)
    for (int e = 0; e < kCholTaskInts; e++) out.task[kCholTaskInts * t + e] = q[t].rec[e];
  return out;
}

// the look-ahead of chol_queue for a launch of `workers` worker
// workgroups beside `groups` chain workgroups, from the bound alone: 12
// columns, fewer when the launch is small (more than groups * (3 * ahead +
// 3) workers are needed), none below 8 workers per group
inline int chol_lookahead(int workers, int groups = 1) {
  if (groups < 1) groups = 1;
  int a = (workers / groups - 4) / 4;
  return a < 0 ? 0 : (a > 12 ? 12 : a);
}

// Exact check that the persistent launch completes: `workers` workgroups
// take tasks in queue order and each holds its task until every tile,
// pre-update and chain step it reads is complete; the chain workgroups run
// their steps in order.  Tasks are taken in order, so when the first p are
// taken, the tasks among them that can never complete without a later task
// are the same whatever order the others finish in; the launch stalls
// exactly when those fill every workgroup.  The greedy simulation below
// finishes every task if and only if that never happens.  With ahead = 0
// every task reads only earlier tasks, so the launch always completes.
inline bool chol_queue_completes(int NT, const int* first, int max_groups, int ahead, int workers,
                                 int split_min = kCholSplitMin) {
  const CholQueue cq = chol_queue(NT, ahead, first, max_groups, split_min);
  const int nt = cq.ntasks;
  const int* raw = cq.task.data();
  const std::vector<int> starts = chol_chain_groups(NT, first, max_groups);
  auto F = [&](int i) { return chol_first(first, i); };
  // the task that completes tile (i, k), k < i: sub-diagonal or tile task
  std::vector<long long> row_base(NT + 1, 0);
  for (int i = 0; i < NT; i++) row_base[i + 1] = row_base[i] + (i - F(i) + 1);
  std::vector<int> tile_task(row_base[NT], -1), pre_sub(NT, -1), pre_diag(NT, -1), part(cq.nparts, -1);
  auto tslot = [&](int i, int k) { return row_base[i] + (k - F(i)); };
  for (int q = 0; q < nt; q++) {
    const int* r = raw + kCholTaskInts * q;
    if (r[0] == CT_TILE || r[0] == CT_SUB) tile_task[tslot(r[1], r[2])] = q;
    else if (r[0] == CT_PRE_SUB) pre_sub[r[2]] = q;
    else if (r[0] == CT_PRE_DIAG) pre_diag[r[2]] = q;
    else part[r[3]] = q;
  }
  const int nn = nt + NT;  // tasks, then chain steps
  std::vector<int> rem(nn, 0);
  std::vector<std::vector<int>> users(nn);
  auto dep = [&](int v, int u) {
    if (u < 0) return;
    rem[v]++;
    users[u].push_back(v);
  };
  // the reads of a sum over rows (r0, r1), k in [lo, hi)
  auto sum_deps = [&](int v, int r0, int r1, int lo, int hi) {
    for (int k = lo; k < hi; k++) {
      dep(v, tile_task[tslot(r0, k)]);
      if (r1 != r0) dep(v, tile_task[tslot(r1, k)]);
    }
  };
  // a sum task: its chunks, then its last chunk
  auto owner_deps = [&](int v, const int* r, int r0, int r1, int kend) {
    const int k0 = std::max(F(r0), F(r1));
    int lo = k0;
    if (r[4] > 0) {
      for (int p = r[3]; p < r[3] + r[4]; p++) dep(v, part[p]);
      lo = std::max(k0, (kend - 1) / kCholPartLen * kCholPartLen);
    }
    sum_deps(v, r0, r1, lo, kend);
  };
  for (int q = 0; q < nt; q++) {
    const int* r = raw + kCholTaskInts * q;
    const int i = r[1], j = r[2];
    if (r[0] == CT_TILE) {
      owner_deps(q, r, i, j, j);
      dep(q, nt + j);
    } else if (r[0] == CT_PRE_SUB) {
      owner_deps(q, r, j + 1, j, j);
    } else if (r[0] == CT_PRE_DIAG) {
      owner_deps(q, r, j + 1, j + 1, j);
    } else if (r[0] == CT_SUB) {
      dep(q, pre_sub[j]);
      dep(q, nt + j);
    } else {
      const int c = r[4];
      sum_deps(q, i, j, std::max(std::max(F(i), F(j)), c * kCholPartLen), (c + 1) * kCholPartLen);
    }
  }
  size_t g = 0;
  for (int j = 0; j < NT; j++) {
    while (g + 1 < starts.size() && starts[g + 1] <= j) g++;
    if (j > starts[g]) dep(nt + j, nt + j - 1);
    if (j >= 1 && F(j) <= j - 1) dep(nt + j, pre_sub[j - 1]);
    if (j >= 1 && F(j) < j - 1) dep(nt + j, pre_diag[j - 1]);
  }
  std::vector<char> taken(nn, 0);
  std::vector<int> ready;
  for (int j = 0; j < NT; j++) {
    taken[nt + j] = 1;
    if (rem[nt + j] == 0) ready.push_back(nt + j);
  }
  int next = 0, free_wg = workers, ndone = 0;
  while (true) {
    while (free_wg > 0 && next < nt) {
      taken[next] = 1;
      free_wg--;
      if (rem[next] == 0) ready.push_back(next);
      next++;
    }
    if (ready.empty()) break;
    const int v = ready.back();
    ready.pop_back();
    ndone++;
    if (v < nt) free_wg++;
    for (int u : users[v])
      if (--rem[u] == 0 && taken[u]) ready.push_back(u);
  }
  return ndone == nn;
}

// the look-ahead the library uses: the bound of chol_lookahead, lowered to
// the largest one whose launch completes (chol_queue_completes) when the
// task graph is small enough to check (at most `max_deps` reads).  The
// bound is also the faster choice measured: with 16 chains (bound 2) a
// look-ahead of 5, which the check allows, holds more workers in waiting
// hoisted tasks (4000-camera path: 2.5 -> 3.7 ms).
inline int chol_pick_lookahead(int NT, const int* first, int max_groups, int workers,
                               long long max_deps = 20000000, int split_min = kCholSplitMin) {
  const std::vector<int> starts = chol_chain_groups(NT, first, max_groups);
  const int groups = (int)starts.size() - 1;
  long long reads = 0;
  for (int i = 0; i < NT; i++)
    for (int j = chol_first(first, i); j <= i && reads <= max_deps; j++)
      reads += 2LL * (j - std::max(chol_first(first, i), chol_first(first, j))) + 2;
  const int bound = chol_lookahead(workers, groups);
  if (reads > max_deps || workers < 1) return bound;
  for (int a = bound; a > 0; a--)
    if (chol_queue_completes(NT, first, max_groups, a, workers, split_min)) return a;
  return 0;
}

}  // namespace nsn
