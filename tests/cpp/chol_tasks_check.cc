// Host check of the tile Cholesky's worker queue (nelson_amd/csrc/chol_tasks.h).
//
// 1. Every look-ahead gives the same tasks as the plain column order, the
//    sub-diagonal tile of column j follows the pre-update it reads, the tasks
//    that are not queued early keep the column order, and the early ones come
//    at most `ahead` columns early.
// 2. A simulation of the persistent launch: W workgroups take tasks in queue
//    order and each holds its task until every tile, pre-update and chain
//    step the task reads is complete; the diagonal chain (workgroup 0) runs
//    beside them.  With the look-ahead the library uses for W workers
//    (chol_lookahead) every task completes for every W; a look-ahead far
//    beyond that bound deadlocks the simulated launch with few workers,
//    which shows the simulation sees what it should.
//
// Output: one line per failure, then "OK" or "FAIL"; exit status 0 / 1.
#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include <vector>

#include "chol_tasks.h"

using nsn::chol_lookahead;
using nsn::chol_task_list;

namespace {

struct Task {
  int kind, i, j;
};

std::vector<Task> tasks_of(int NT, int ahead, const int* first = nullptr) {
  int n = 0;
  chol_task_list(NT, ahead, nullptr, &n, first);
  std::vector<int> raw(3 * (size_t)(n > 0 ? n : 1));
  chol_task_list(NT, ahead, raw.data(), &n, first);
  std::vector<Task> t(n);
  for (int q = 0; q < n; q++) t[q] = {raw[3 * q], raw[3 * q + 1], raw[3 * q + 2]};
  return t;
}

bool early(const Task& t) { return t.kind == 1 || t.kind == 2 || (t.kind == 0 && t.i == t.j + 2); }

int check_order(int NT, int ahead) {
  int bad = 0;
  const std::vector<Task> a = tasks_of(NT, 0), b = tasks_of(NT, ahead);
  std::map<std::tuple<int, int, int>, int> ca, cb;
  for (const Task& t : a) ca[{t.kind, t.i, t.j}]++;
  for (const Task& t : b) cb[{t.kind, t.i, t.j}]++;
  if (ca != cb) {
    std::printf("NT=%d ahead=%d: task set differs from the plain order\n", NT, ahead);
    bad++;
  }
  std::vector<int> pos_pre(NT, -1), pos_sub(NT, -1);
  int col = -1;  // column of the last task that is not queued early
  for (int q = 0; q < (int)b.size(); q++) {
    const Task& t = b[q];
    if (t.kind == 1) pos_pre[t.j] = q;
    if (t.kind == 3) pos_sub[t.j] = q;
    if (!early(t)) {
      if (t.j < col) {
        std::printf("NT=%d ahead=%d: task %d (column %d) after column %d\n", NT, ahead, q, t.j, col);
        bad++;
      }
      col = t.j;
    } else if (t.j > col + 1 + ahead) {  // queued with (before) the tasks of column col + 1
      std::printf("NT=%d ahead=%d: task %d of column %d queued more than %d columns early\n", NT, ahead, q, t.j,
                  ahead);
      bad++;
    }
  }
  for (int j = 0; j + 1 < NT; j++)
    if (pos_pre[j] < 0 || pos_sub[j] < pos_pre[j]) {
      std::printf("NT=%d ahead=%d: sub-diagonal tile of column %d before its pre-update\n", NT, ahead, j);
      bad++;
    }
  return bad;
}

// an envelope: first[i] <= i - 2, the last row dense, otherwise a band of
// random width (some rows reaching back further)
std::vector<int> random_envelope(int NT, unsigned seed) {
  std::vector<int> f(NT, 0);
  unsigned x = seed * 2654435761u + 12345u;
  for (int i = 0; i < NT; i++) {
    x = x * 1103515245u + 12345u;
    const int back = 2 + (int)((x >> 16) % 6) + (((x >> 8) % 7) == 0 ? (int)((x >> 20) % 20) : 0);
    f[i] = i - back < 0 ? 0 : i - back;
  }
  if (NT > 0) f[NT - 1] = 0;
  return f;
}

// true when every task of the simulated launch completes
bool simulate(int NT, int ahead, int W, const int* first = nullptr) {
  const std::vector<Task> t = tasks_of(NT, ahead, first);
  const int nt = (int)t.size();
  // nodes: tasks 0..nt-1, then chain steps nt..nt+NT-1
  std::map<std::tuple<int, int, int>, int> id;
  for (int q = 0; q < nt; q++) id[{t[q].kind, t[q].i, t[q].j}] = q;
  auto tile = [&](int i, int k) { return i == k + 1 ? id.at({3, i, k}) : id.at({0, i, k}); };
  auto chain = [&](int j) { return nt + j; };
  std::vector<std::vector<int>> deps(nt + NT);
  for (int q = 0; q < nt; q++) {
    const Task& x = t[q];
    std::vector<int>& d = deps[q];
    auto k0 = [&](int i, int j) { return first ? std::max(first[i], first[j]) : 0; };
    if (x.kind == 0) {  // C_ij -= sum_k L_ik L_jk^T, then the inverse of step j
      for (int k = k0(x.i, x.j); k < x.j; k++) d.push_back(tile(x.i, k)), d.push_back(tile(x.j, k));
      d.push_back(chain(x.j));
    } else if (x.kind == 1) {  // pre-update of (j+1, j)
      for (int k = k0(x.j + 1, x.j); k < x.j; k++) d.push_back(tile(x.j + 1, k)), d.push_back(tile(x.j, k));
    } else if (x.kind == 2) {  // pre-update of (j+1, j+1)
      for (int k = k0(x.j + 1, x.j + 1); k < x.j; k++) d.push_back(tile(x.j + 1, k));
    } else {  // sub-diagonal tile: the pre-update and the inverse of step j
      d.push_back(id.at({1, x.j + 1, x.j}));
      d.push_back(chain(x.j));
    }
  }
  for (int j = 0; j < NT; j++) {  // step j: step j-1 and the pre-updates of column j-1
    if (j >= 1) {
      deps[chain(j)].push_back(chain(j - 1));
      deps[chain(j)].push_back(id.at({1, j, j - 1}));
      if (j >= 2) deps[chain(j)].push_back(id.at({2, j, j - 1}));
    }
  }
  const int nn = nt + NT;
  std::vector<std::vector<int>> users(nn);
  std::vector<int> rem(nn);
  for (int v = 0; v < nn; v++) {
    rem[v] = (int)deps[v].size();
    for (int u : deps[v]) users[u].push_back(v);
  }
  std::vector<char> taken(nn, 0), done(nn, 0);
  std::vector<int> ready;  // taken tasks and chain steps whose dependencies are complete
  for (int j = 0; j < NT; j++) {
    taken[chain(j)] = 1;
    if (rem[chain(j)] == 0) ready.push_back(chain(j));
  }
  int next = 0, free_wg = W, ndone = 0;
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
    done[v] = 1;
    ndone++;
    if (v < nt) free_wg++;
    for (int u : users[v])
      if (--rem[u] == 0 && taken[u]) ready.push_back(u);
  }
  return ndone == nn;
}

}  // namespace

int main() {
  int bad = 0;
  for (int NT = 1; NT <= 40; NT++)
    for (int ahead : {0, 1, 2, 5, 12}) bad += check_order(NT, ahead);
  for (int NT : {94, 120, 282}) bad += check_order(NT, 12);
  // the look-ahead the library uses, for every launch width
  for (int NT : {2, 3, 5, 8, 13, 21, 40, 94})
    for (int W = 1; W <= 255; W++) {
      int n = 0;
      chol_task_list(NT, 0, nullptr, &n);
      const int workers = W < n ? W : n;  // the launch never has more workers than tasks
      if (workers < 1) continue;
      if (!simulate(NT, chol_lookahead(workers), workers)) {
        std::printf("deadlock: NT=%d workers=%d ahead=%d\n", NT, workers, chol_lookahead(workers));
        bad++;
      }
    }
  // envelopes (banded rows, a dense last row): only tiles inside get a task,
  // every tile a task reads is inside, and the launch completes
  for (int NT : {5, 13, 40, 94})
    for (unsigned seed = 1; seed <= 4; seed++) {
      const std::vector<int> f = random_envelope(NT, seed);
      for (const Task& x : tasks_of(NT, 12, f.data()))
        if (x.kind == 0 && f[x.i] > x.j) {
          std::printf("NT=%d seed=%u: task for tile (%d, %d) outside the envelope\n", NT, seed, x.i, x.j);
          bad++;
        }
      for (int W : {1, 2, 7, 20, 60, 255}) {
        int n = 0;
        chol_task_list(NT, 0, nullptr, &n, f.data());
        const int workers = W < n ? W : n;
        if (workers < 1) continue;
        if (!simulate(NT, chol_lookahead(workers), workers, f.data())) {
          std::printf("deadlock: NT=%d seed=%u workers=%d (envelope)\n", NT, seed, workers);
          bad++;
        }
      }
    }
  // the simulation does detect a queue that can deadlock
  if (simulate(40, 12, 4)) {
    std::printf("the simulation found no deadlock with 4 workers and ahead=12\n");
    bad++;
  }
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}
