// Host check of the tile Cholesky's worker queue and diagonal chains
// (nelson_amd/csrc/chol_tasks.h).
//
// 1. Dense S: every look-ahead gives the same tasks as the plain column
//    order, the sub-diagonal tile of column j follows the pre-update it
//    reads, the tasks that are not queued early keep the column order, and
//    the early ones come at most `ahead` columns early.
// 2. Any envelope: every task that is not queued early comes after every
//    task it reads, and every task reads only tiles inside the envelope.
// 3. A simulation of the persistent launch: W workgroups take tasks in queue
//    order and each holds its task until every tile, pre-update and chain
//    step the task reads is complete; the chain workgroups run their steps
//    in order beside them.  With the look-ahead bound (chol_lookahead) every
//    task completes for every W, on dense S, on banded envelopes and on
//    envelopes of independent chains (the nested-dissection layout: leaves,
//    then separators whose rows reach over their subtree); a look-ahead far
//    beyond that bound deadlocks the simulated launch with few workers,
//    which shows the simulation sees what it should.
// 4. The library's own exact check (chol_queue_completes) agrees with this
//    simulation, and the look-ahead the library picks with it
//    (chol_pick_lookahead) completes here.
//
// Output: one line per failure, then "OK" or "FAIL"; exit status 0 / 1.
#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>
#include <vector>

#include "chol_tasks.h"

using nsn::chol_chain_groups;
using nsn::chol_lookahead;
using nsn::kCholPartLen;
using nsn::kCholTaskInts;

namespace {

struct Task {
  int kind, i, j, a, b;
};

std::vector<Task> tasks_of(int NT, int ahead, const int* first = nullptr, int groups = 1) {
  const nsn::CholQueue cq = nsn::chol_queue(NT, ahead, first, groups);
  std::vector<Task> t(cq.ntasks);
  for (int q = 0; q < cq.ntasks; q++) {
    const int* r = cq.task.data() + kCholTaskInts * q;
    t[q] = {r[0], r[1], r[2], r[3], r[4]};
  }
  return t;
}

// queued ahead of its column: the hoisted pre-updates and tile (j+2, j),
// and the chunks of long sums (queued when their own tiles can be ready)
bool early(const Task& t) { return t.kind == 1 || t.kind == 2 || (t.kind == 0 && t.i == t.j + 2) || t.kind == 4; }

int F(const int* first, int i) { return first ? first[i] : 0; }

int check_order(int NT, int ahead) {
  int bad = 0;
  const std::vector<Task> a = tasks_of(NT, 0), b = tasks_of(NT, ahead);
  std::map<std::tuple<int, int, int, int, int>, int> ca, cb;
  for (const Task& t : a) ca[{t.kind, t.i, t.j, t.a, t.b}]++;
  for (const Task& t : b) cb[{t.kind, t.i, t.j, t.a, t.b}]++;
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
    } else if (t.kind != 4 && t.j > col + 1 + ahead) {  // queued with (before) the tasks of column col + 1
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

// the task ids and chain steps a task reads (nodes: tasks 0..nt-1, then
// chain steps nt..nt+NT-1); -1 where a read tile has no task
struct Graph {
  std::vector<std::vector<int>> deps;
  int nt = 0, missing = 0;
};

Graph build_graph(int NT, const std::vector<Task>& t, const int* first, const std::vector<int>& starts) {
  Graph gr;
  const int nt = (int)t.size();
  gr.nt = nt;
  std::map<std::tuple<int, int, int>, int> id;
  std::map<int, int> part;  // chunk scratch slot -> task
  for (int q = 0; q < nt; q++) {
    if (t[q].kind == 4) {
      if (part.count(t[q].a)) gr.missing++;  // two chunks in one slot
      part[t[q].a] = q;
    } else {
      id[{t[q].kind, t[q].i, t[q].j}] = q;
    }
  }
  auto get = [&](int kind, int i, int j) {
    auto it = id.find({kind, i, j});
    if (it == id.end()) {
      gr.missing++;
      return -1;
    }
    return it->second;
  };
  auto tile = [&](int i, int k) {
    if (k < F(first, i) || k < F(first, k)) gr.missing++;  // outside the envelope
    return i == k + 1 ? get(3, i, k) : get(0, i, k);
  };
  auto chain = [&](int j) { return nt + j; };
  gr.deps.assign(nt + NT, {});
  // a sum over rows (r0, r1), k in [k0, kend): chunks a .. a+b-1 of the rows
  // (each the chunk c0 + m, c0 the chunk of k0) and its own last chunk
  auto sum = [&](std::vector<int>& d, const Task& x, int r0, int r1, int kend) {
    const int k0 = std::max(F(first, r0), F(first, r1));
    int lo = k0;
    if (x.b > 0) {
      const int c0 = k0 / kCholPartLen;
      for (int m = 0; m < x.b; m++) {
        auto it = part.find(x.a + m);
        if (it == part.end()) {
          gr.missing++;
          continue;
        }
        const Task& pt = t[it->second];
        if (pt.i != r0 || pt.j != r1 || pt.b != c0 + m) gr.missing++;  // the wrong chunk
        d.push_back(it->second);
      }
      lo = std::max(k0, (c0 + x.b) * kCholPartLen);
      if ((kend - 1) / kCholPartLen != c0 + x.b) gr.missing++;  // the last chunk is not its own
    }
    for (int k = lo; k < kend; k++) {
      d.push_back(tile(r0, k));
      if (r1 != r0) d.push_back(tile(r1, k));
    }
  };
  for (int q = 0; q < nt; q++) {
    const Task& x = t[q];
    std::vector<int>& d = gr.deps[q];
    if (x.kind == 0) {  // C_ij -= sum_k L_ik L_jk^T, then the inverse of step j
      sum(d, x, x.i, x.j, x.j);
      d.push_back(chain(x.j));
    } else if (x.kind == 1) {  // pre-update of (j+1, j)
      sum(d, x, x.j + 1, x.j, x.j);
    } else if (x.kind == 2) {  // pre-update of (j+1, j+1)
      sum(d, x, x.j + 1, x.j + 1, x.j);
    } else if (x.kind == 3) {  // sub-diagonal tile: the pre-update and the inverse of step j
      d.push_back(get(1, x.j + 1, x.j));
      d.push_back(chain(x.j));
    } else {  // chunk b of the rows (i, j)
      const int lo = std::max(std::max(F(first, x.i), F(first, x.j)), x.b * kCholPartLen);
      for (int k = lo; k < (x.b + 1) * kCholPartLen; k++) {
        d.push_back(tile(x.i, k));
        if (x.i != x.j) d.push_back(tile(x.j, k));
      }
    }
  }
  // chain step j: the previous step of its group (the workgroup runs them in
  // order) and, unless j starts a chain, the pre-updates of column j - 1
  size_t g = 0;
  for (int j = 0; j < NT; j++) {
    while (g + 1 < starts.size() && starts[g + 1] <= j) g++;
    if (j > starts[g]) gr.deps[chain(j)].push_back(chain(j - 1));
    if (F(first, j) <= j - 1) {
      gr.deps[chain(j)].push_back(get(1, j, j - 1));
      if (F(first, j) < j - 1) gr.deps[chain(j)].push_back(get(2, j, j - 1));
    }
  }
  for (auto& d : gr.deps) d.erase(std::remove(d.begin(), d.end(), -1), d.end());
  return gr;
}

// every task that is not queued early (with no look-ahead: every task) comes
// after every task it reads, and no task reads a tile outside the envelope
// or without a task
int check_deps(int NT, const int* first, int groups, int ahead, const char* what) {
  const std::vector<Task> t = tasks_of(NT, ahead, first, groups);
  const std::vector<int> starts = chol_chain_groups(NT, first, groups);
  Graph gr = build_graph(NT, t, first, starts);
  int bad = 0;
  if (gr.missing) {
    std::printf("%s NT=%d: %d reads of tiles outside the envelope or without a task\n", what, NT, gr.missing);
    bad++;
  }
  for (int q = 0; q < gr.nt; q++) {
    // with no look-ahead the queue is in dependency order; otherwise only
    // the tasks queued with their column are
    if (ahead > 0 && early(t[q])) continue;
    for (int u : gr.deps[q])
      if (u < gr.nt && u > q) {
        std::printf("%s NT=%d: task %d (%d, %d, %d) reads task %d queued after it\n", what, NT, q, t[q].kind, t[q].i,
                    t[q].j, u);
        bad++;
        break;
      }
  }
  // tiles inside the envelope all have a task (or are diagonal)
  std::map<std::pair<int, int>, int> have;
  for (const Task& x : t)
    if (x.kind == 0 || x.kind == 3) have[{x.i, x.j}]++;
  for (int i = 0; i < NT; i++)
    for (int j = F(first, i); j < i; j++)
      if (have[{i, j}] != 1) {
        std::printf("%s NT=%d: tile (%d, %d) has %d tasks\n", what, NT, i, j, have[{i, j}]);
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

// the nested-dissection layout: a part [lo, hi) is a banded leaf (a chain:
// first[lo] = lo, later rows reach back at most a random band, never before
// lo) or [A, B, S] with A and B parts built the same way and the separator
// rows S reaching back to lo (fill over the whole part); the last row is dense
struct Rng {
  unsigned x;
  unsigned next() { return x = x * 1103515245u + 12345u, x >> 8; }
};

void nd_part(std::vector<int>& f, int lo, int hi, int depth, Rng& r) {
  const int len = hi - lo;
  if (depth == 0 || len < 8) {
    const int band = 2 + (int)(r.next() % 5);
    for (int i = lo; i < hi; i++) f[i] = std::max(lo, i - band);
    return;
  }
  const int sep = 1 + (int)(r.next() % std::max(1, len / 5));
  const int m = lo + std::max(2, (len - sep) / 2 + (int)(r.next() % 3) - 1);
  const int s = hi - sep;
  if (m >= s) {
    nd_part(f, lo, hi, 0, r);
    return;
  }
  nd_part(f, lo, m, depth - 1, r);
  nd_part(f, m, s, depth - 1, r);
  for (int i = s; i < hi; i++) f[i] = lo;
}

std::vector<int> nd_envelope(int NT, unsigned seed, int depth) {
  std::vector<int> f(NT, 0);
  Rng r{seed * 2654435761u + 77u};
  nd_part(f, 0, NT, depth, r);
  if (NT > 0) f[NT - 1] = 0;
  return f;
}

// true when every task of the simulated launch completes
bool simulate(int NT, int ahead, int W, const int* first = nullptr, int groups = 1) {
  const std::vector<Task> t = tasks_of(NT, ahead, first, groups);
  const std::vector<int> starts = chol_chain_groups(NT, first, groups);
  Graph gr = build_graph(NT, t, first, starts);
  if (gr.missing) return false;
  const int nt = gr.nt, nn = nt + NT;
  std::vector<std::vector<int>> users(nn);
  std::vector<int> rem(nn);
  for (int v = 0; v < nn; v++) {
    rem[v] = (int)gr.deps[v].size();
    for (int u : gr.deps[v]) users[u].push_back(v);
  }
  std::vector<char> taken(nn, 0);
  std::vector<int> ready;  // taken tasks and chain steps whose dependencies are complete
  for (int j = 0; j < NT; j++) {
    taken[nt + j] = 1;
    if (rem[nt + j] == 0) ready.push_back(nt + j);
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
    ndone++;
    if (v < nt) free_wg++;
    for (int u : users[v])
      if (--rem[u] == 0 && taken[u]) ready.push_back(u);
  }
  return ndone == nn;
}

int check_launches(int NT, const int* first, int groups, const char* what) {
  int bad = 0;
  const int ng = (int)chol_chain_groups(NT, first, groups).size() - 1;
  const int n = nsn::chol_queue(NT, 0, first, groups).ntasks;
  for (int W = 1; W <= 256 - ng; W++) {
    const int workers = W < n ? W : n;  // the launch never has more workers than tasks
    if (workers < 1) continue;
    if (!simulate(NT, chol_lookahead(workers, ng), workers, first, groups)) {
      std::printf("deadlock: %s NT=%d groups=%d workers=%d ahead=%d\n", what, NT, ng, workers,
                  chol_lookahead(workers, ng));
      bad++;
    }
    if (W > 64) W += 6;  // every width up to 64, then a sample
  }
  return bad;
}

}  // namespace

int main() {
  int bad = 0;
  for (int NT = 1; NT <= 40; NT++)
    for (int ahead : {0, 1, 2, 5, 12}) bad += check_order(NT, ahead);
  for (int NT : {94, 120, 282}) bad += check_order(NT, 12);
  // the look-ahead the library uses, for every launch width
  for (int NT : {2, 3, 5, 8, 13, 21, 40, 94}) bad += check_launches(NT, nullptr, 1, "dense");
  // banded envelopes with a dense last row
  for (int NT : {5, 13, 40, 94})
    for (unsigned seed = 1; seed <= 4; seed++) {
      const std::vector<int> f = random_envelope(NT, seed);
      bad += check_deps(NT, f.data(), 1, 12, "band");
      bad += check_launches(NT, f.data(), 1, "band");
    }
  // independent chains (nested dissection), several chain workgroups, and
  // more chains than chain workgroups (merged groups)
  for (int NT : {6, 19, 40, 94, 160})
    for (unsigned seed = 1; seed <= 4; seed++)
      for (int depth : {1, 2, 3}) {
        const std::vector<int> f = nd_envelope(NT, seed, depth);
        for (int groups : {1, 2, 3, 16}) {
          for (int ahead : {0, 3, 12}) bad += check_deps(NT, f.data(), groups, ahead, "nd");
          bad += check_launches(NT, f.data(), groups, "nd");
        }
      }
  {
    // chains are found where the envelope has no tile left of the diagonal
    const std::vector<int> f = nd_envelope(94, 3, 2);
    const std::vector<int> st = chol_chain_groups(94, f.data(), 16);
    int starts = 0;
    for (int j = 0; j < 94; j++) starts += f[j] == j;
    if ((int)st.size() - 1 != starts) {
      std::printf("nd: %d chain starts, %d groups\n", starts, (int)st.size() - 1);
      bad++;
    }
    if (starts < 3) {
      std::printf("nd: the test envelope has only %d chains\n", starts);
      bad++;
    }
  }
  // the library's own check (chol_queue_completes) agrees with this
  // simulation, and the look-ahead it picks (chol_pick_lookahead) completes
  // here for every width
  for (int NT : {13, 40, 94})
    for (unsigned seed = 1; seed <= 3; seed++)
      for (int depth : {0, 2, 3}) {
        const std::vector<int> f = depth ? nd_envelope(NT, seed, depth) : random_envelope(NT, seed);
        for (int groups : {1, 16})
          for (int W : {1, 3, 8, 20, 60, 120, 240}) {
            for (int ahead : {0, 4, 12})
              if (nsn::chol_queue_completes(NT, f.data(), groups, ahead, W) !=
                  simulate(NT, ahead, W, f.data(), groups)) {
                std::printf("chol_queue_completes disagrees: NT=%d seed=%u depth=%d groups=%d W=%d ahead=%d\n", NT,
                            seed, depth, groups, W, ahead);
                bad++;
              }
            const int a = nsn::chol_pick_lookahead(NT, f.data(), groups, W);
            if (!simulate(NT, a, W, f.data(), groups)) {
              std::printf("picked look-ahead %d deadlocks: NT=%d seed=%u depth=%d groups=%d W=%d\n", a, NT, seed,
                          depth, groups, W);
              bad++;
            }
          }
      }
  // the simulation does detect a queue that can deadlock, and so does the
  // library's check
  if (simulate(40, 12, 4) || nsn::chol_queue_completes(40, nullptr, 1, 12, 4)) {
    std::printf("the simulation found no deadlock with 4 workers and ahead=12\n");
    bad++;
  }
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}
