// Elimination schedule of the coarse-level tiled Cholesky.
//
// A schedule is a sequence of steps; a step eliminates a set of tile columns
// K, each with its panel (the tile rows I > K with L_IK != 0).  The columns
// of one step are independent (no panel of one contains another), so one
// launch applies all their updates: one task per target tile (I, J), I >= J,
// of the union of their panels, summing the updates of every column whose
// panel holds both I and J.
#include "coarse_schedule.hpp"

#include <algorithm>
#include <cstdio>

namespace t3d {

namespace {

struct Step {
  std::vector<std::pair<int, std::vector<int>>> src;  // (K, panel rows I > K)
};

// banded order: one column per step, panel K+1 .. last tile within bw
std::vector<Step> band_steps(int nblk, int bw) {
  std::vector<Step> st(nblk);
  for (int b = 0; b < nblk; b++) {
    const int t1 = std::min(nblk - 1, (b * kTile + kTile - 1 + bw) / kTile);
    std::vector<int> pan;
    for (int I = b + 1; I <= t1; I++) pan.push_back(I);
    st[b].src.push_back({b, pan});
  }
  return st;
}

// Nested dissection of a chain of planes by odd-even reduction: eliminate
// every other plane of the remaining chain (an independent set), which
// couples their neighbours, and repeat; about log2(planes) levels of tp
// tile columns each.  slot[p] = position of plane p in the elimination order.
std::vector<Step> plane_steps(int planes, int tp, std::vector<int> &slot) {
  std::vector<std::vector<int>> levels;
  std::vector<int> rem(planes);
  for (int p = 0; p < planes; p++) rem[p] = p;
  while (!rem.empty()) {
    std::vector<int> lev, keep;
    for (size_t a = 0; a < rem.size(); a++) (a % 2 == 0 ? lev : keep).push_back(rem[a]);
    levels.push_back(lev);
    rem = keep;
  }
  slot.assign(planes, 0);
  int next = 0;
  for (const auto &lev : levels)
    for (int p : lev) slot[p] = next++;
  // symbolic elimination on the plane graph
  std::vector<std::vector<int>> adj(planes);
  for (int p = 0; p + 1 < planes; p++) {
    adj[p].push_back(p + 1);
    adj[p + 1].push_back(p);
  }
  std::vector<Step> st;
  for (const auto &lev : levels) {
    std::vector<std::vector<int>> nb(lev.size());
    for (size_t a = 0; a < lev.size(); a++) {
      const int p = lev[a];
      nb[a] = adj[p];
      std::sort(nb[a].begin(), nb[a].end(), [&](int u, int v) { return slot[u] < slot[v]; });
      for (int u : nb[a]) {
        auto &au = adj[u];
        au.erase(std::remove(au.begin(), au.end(), p), au.end());
        for (int v : nb[a])
          if (v != u && std::find(au.begin(), au.end(), v) == au.end()) au.push_back(v);
      }
    }
    for (int c = 0; c < tp; c++) {
      Step e;
      for (size_t a = 0; a < lev.size(); a++) {
        const int p = lev[a];
        std::vector<int> pan;
        for (int c2 = c + 1; c2 < tp; c2++) pan.push_back(slot[p] * tp + c2);
        for (int u : nb[a])
          for (int c2 = 0; c2 < tp; c2++) pan.push_back(slot[u] * tp + c2);
        e.src.push_back({slot[p] * tp + c, pan});
      }
      st.push_back(e);
    }
  }
  return st;
}

}  // namespace

bool build_coarse_schedule(int n, int bw, int planes, CoarseSchedule &out) {
  out = CoarseSchedule();
  out.n = n;
  out.pos.resize(n);
  std::vector<Step> steps;
  if (planes > 1 && n % planes == 0) {
    const int pm = n / planes, tp = (pm + kTile - 1) / kTile;
    std::vector<int> slot;
    steps = plane_steps(planes, tp, slot);
    out.np = planes * tp * kTile;
    for (int d = 0; d < n; d++) out.pos[d] = slot[d / pm] * tp * kTile + d % pm;
  } else {
    out.np = (n + kTile - 1) / kTile * kTile;
    steps = band_steps(out.np / kTile, std::max(0, bw));
    for (int d = 0; d < n; d++) out.pos[d] = d;
  }
  const int nblk = out.nblk = out.np / kTile;
  auto &tasks = out.tasks;
  std::vector<std::vector<int>> lcols(nblk);  // K < I with L_IK != 0
  auto factor_only = [&](int K) {
    CholTask t{};
    t.I = t.J = K;
    t.flags = kTaskFactor;
    tasks.push_back(t);
  };
  // launch 0 factors the first step's diagonal tiles
  out.launch_first.push_back(0);
  for (const auto &sk : steps[0].src) factor_only(sk.first);
  out.launch_count.push_back((int)tasks.size());
  for (size_t t = 0; t < steps.size(); t++) {
    std::vector<std::pair<std::pair<int, int>, std::vector<int>>> tgt;  // ((I, J), columns)
    auto find = [&](int I, int J) -> std::vector<int> & {
      for (auto &x : tgt)
        if (x.first.first == I && x.first.second == J) return x.second;
      tgt.push_back({{I, J}, {}});
      return tgt.back().second;
    };
    for (const auto &sk : steps[t].src) {
      for (int I : sk.second) lcols[I].push_back(sk.first);
      for (int I : sk.second)
        for (int J : sk.second)
          if (J <= I) find(I, J).push_back(sk.first);
    }
    std::vector<int> next;
    if (t + 1 < steps.size())
      for (const auto &sk : steps[t + 1].src) next.push_back(sk.first);
    std::vector<CholTask> fac, rest;
    std::vector<int> covered;
    for (const auto &x : tgt) {
      if ((int)x.second.size() > kMaxSrc) {
        char msg[96];
        std::snprintf(msg, sizeof(msg), "%d updates of one tile in one step",
                      (int)x.second.size());
        out.error = msg;
        return false;
      }
      CholTask c{};
      c.I = x.first.first;
      c.J = x.first.second;
      c.nsrc = (int)x.second.size();
      for (int q = 0; q < c.nsrc; q++) c.K[q] = x.second[q];
      const bool nx = c.I == c.J && std::find(next.begin(), next.end(), c.I) != next.end();
      c.flags = nx ? kTaskFactor : 0;
      if (nx) covered.push_back(c.I);
      (nx ? fac : rest).push_back(c);
    }
    const int first = (int)tasks.size();
    for (int K : next)  // not updated in this step: complete already
      if (std::find(covered.begin(), covered.end(), K) == covered.end()) factor_only(K);
    tasks.insert(tasks.end(), fac.begin(), fac.end());  // the critical path first
    tasks.insert(tasks.end(), rest.begin(), rest.end());
    if ((int)tasks.size() > first) {
      out.launch_first.push_back(first);
      out.launch_count.push_back((int)tasks.size() - first);
    }
  }
  // non-zero tile rows of each column of W = L^{-1}: forward reach through
  // the tiles of L
  out.colinfo.assign(2 * nblk, 0);
  for (int J = 0; J < nblk; J++) {
    std::vector<char> in(nblk, 0);
    in[J] = 1;
    const int r0 = (int)out.rows.size() / 4;
    for (int I = J + 1; I < nblk; I++) {
      std::vector<int> ks;
      for (int K : lcols[I])
        if (in[K]) ks.push_back(K);
      if (ks.empty()) continue;
      std::sort(ks.begin(), ks.end());
      out.rows.insert(out.rows.end(), {I, (int)out.klist.size(), (int)ks.size(), 0});
      out.klist.insert(out.klist.end(), ks.begin(), ks.end());
      in[I] = 1;
    }
    out.colinfo[2 * J] = r0;
    out.colinfo[2 * J + 1] = (int)out.rows.size() / 4 - r0;
  }
  return true;
}

}  // namespace t3d
