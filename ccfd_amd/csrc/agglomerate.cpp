// agglomerate.cpp - element agglomeration for the multigrid preconditioner
// (host side); what the levels hold is described in agglomerate.hpp.
#include "agglomerate.hpp"

#include <algorithm>
#include <tuple>

namespace ccfd {

namespace {

// weighted element graph without self-loops, CSR
struct Graph {
  int n = 0;
  std::vector<int> start, adj;
  std::vector<double> wt;
};

// heavy-edge matching in element order: every unmatched element is paired
// with its most strongly coupled unmatched neighbour (or stays alone)
std::vector<int> match(const Graph& g, int& nOut) {
  std::vector<int> m(g.n, -1);
  int k = 0;
  for (int i = 0; i < g.n; ++i) {
    if (m[i] >= 0) continue;
    int best = -1;
    double bw = -1.0;
    for (int t = g.start[i]; t < g.start[i + 1]; ++t) {
      const int j = g.adj[t];
      if (m[j] < 0 && g.wt[t] > bw) {
        bw = g.wt[t];
        best = j;
      }
    }
    m[i] = k;
    if (best >= 0) m[best] = k;
    ++k;
  }
  nOut = k;
  return m;
}

// the graph of the groups: weights summed over the edges between groups
Graph contract(const Graph& g, const std::vector<int>& map, int nc) {
  std::vector<std::tuple<int, int, double>> e;
  e.reserve(g.adj.size());
  for (int i = 0; i < g.n; ++i)
    for (int t = g.start[i]; t < g.start[i + 1]; ++t) {
      const int a = map[i], b = map[g.adj[t]];
      if (a != b) e.emplace_back(a, b, g.wt[t]);
    }
  std::sort(e.begin(), e.end(), [](const auto& x, const auto& y) {
    return std::get<0>(x) != std::get<0>(y) ? std::get<0>(x) < std::get<0>(y)
                                            : std::get<1>(x) < std::get<1>(y);
  });
  Graph c;
  c.n = nc;
  c.start.assign(nc + 1, 0);
  for (size_t k = 0; k < e.size(); ++k) {
    const int a = std::get<0>(e[k]), b = std::get<1>(e[k]);
    if (k > 0 && std::get<0>(e[k - 1]) == a && std::get<1>(e[k - 1]) == b) {
      c.wt.back() += std::get<2>(e[k]);
      continue;
    }
    c.adj.push_back(b);
    c.wt.push_back(std::get<2>(e[k]));
    c.start[a + 1]++;
  }
  for (int a = 0; a < nc; ++a) c.start[a + 1] += c.start[a];
  return c;
}

template <typename T>
void to_csr(int n, const std::vector<std::pair<int, T>>& items,
            std::vector<int64_t>& start, std::vector<T>& out) {
  start.assign(n + 1, 0);
  for (const auto& it : items) start[it.first + 1]++;
  for (int a = 0; a < n; ++a) start[a + 1] += start[a];
  out.assign(items.size(), T());
  std::vector<int64_t> pos(start.begin(), start.end() - 1);
  for (const auto& it : items) out[pos[it.first]++] = it.second;
}

}  // namespace

void build_agglomeration(int N, int Np, int S, const int* link,
                         const double* w, int minCells, int maxLevels,
                         std::vector<AggLevel>& levels) {
  levels.clear();
  // level-0 graph: one edge per face to another element (two faces to
  // one element add up)
  Graph g;
  {
    std::vector<int> id(N);
    for (int i = 0; i < N; ++i) id[i] = i;
    Graph raw;
    raw.n = N;
    raw.start.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) {
      for (int s = 0; s < S; ++s) {
        const int lk = link[(size_t)s * Np + i];
        if (lk < 0 || (lk >> 2) == i) continue;
        raw.adj.push_back(lk >> 2);
        raw.wt.push_back(w[(size_t)s * Np + i]);
      }
      raw.start[i + 1] = (int)raw.adj.size();
    }
    g = contract(raw, id, N);
  }
  int nPrev = N;
  for (int lev = 1; lev <= maxLevels && nPrev > minCells; ++lev) {
    int n1 = 0, n2 = 0;
    const std::vector<int> m1 = match(g, n1);
    const Graph g1 = contract(g, m1, n1);
    const std::vector<int> m2 = match(g1, n2);
    Graph g2 = contract(g1, m2, n2);
    if (n2 > nPrev - nPrev / 8) break;   // no longer shrinking
    AggLevel L;
    L.n = n2;
    L.aggOf.resize(nPrev);
    for (int p = 0; p < nPrev; ++p) L.aggOf[p] = m2[m1[p]];
    // members
    {
      std::vector<std::pair<int, int>> it(nPrev);
      for (int p = 0; p < nPrev; ++p) it[p] = {L.aggOf[p], p};
      std::vector<int64_t> st;
      to_csr(n2, it, st, L.mem);
      L.memStart.assign(st.begin(), st.end());
    }
    // off-diagonal pattern: the coarse graph's edges
    L.rowStart = g2.start;
    L.col = g2.adj;
    L.nnz = (int)L.col.size();
    auto nz_of = [&](int a, int b) {
      const auto first = L.col.begin() + L.rowStart[a];
      const auto last = L.col.begin() + L.rowStart[a + 1];
      return (int64_t)(std::lower_bound(first, last, b) - L.col.begin());
    };
    // which blocks of the level above sum into which block here
    std::vector<std::pair<int, int64_t>> dItems, oItems;
    dItems.reserve(nPrev);
    for (int p = 0; p < nPrev; ++p) dItems.push_back({L.aggOf[p], ((int64_t)p << 1) | 1});
    auto off_entry = [&](int p, int q, int64_t idx) {
      const int a = L.aggOf[p], b = L.aggOf[q];
      if (a == b)
        dItems.push_back({a, idx << 1});
      else
        oItems.push_back({(int)nz_of(a, b), idx << 1});
    };
    if (lev == 1) {
      for (int i = 0; i < N; ++i)
        for (int s = 0; s < S; ++s) {
          const int lk = link[(size_t)s * Np + i];
          if (lk < 0 || (lk >> 2) == i) continue;
          off_entry(i, lk >> 2, (int64_t)s * Np + i);
        }
    } else {
      const AggLevel& P = levels.back();
      for (int p = 0; p < nPrev; ++p)
        for (int t = P.rowStart[p]; t < P.rowStart[p + 1]; ++t)
          off_entry(p, P.col[t], t);
    }
    std::stable_sort(dItems.begin(), dItems.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    std::stable_sort(oItems.begin(), oItems.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    to_csr(n2, dItems, L.dSrcStart, L.dSrc);
    to_csr(L.nnz, oItems, L.oSrcStart, L.oSrc);
    // greedy distance-1 colouring for the Gauss-Seidel sweeps, in
    // breadth-first order (few colours, two on a bipartite graph)
    {
      std::vector<int> order;
      order.reserve(n2);
      std::vector<char> seen(n2, 0);
      for (int r = 0; r < n2; ++r) {
        if (seen[r]) continue;
        seen[r] = 1;
        size_t head = order.size();
        order.push_back(r);
        while (head < order.size()) {
          const int u = order[head++];
          for (int t = L.rowStart[u]; t < L.rowStart[u + 1]; ++t)
            if (!seen[L.col[t]]) {
              seen[L.col[t]] = 1;
              order.push_back(L.col[t]);
            }
        }
      }
      std::vector<int> color(n2, -1);
      std::vector<char> used;
      int ncol = 0;
      for (int a : order) {
        used.assign(ncol + 1, 0);
        for (int t = L.rowStart[a]; t < L.rowStart[a + 1]; ++t)
          if (color[L.col[t]] >= 0) used[color[L.col[t]]] = 1;
        int cc = 0;
        while (cc < ncol && used[cc]) ++cc;
        color[a] = cc;
        if (cc == ncol) ++ncol;
      }
      std::vector<std::pair<int, int>> it(n2);
      for (int a = 0; a < n2; ++a) it[a] = {color[a], a};
      std::vector<int64_t> st;
      to_csr(ncol, it, st, L.colorCells);
      L.colorStart.assign(st.begin(), st.end());
    }
    levels.push_back(std::move(L));
    g = std::move(g2);
    nPrev = n2;
  }
}

}  // namespace ccfd
