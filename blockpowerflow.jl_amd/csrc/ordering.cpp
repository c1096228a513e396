// Fill-reducing orderings of a square sparse pattern (on A + A^T), host side.
// They stand in for the cusolverSp host reorderings selected by `ordering` in
// glu_analysis_host (src/cusolverRF.jl:70-95):
//   :AMD / :MDQ -> minimum degree on the elimination graph (exact external
//                  degrees, explicit fill edges; George & Liu 1989)
//   :METIS      -> nested dissection by BFS level-structure bisection
//   :RCM        -> reverse Cuthill-McKee from a pseudo-peripheral node
// Each returns perm with perm[k] = original index of the k-th eliminated node.
#include "ordering.h"

#include <algorithm>
#include <numeric>
#include <queue>

namespace bpf {

Graph symmetric_graph(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx) {
  Graph G;
  G.n = n;
  std::vector<int> deg(n, 0);
  for (int i = 0; i < n; ++i)
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      if (j != i) { ++deg[i]; ++deg[j]; }
    }
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i) adj[i].reserve(deg[i]);
  for (int i = 0; i < n; ++i)
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      if (j != i) { adj[i].push_back(j); adj[j].push_back(i); }
    }
  G.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    auto &a = adj[i];
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    G.ptr[i + 1] = G.ptr[i] + (int)a.size();
  }
  G.adj.reserve(G.ptr[n]);
  for (int i = 0; i < n; ++i) G.adj.insert(G.adj.end(), adj[i].begin(), adj[i].end());
  return G;
}

// BFS from s over nodes with mask[v] == tag; returns level of each reached node.
static int bfs_levels(const Graph &G, int s, const std::vector<int> &mask, int tag,
                      std::vector<int> &level, std::vector<int> &order) {
  order.clear();
  order.push_back(s);
  level[s] = 0;
  int depth = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    const int v = order[h];
    for (int k = G.ptr[v]; k < G.ptr[v + 1]; ++k) {
      const int u = G.adj[k];
      if (mask[u] == tag && level[u] < 0) {
        level[u] = level[v] + 1;
        depth = std::max(depth, level[u]);
        order.push_back(u);
      }
    }
  }
  return depth;
}

// Pseudo-peripheral node of the component containing s (George-Liu).
static int peripheral(const Graph &G, int s, const std::vector<int> &mask, int tag,
                      std::vector<int> &level, std::vector<int> &order) {
  int best = s, depth = -1;
  for (int it = 0; it < 8; ++it) {
    for (int v : order) level[v] = -1;
    level[best] = -1;
    const int d = bfs_levels(G, best, mask, tag, level, order);
    if (d <= depth) break;
    depth = d;
    // farthest node of minimum degree
    int cand = best, cdeg = 1 << 30;
    for (int v : order)
      if (level[v] == d) {
        const int dg = G.ptr[v + 1] - G.ptr[v];
        if (dg < cdeg) { cdeg = dg; cand = v; }
      }
    best = cand;
  }
  for (int v : order) level[v] = -1;
  return best;
}

std::vector<int> rcm_order(const Graph &G) {
  const int n = G.n;
  std::vector<int> perm, mask(n, 0), level(n, -1), order;
  std::vector<char> seen(n, 0);
  perm.reserve(n);
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    order.clear();
    const int s = peripheral(G, s0, mask, 0, level, order);
    // Cuthill-McKee: BFS, neighbours by increasing degree
    const size_t start = perm.size();
    perm.push_back(s);
    seen[s] = 1;
    for (size_t h = start; h < perm.size(); ++h) {
      const int v = perm[h];
      std::vector<int> nb;
      for (int k = G.ptr[v]; k < G.ptr[v + 1]; ++k)
        if (!seen[G.adj[k]]) { nb.push_back(G.adj[k]); seen[G.adj[k]] = 1; }
      std::sort(nb.begin(), nb.end(), [&](int a, int b) {
        return (G.ptr[a + 1] - G.ptr[a]) < (G.ptr[b + 1] - G.ptr[b]);
      });
      perm.insert(perm.end(), nb.begin(), nb.end());
    }
    std::reverse(perm.begin() + start, perm.end());
  }
  return perm;
}

std::vector<int> mindeg_order(const Graph &G) {
  // Elimination graph with explicit fill; degree buckets keep the minimum.
  const int n = G.n;
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i) adj[i].assign(G.adj.begin() + G.ptr[i], G.adj.begin() + G.ptr[i + 1]);
  std::vector<int> deg(n), perm;
  std::vector<char> done(n, 0);
  perm.reserve(n);
  // bucket lists (doubly linked)
  std::vector<int> head(n + 1, -1), next(n, -1), prev(n, -1);
  auto insert = [&](int v) {
    const int d = deg[v];
    prev[v] = -1;
    next[v] = head[d];
    if (head[d] >= 0) prev[head[d]] = v;
    head[d] = v;
  };
  auto remove = [&](int v) {
    if (prev[v] >= 0) next[prev[v]] = next[v];
    else head[deg[v]] = next[v];
    if (next[v] >= 0) prev[next[v]] = prev[v];
  };
  for (int i = 0; i < n; ++i) { deg[i] = (int)adj[i].size(); insert(i); }
  int mind = 0;
  std::vector<int> merged;
  for (int step = 0; step < n; ++step) {
    while (head[mind] < 0) ++mind;
    const int v = head[mind];
    remove(v);
    done[v] = 1;
    perm.push_back(v);
    const std::vector<int> nv = adj[v];   // live neighbours of v
    for (int u : nv) {
      // adj[u] <- (adj[u] \ {v}) U (nv \ {u})
      merged.clear();
      const auto &au = adj[u];
      size_t a = 0, b = 0;
      while (a < au.size() || b < nv.size()) {
        int x;
        if (b >= nv.size() || (a < au.size() && au[a] < nv[b])) x = au[a++];
        else if (a >= au.size() || nv[b] < au[a]) x = nv[b++];
        else { x = au[a++]; ++b; }
        if (x != v && x != u && !done[x]) merged.push_back(x);
      }
      remove(u);
      adj[u].swap(merged);
      deg[u] = (int)adj[u].size();
      insert(u);
      if (deg[u] < mind) mind = deg[u];
    }
    std::vector<int>().swap(adj[v]);
  }
  return perm;
}

// Nested dissection: split each component at the middle level of a BFS from a
// pseudo-peripheral node; separators are numbered last.  Small pieces are
// ordered by minimum degree of their induced subgraph.
std::vector<int> nd_order(const Graph &G, int leaf) {
  const int n = G.n;
  std::vector<int> mask(n, 0), level(n, -1), order, perm(n), idx(n, -1);
  int next_tag = 1;
  // Work stack of (tag, nodes); output filled from the back for separators.
  struct Part { std::vector<int> nodes; };
  std::vector<Part> stack;
  {
    Part all;
    all.nodes.resize(n);
    std::iota(all.nodes.begin(), all.nodes.end(), 0);
    stack.push_back(std::move(all));
  }
  int front = 0, back = n;
  while (!stack.empty()) {
    Part P = std::move(stack.back());
    stack.pop_back();
    if (P.nodes.empty()) continue;
    const int tag = next_tag++;
    for (int v : P.nodes) mask[v] = tag;
    if ((int)P.nodes.size() <= leaf) {
      // order the leaf by minimum degree of its induced subgraph
      std::vector<int> gid(P.nodes);
      for (size_t q = 0; q < gid.size(); ++q) idx[gid[q]] = (int)q;
      Graph H;
      H.n = (int)gid.size();
      H.ptr.assign(H.n + 1, 0);
      for (int q = 0; q < H.n; ++q) {
        const int v = gid[q];
        for (int k = G.ptr[v]; k < G.ptr[v + 1]; ++k)
          if (mask[G.adj[k]] == tag) H.adj.push_back(idx[G.adj[k]]);
        H.ptr[q + 1] = (int)H.adj.size();
      }
      auto lp = mindeg_order(H);
      for (int q : lp) perm[front++] = gid[q];
      continue;
    }
    // BFS from a pseudo-peripheral node of the (possibly disconnected) part
    order.clear();
    const int s = peripheral(G, P.nodes[0], mask, tag, level, order);
    const int depth = bfs_levels(G, s, mask, tag, level, order);
    if (order.size() < P.nodes.size()) {
      // disconnected: split off the reached component and recurse on both
      Part a, b;
      for (int v : P.nodes) (level[v] >= 0 ? a : b).nodes.push_back(v);
      for (int v : P.nodes) level[v] = -1;
      stack.push_back(std::move(b));
      stack.push_back(std::move(a));
      continue;
    }
    if (depth < 2) {   // too shallow to bisect: natural order
      for (int v : P.nodes) level[v] = -1;
      for (int v : P.nodes) perm[front++] = v;
      continue;
    }
    const int mid = depth / 2;
    Part a, b;
    std::vector<int> sep;
    for (int v : P.nodes) {
      if (level[v] < mid) a.nodes.push_back(v);
      else if (level[v] > mid) b.nodes.push_back(v);
      else sep.push_back(v);
    }
    for (int v : P.nodes) level[v] = -1;
    for (auto it = sep.rbegin(); it != sep.rend(); ++it) perm[--back] = *it;
    for (int v : sep) mask[v] = -1;
    stack.push_back(std::move(b));
    stack.push_back(std::move(a));
  }
  // front == back here; separators occupy [back, n) in reverse discovery order
  return perm;
}

}  // namespace bpf
