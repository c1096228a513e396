// Self-check of the host-only code (orderings, host sparse LU, level
// analysis, elimination plans, the batched solve's thread pool), built with
// g++ alone under the address / undefined-behaviour or thread sanitizers.
// Every result is checked against the property that defines it:
//   - symmetric_graph: the off-diagonal pattern of A + A^T, sorted, unique;
//   - rcm / minimum-degree / nested-dissection orderings: permutations;
//   - host_lu: L U = A[P, Q] to rounding (dense product on small n), P and Q
//     permutations, a structurally singular matrix raises BPF_E_SINGULAR;
//   - analyze_triop: levels partition the operator's rows, every dependency
//     lies in an earlier level and every row past level 0 has one in the
//     level before (which fixes the levels uniquely), each stored entry
//     gathers the factor entry of its (row, column), relabelled operators
//     are the plain ones with indices mapped;
//   - elim_plan_host: the update positions a linear search of each row finds;
//   - run_pool: every item runs exactly once; an error raised by one item on
//     a worker reaches the caller with its message.
// Exit status 0 when every check holds.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <set>

#include "host_common.h"
#include "ordering.h"

using namespace bpf;

static int g_fail = 0, g_checks = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    ++g_checks;                                           \
    if (!(cond)) {                                        \
      if (g_fail < 20) {                                  \
        fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
        fprintf(stderr, __VA_ARGS__);                     \
        fputc('\n', stderr);                              \
      }                                                   \
      ++g_fail;                                           \
    }                                                     \
  } while (0)

struct Csr {
  int n = 0;
  std::vector<int> rowptr{0}, colidx, diagpos;
  std::vector<double> val;
};

// Random pattern with sorted rows, a diagonal in every row (unless
// drop_diag), about `deg` off-diagonal entries per row, some empty rows of
// off-diagonals; values diagonally weighted.
static Csr random_csr(int n, int deg, unsigned seed, bool drop_diag = false) {
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> col(0, std::max(0, n - 1));
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  Csr A;
  A.n = n;
  A.rowptr.assign(n + 1, 0);
  A.diagpos.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    std::set<int> cols;
    if (!drop_diag) cols.insert(i);
    const int d = (i % 7 == 3) ? 0 : deg;
    for (int t = 0; t < d; ++t) cols.insert(col(rng));
    if (i > 0 && i % 5 == 0) cols.insert(i - 1);   // some chains -> many levels
    for (int j : cols) {
      if (j == i) A.diagpos[i] = (int)A.colidx.size();
      A.colidx.push_back(j);
      A.val.push_back(j == i ? 4.0 + deg + u(rng) : u(rng));
    }
    A.rowptr[i + 1] = (int)A.colidx.size();
  }
  return A;
}

static bool is_perm(const std::vector<int> &p, int n) {
  if ((int)p.size() != n) return false;
  std::vector<char> seen(n, 0);
  for (int v : p) {
    if (v < 0 || v >= n || seen[v]) return false;
    seen[v] = 1;
  }
  return true;
}

static void check_graph_orderings(const Csr &A) {
  const int n = A.n;
  const Graph G = symmetric_graph(n, A.rowptr, A.colidx);
  CHECK((int)G.ptr.size() == n + 1 && G.ptr[0] == 0 && G.ptr[n] == (int)G.adj.size(), "graph ptr n=%d", n);
  std::vector<std::set<int>> want(n);
  for (int i = 0; i < n; ++i)
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
      if (A.colidx[k] != i) {
        want[i].insert(A.colidx[k]);
        want[A.colidx[k]].insert(i);
      }
  for (int i = 0; i < n; ++i) {
    std::vector<int> got(G.adj.begin() + G.ptr[i], G.adj.begin() + G.ptr[i + 1]);
    CHECK(got == std::vector<int>(want[i].begin(), want[i].end()), "adjacency of %d (n=%d)", i, n);
  }
  CHECK(is_perm(rcm_order(G), n), "rcm order n=%d", n);
  CHECK(is_perm(mindeg_order(G), n), "minimum-degree order n=%d", n);
  CHECK(is_perm(nd_order(G, 8), n), "nested-dissection order (leaf 8) n=%d", n);
  CHECK(is_perm(nd_order(G), n), "nested-dissection order n=%d", n);
}

static void check_host_lu(const Csr &A) {
  const int n = A.n;
  const std::vector<int> Qs = mindeg_order(symmetric_graph(n, A.rowptr, A.colidx));
  const HostLU F = host_lu(n, A.rowptr, A.colidx, A.val, Qs, 1e-14);
  CHECK(is_perm(F.P, n) && is_perm(F.Q, n), "LU permutations n=%d", n);
  CHECK((int)F.frow.size() == n + 1 && F.frow[n] == (int)F.fcol.size() && F.fcol.size() == F.fval.size(),
        "LU storage n=%d", n);
  if (n > 160) return;
  std::vector<double> Ad((size_t)n * n, 0.0), L((size_t)n * n, 0.0), U((size_t)n * n, 0.0);
  double amax = 0.0;
  for (int i = 0; i < n; ++i)
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
      Ad[(size_t)i * n + A.colidx[k]] = A.val[k];
      amax = std::max(amax, std::fabs(A.val[k]));
    }
  for (int k = 0; k < n; ++k) {
    L[(size_t)k * n + k] = 1.0;
    for (int e = F.frow[k]; e < F.frow[k + 1]; ++e) {
      const int j = F.fcol[e];
      CHECK(j >= 0 && j < n && (e == F.frow[k] || F.fcol[e - 1] < j), "LU row %d columns sorted", k);
      if (j < k) L[(size_t)k * n + j] = F.fval[e];
      else U[(size_t)k * n + j] = F.fval[e];
    }
  }
  double err = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k <= std::min(i, j); ++k) s += L[(size_t)i * n + k] * U[(size_t)k * n + j];
      err = std::max(err, std::fabs(s - Ad[(size_t)F.P[i] * n + F.Q[j]]));
    }
  CHECK(err <= 1e-12 * n * amax, "||LU - A[P,Q]||_max = %.3e (n=%d)", err, n);
}

static void check_singular() {
  // column 2 has no entries: structurally singular
  Csr A;
  A.n = 3;
  A.rowptr = {0, 2, 3, 5};
  A.colidx = {0, 1, 1, 0, 1};
  A.val = {1.0, 2.0, 3.0, 4.0, 5.0};
  bool thrown = false;
  try {
    host_lu(3, A.rowptr, A.colidx, A.val, {0, 1, 2}, 0.0);
  } catch (const Error &e) {
    thrown = e.code == BPF_E_SINGULAR && std::strlen(last_error()) > 0;
  }
  CHECK(thrown, "structurally singular matrix raises BPF_E_SINGULAR");
}

// Level analysis of one kind / window, optionally relabelled.
static void check_triop(const Csr &A, int kind, const TriWindow &win) {
  const int n = A.n;
  TriLevels op;
  analyze_triop(op, n, A.rowptr, A.colidx, kind, win);
  const int row_lo = win.row_lo, row_hi = win.row_hi < 0 ? n : win.row_hi;
  const int col_lo = win.col_lo, col_hi = win.col_hi < 0 ? n : win.col_hi;
  const std::vector<int> *rl = win.relabel;
  std::vector<int> inv(n);
  for (int i = 0; i < n; ++i) inv[rl ? (*rl)[i] : i] = i;   // storage -> original
  // Expected entries per original row: (dependency, gathered position).
  std::vector<std::vector<std::pair<int, int>>> want(n);
  auto keep = [&](int i, int j) {
    if (j < col_lo || j >= col_hi) return false;
    return kind == 1 ? j > i : j < i;
  };
  for (int i = 0; i < n; ++i) {
    if (kind != 3 && (i < row_lo || i >= row_hi)) continue;
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
      const int j = A.colidx[k];
      if (!keep(i, j)) continue;
      if (kind == 3) want[j].push_back({i, k});
      else want[i].push_back({j, k});
    }
  }
  std::vector<char> member(n, 0);
  int nrows = 0;
  for (int i = 0; i < n; ++i) {
    member[i] = (kind == 3 || (i >= row_lo && i < row_hi)) && !(win.skip_empty && want[i].empty());
    nrows += member[i];
  }
  CHECK(op.n == nrows && op.ndim == n, "kind %d: %d rows, want %d", kind, op.n, nrows);
  CHECK((int)op.h_lvl.size() == op.nlevels + 1 && op.h_lvl[0] == 0 && op.h_lvl[op.nlevels] == op.n,
        "kind %d: level offsets", kind);
  for (int l = 0; l < op.nlevels; ++l) CHECK(op.h_lvl[l] < op.h_lvl[l + 1], "kind %d: empty level %d", kind, l);
  CHECK((int)op.h_order.size() == op.n && (int)op.h_ptr.size() == op.n + 1 &&
        (int)op.h_dgather.size() == op.n && op.h_col.size() == op.h_gather.size(), "kind %d: sizes", kind);
  if (g_fail) return;
  std::vector<int> level(n, -1);   // by original row
  for (int l = 0; l < op.nlevels; ++l)
    for (int t = op.h_lvl[l]; t < op.h_lvl[l + 1]; ++t) {
      const int s = op.h_order[t];
      CHECK(s >= 0 && s < n && level[inv[s]] < 0 && member[inv[s]], "kind %d: stored row %d", kind, t);
      if (s >= 0 && s < n) level[inv[s]] = l;
    }
  if (g_fail) return;
  int max_len = 0;
  for (int l = 0; l < op.nlevels; ++l)
    for (int t = op.h_lvl[l]; t < op.h_lvl[l + 1]; ++t) {
      const int i = inv[op.h_order[t]];
      std::vector<std::pair<int, int>> got;
      bool prev = false;
      for (int o = op.h_ptr[t]; o < op.h_ptr[t + 1]; ++o) {
        const int dep = inv[op.h_col[o]];
        got.push_back({dep, op.h_gather[o]});
        if (member[dep]) {
          CHECK(level[dep] >= 0 && level[dep] < l, "kind %d: row %d (level %d) depends on %d (level %d)",
                kind, i, l, dep, level[dep]);
          prev = prev || level[dep] == l - 1;
        }
      }
      CHECK(l == 0 || prev, "kind %d: row %d at level %d has no dependency at level %d", kind, i, l, l - 1);
      CHECK(got == want[i], "kind %d: entries of row %d", kind, i);
      max_len = std::max(max_len, op.h_ptr[t + 1] - op.h_ptr[t]);
      if (A.diagpos[i] >= 0) CHECK(op.h_dgather[t] == A.diagpos[i], "kind %d: diagonal of row %d", kind, i);
      if (rl)   // each level's rows in storage order
        CHECK(t == op.h_lvl[l] || op.h_order[t - 1] < op.h_order[t], "kind %d: level %d order", kind, l);
    }
  CHECK(op.max_row_nnz == max_len, "kind %d: max row length", kind);
}

static void check_triops(const Csr &A, unsigned seed) {
  const int n = A.n;
  for (int kind = 0; kind < 4; ++kind) check_triop(A, kind, TriWindow());
  std::mt19937 rng(seed);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::shuffle(perm.begin(), perm.end(), rng);
  for (int kind = 0; kind < 4; ++kind) {
    TriWindow w;
    w.relabel = &perm;
    check_triop(A, kind, w);
  }
  if (n >= 4) {
    TriWindow w;
    w.row_lo = n / 4;
    w.row_hi = 3 * n / 4;
    w.col_lo = n / 8;
    w.skip_empty = true;
    w.unit = true;
    check_triop(A, 0, w);
    check_triop(A, 1, w);
    w.skip_empty = false;
    check_triop(A, 2, w);
  }
  if (n == 0) return;
  // the same pattern without row n/2's diagonal
  Csr B = A;
  const int r = n / 2, d = A.diagpos[r];
  B.colidx.erase(B.colidx.begin() + d);
  for (int i = r + 1; i <= n; ++i) --B.rowptr[i];
  bool thrown = false;
  try {
    TriLevels op;
    analyze_triop(op, B.n, B.rowptr, B.colidx, 1, TriWindow());
  } catch (const Error &e) {
    thrown = e.code == BPF_E_SINGULAR;
  }
  CHECK(thrown, "missing diagonal raises BPF_E_SINGULAR");
}

static void check_elim_plan(const Csr &A) {
  const int n = A.n;
  std::vector<PivInfo> pinfo;
  std::vector<int> map;
  elim_plan_host(n, A.rowptr, A.colidx, A.diagpos, pinfo, map);
  CHECK((int)pinfo.size() == std::max(1, A.rowptr[n]) && !map.empty() && map.back() == -1, "plan sizes n=%d", n);
  size_t used = 0;
  for (int i = 0; i < n; ++i) {
    const int r0 = A.rowptr[i], r1 = A.rowptr[i + 1];
    for (int e = r0; e < A.diagpos[i]; ++e) {
      const int k = A.colidx[e];
      const PivInfo &pi = pinfo[e];
      CHECK(pi.diag == A.diagpos[k] && pi.end == A.rowptr[k + 1] && pi.row == k, "pivot (%d,%d) extent", i, k);
      for (int j = A.diagpos[k] + 1; j < A.rowptr[k + 1]; ++j) {
        int want = -1;
        for (int f = r0; f < r1; ++f)
          if (A.colidx[f] == A.colidx[j]) want = f - r0;
        const int m = pi.moff + (j - A.diagpos[k] - 1);
        CHECK(m >= 0 && m + 1 < (int)map.size() && map[m] == want, "update (%d,%d) via %d", i, A.colidx[j], k);
        ++used;
      }
    }
  }
  CHECK(used + 1 == map.size(), "plan holds %zu updates, want %zu", map.size() - 1, used);
}

static void check_pool() {
  for (int nthr : {1, 2, 3, 8}) {
    const int n = 997;
    std::vector<int> hits(n, 0);
    std::atomic<int> started{0};
    run_pool(n, nthr, [&](int i) { hits[i] += 1; }, [&] { started.fetch_add(1); });
    bool once = true;
    for (int i = 0; i < n; ++i) once = once && hits[i] == 1;
    CHECK(once, "run_pool(%d threads): every item once", nthr);
    CHECK(nthr <= 1 ? started == 0 : started == std::min(nthr, n), "run_pool(%d threads): %d starts", nthr,
          started.load());
    bool thrown = false;
    std::vector<int> ran(n, 0);
    try {
      run_pool(n, nthr, [&](int i) {
        ran[i] = 1;
        if (i == 37) BPF_THROW(BPF_E_INVALID, "item %d failed", i);
      }, [] {});
    } catch (const Error &e) {
      thrown = e.code == BPF_E_INVALID && std::strcmp(last_error(), "item 37 failed") == 0;
    }
    CHECK(thrown, "run_pool(%d threads): the error of item 37 reaches the caller (message \"%s\")", nthr,
          last_error());
    CHECK(ran[37] == 1, "run_pool(%d threads): item 37 ran", nthr);
  }
}

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 6;
  check_singular();
  check_pool();
  const int sizes[] = {0, 1, 2, 5, 17, 64, 150, 700, 3000};
  unsigned seed = 11;
  for (int r = 0; r < rounds; ++r)
    for (int n : sizes) {
      ++seed;
      const Csr A = random_csr(n, 1 + (r % 4), seed);
      check_graph_orderings(A);
      if (n > 0 && n <= 700) check_host_lu(A);   // random patterns fill in almost densely
      check_triops(A, seed);
      check_elim_plan(A);
    }
  printf("host_check: %d checks, %d failures\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
