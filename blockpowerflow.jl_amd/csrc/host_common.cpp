// Host-only analysis: error messages, level sets of triangular operators
// (Anderson & Saad 1989 level scheduling) and elimination plans.  Uses no
// HIP.
#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>

#include "host_common.h"

namespace bpf {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

void elim_plan_host(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
                    const std::vector<int> &diagpos, std::vector<PivInfo> &pinfo,
                    std::vector<int> &map) {
  const int nnz = rowptr[n];
  pinfo.assign(std::max(1, nnz), PivInfo{0, 0, 0, 0});
  size_t total = 0;
  for (int i = 0; i < n; ++i)
    for (int e = rowptr[i]; e < diagpos[i]; ++e) {
      const int k = colidx[e];
      total += (size_t)(rowptr[k + 1] - diagpos[k] - 1);
    }
  BPF_REQUIRE(total < (size_t)INT32_MAX - 256, "elimination plan of %zu updates too large", total);
  map.clear();
  map.reserve(total + 1);
  for (int i = 0; i < n; ++i) {
    const int r0 = rowptr[i], r1 = rowptr[i + 1];
    for (int e = r0; e < diagpos[i]; ++e) {
      const int k = colidx[e];
      pinfo[e] = PivInfo{diagpos[k], rowptr[k + 1], (int)map.size(), k};
      // row i's columns after position e are the ones above column k
      const int *b = colidx.data() + e + 1, *end = colidx.data() + r1;
      for (int j = diagpos[k] + 1; j < rowptr[k + 1]; ++j) {
        const int *f = std::lower_bound(b, end, colidx[j]);
        map.push_back(f != end && *f == colidx[j] ? (int)(f - colidx.data()) - r0 : -1);
      }
    }
  }
  map.push_back(-1);   // a pivot row without upper entries still reads one entry
}

void analyze_triop(TriLevels &op, int n, const std::vector<int> &rowptr,
                   const std::vector<int> &colidx, int kind, const TriWindow &win) {
  const int row_lo = win.row_lo, row_hi = win.row_hi < 0 ? n : win.row_hi;
  const int col_lo = win.col_lo, col_hi = win.col_hi < 0 ? n : win.col_hi;
  BPF_REQUIRE(kind != 3 || (row_lo == 0 && row_hi == n && col_lo == 0 && col_hi == n),
              "windowed transpose operator");
  op.unit = (kind == 0) || win.unit;
  op.lower = (kind == 0 || kind == 2);
  auto keep = [&](int i, int j) {
    if (j < col_lo || j >= col_hi) return false;
    if (kind == 0 || kind == 2) return j < i;
    if (kind == 1) return j > i;
    return j < i;
  };
  // Per-row entry lists (positions into the combined array) and diagonals.
  std::vector<int> dpos(n, -1);
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < n; ++i) {
    const bool in_rows = kind == 3 || (i >= row_lo && i < row_hi);
    if (!in_rows) continue;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      if (j == i) dpos[i] = k;
      if (!keep(i, j)) continue;
      if (kind == 3) ++cnt[j];   // transpose: entry (i,j) -> row j
      else ++cnt[i];
    }
  }
  if (!op.unit)
    for (int i = row_lo; i < row_hi; ++i)
      if (dpos[i] < 0) BPF_THROW(BPF_E_SINGULAR, "structurally zero diagonal in row %d", i);
  // Row lists in original numbering (CSR of the operator).
  std::vector<int> optr(n + 1, 0);
  for (int i = 0; i < n; ++i) optr[i + 1] = optr[i] + cnt[i];
  std::vector<int> ocol(optr[n]), opos(optr[n]);
  std::vector<int> fill(optr.begin(), optr.end() - 1);
  for (int i = 0; i < n; ++i) {
    const bool in_rows = kind == 3 || (i >= row_lo && i < row_hi);
    if (!in_rows) continue;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      if (!keep(i, j)) continue;
      if (kind == 3) { ocol[fill[j]] = i; opos[fill[j]++] = k; }
      else { ocol[fill[i]] = j; opos[fill[i]++] = k; }
    }
  }
  // Rows of the operator: the window's rows (all rows for a transpose),
  // optionally without rows that have no entries.
  std::vector<char> member(n, 0);
  for (int i = 0; i < n; ++i) {
    const bool in_rows = kind == 3 || (i >= row_lo && i < row_hi);
    member[i] = in_rows && !(win.skip_empty && cnt[i] == 0);
  }
  // Levels; dependencies outside the operator's rows are already solved.
  std::vector<int> level(n, -1);
  int nlev = 0;
  auto lev = [&](int i) {
    int lv = 0;
    for (int k = optr[i]; k < optr[i + 1]; ++k)
      if (member[ocol[k]]) lv = std::max(lv, level[ocol[k]] + 1);
    level[i] = lv;
    nlev = std::max(nlev, lv + 1);
  };
  if (op.lower) {
    for (int i = 0; i < n; ++i)
      if (member[i]) lev(i);
  } else {
    for (int i = n - 1; i >= 0; --i)
      if (member[i]) lev(i);
  }
  int nrows = 0;
  for (int i = 0; i < n; ++i) nrows += member[i];
  op.n = nrows;
  op.ndim = n;
  if (nrows == 0) nlev = 0;
  op.nlevels = nlev;
  op.h_lvl.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i)
    if (member[i]) op.h_lvl[level[i] + 1]++;
  for (int l = 0; l < nlev; ++l) op.h_lvl[l + 1] += op.h_lvl[l];
  op.h_order.assign(nrows, 0);
  {
    std::vector<int> f(op.h_lvl.begin(), op.h_lvl.end() - 1);
    for (int i = 0; i < n; ++i)
      if (member[i]) op.h_order[f[level[i]]++] = i;
  }
  const std::vector<int> *rl = win.relabel;
  if (rl) {
    BPF_REQUIRE((int)rl->size() == n, "relabelling of %zu rows for an operator on %d", rl->size(), n);
    // each level's rows in storage order, so neighbouring rows (and the rows
    // they gather) are near each other in memory
    for (int l = 0; l < nlev; ++l)
      std::sort(op.h_order.begin() + op.h_lvl[l], op.h_order.begin() + op.h_lvl[l + 1],
                [&](int a, int b) { return (*rl)[a] < (*rl)[b]; });
  }
  n = nrows;
  op.h_ptr.assign(n + 1, 0);
  op.max_row_nnz = 0;
  for (int t = 0; t < n; ++t) {
    const int i = op.h_order[t];
    const int len = optr[i + 1] - optr[i];
    op.h_ptr[t + 1] = op.h_ptr[t] + len;
    op.max_row_nnz = std::max(op.max_row_nnz, len);
  }
  op.h_col.resize(op.h_ptr[n]);
  op.h_gather.resize(op.h_ptr[n]);
  op.h_dgather.assign(n, 0);
  for (int t = 0; t < n; ++t) {
    const int i = op.h_order[t];
    int o = op.h_ptr[t];
    for (int k = optr[i]; k < optr[i + 1]; ++k, ++o) {
      op.h_col[o] = rl ? (*rl)[ocol[k]] : ocol[k];
      op.h_gather[o] = opos[k];
    }
    op.h_dgather[t] = dpos[i] < 0 ? 0 : dpos[i];
  }
  if (rl)
    for (int t = 0; t < n; ++t) op.h_order[t] = (*rl)[op.h_order[t]];
}

}  // namespace bpf
