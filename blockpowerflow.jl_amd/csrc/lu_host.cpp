// Host sparse LU with partial pivoting (left-looking, Gilbert & Peierls 1988;
// column-by-column sparse triangular solves with a depth-first reach on the
// graph of L).  This is the analysis half of the refactorization path: it
// fixes P, Q and the L\U pattern that the device refactorization reuses,
// standing in for cusolverSpXcsrluAnalysisHost / DcsrluFactorHost /
// DcsrluExtractHost (src/cusolverRF.jl:127-204).
#include <algorithm>
#include <cmath>

#include "host_common.h"
#include "ordering.h"

namespace bpf {

HostLU host_lu(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
               const std::vector<double> &val, const std::vector<int> &Qs, double tol,
               double pivot_threshold) {
  // Columns of A (CSC) for column access.
  std::vector<int> cptr(n + 1, 0), crow(rowptr[n]);
  std::vector<double> cval(rowptr[n]);
  for (int k = 0; k < rowptr[n]; ++k) cptr[colidx[k] + 1]++;
  for (int j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  {
    std::vector<int> f(cptr.begin(), cptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int j = colidx[k], pos = f[j]++;
        crow[pos] = i;
        cval[pos] = val[k];
      }
  }
  // Symmetric preordering: B0 = A[Qs, Qs]; row r of B0 is row Qs[r] of A.
  std::vector<int> qinv(n);
  for (int k = 0; k < n; ++k) qinv[Qs[k]] = k;
  // L stored by columns with row indices in B0 numbering; U by columns with
  // row indices = pivot step k.
  std::vector<int> Lp(1, 0), Li, Up(1, 0), Ui;
  std::vector<double> Lx, Ux;
  Li.reserve(4 * (size_t)rowptr[n]);
  Lx.reserve(4 * (size_t)rowptr[n]);
  Ui.reserve(4 * (size_t)rowptr[n]);
  Ux.reserve(4 * (size_t)rowptr[n]);
  std::vector<int> pinv(n, -1);          // B0 row -> pivot step
  std::vector<double> x(n, 0.0);
  std::vector<int> xi(n), stack(n), pstack(n);
  std::vector<char> mark(n, 0);
  HostLU R;
  R.n = n;
  for (int j = 0; j < n; ++j) {
    // --- reach of column j of B0 in the graph of L (DFS, topological order)
    const int acol = Qs[j];
    int top = n;
    for (int p = cptr[acol]; p < cptr[acol + 1]; ++p) {
      const int r0 = qinv[crow[p]];
      if (mark[r0]) continue;
      int head = 0;
      stack[0] = r0;
      while (head >= 0) {
        const int r = stack[head];
        const int k = pinv[r];           // column of L to follow (if pivotal)
        if (!mark[r]) {
          mark[r] = 1;
          pstack[head] = k < 0 ? 0 : Lp[k];
        }
        bool done = true;
        const int pend = k < 0 ? 0 : Lp[k + 1];
        for (int q = pstack[head]; q < pend; ++q) {
          const int r2 = Li[q];
          if (mark[r2]) continue;
          pstack[head] = q + 1;
          stack[++head] = r2;
          done = false;
          break;
        }
        if (done) {
          --head;
          xi[--top] = r;
        }
      }
    }
    for (int q = top; q < n; ++q) mark[xi[q]] = 0;
    // --- numeric sparse triangular solve x = L \ B0[:, j]
    for (int q = top; q < n; ++q) x[xi[q]] = 0.0;
    for (int p = cptr[acol]; p < cptr[acol + 1]; ++p) x[qinv[crow[p]]] += cval[p];
    for (int q = top; q < n; ++q) {
      const int r = xi[q];
      const int k = pinv[r];
      if (k < 0) continue;
      const double xr = x[r];
      for (int t = Lp[k]; t < Lp[k + 1]; ++t) x[Li[t]] -= Lx[t] * xr;
    }
    // --- pivot: largest magnitude among non-pivotal rows, diagonal preferred
    // when it is within the threshold.
    int ipiv = -1;
    double amax = -1.0;
    for (int q = top; q < n; ++q) {
      const int r = xi[q];
      if (pinv[r] < 0 && std::fabs(x[r]) > amax) { amax = std::fabs(x[r]); ipiv = r; }
    }
    if (ipiv < 0) BPF_THROW(BPF_E_SINGULAR, "structurally singular matrix at column %d", j);
    if (pinv[j] < 0 && std::fabs(x[j]) >= pivot_threshold * amax) ipiv = j;
    const double piv = x[ipiv];
    if (std::fabs(piv) <= tol && R.zero_pivot < 0) R.zero_pivot = j;
    if (piv == 0.0) BPF_THROW(BPF_E_SINGULAR, "zero pivot at column %d", j);
    pinv[ipiv] = j;
    // --- U(:, j) = x at pivotal rows; L(:, j) = x / piv at the others
    for (int q = top; q < n; ++q) {
      const int r = xi[q];
      const int k = pinv[r];
      if (k >= 0 && k <= j) {
        Ui.push_back(k);
        Ux.push_back(x[r]);
      } else if (k < 0) {
        Li.push_back(r);
        Lx.push_back(x[r] / piv);
      }
      x[r] = 0.0;
    }
    Lp.push_back((int)Li.size());
    Up.push_back((int)Ui.size());
  }
  // Map L row indices to pivot order.
  for (auto &r : Li) r = pinv[r];
  // Permutations: B = A[P, Q] with P[k] = original row pivoted at step k.
  R.P.resize(n);
  R.Q.resize(n);
  for (int r = 0; r < n; ++r) R.P[pinv[r]] = Qs[r];
  for (int k = 0; k < n; ++k) R.Q[k] = Qs[k];
  // Combined CSR by rows of B: row k holds L(k, j<k) and U(k, j>=k).
  std::vector<int> cnt(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    for (int t = Lp[j]; t < Lp[j + 1]; ++t) cnt[Li[t] + 1]++;
    for (int t = Up[j]; t < Up[j + 1]; ++t) cnt[Ui[t] + 1]++;
  }
  for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
  R.frow = cnt;
  R.fcol.resize(cnt[n]);
  R.fval.resize(cnt[n]);
  std::vector<int> f(cnt.begin(), cnt.end() - 1);
  for (int j = 0; j < n; ++j) {   // columns in increasing order -> rows come out sorted
    // within column j the L rows (> j) and U rows (<= j) are appended in
    // column order, so each row's columns are increasing.
    for (int t = Up[j]; t < Up[j + 1]; ++t) {
      const int k = Ui[t];
      R.fcol[f[k]] = j;
      R.fval[f[k]++] = Ux[t];
    }
    for (int t = Lp[j]; t < Lp[j + 1]; ++t) {
      const int k = Li[t];
      R.fcol[f[k]] = j;
      R.fval[f[k]++] = Lx[t];
    }
  }
  return R;
}

}  // namespace bpf
