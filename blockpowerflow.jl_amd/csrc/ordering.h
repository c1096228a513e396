// Fill-reducing orderings and the host sparse LU behind the refactorization.
#pragma once

#include <vector>

namespace bpf {

struct Graph {
  int n = 0;
  std::vector<int> ptr, adj;   // adjacency of A + A^T without the diagonal
};

Graph symmetric_graph(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx);
std::vector<int> rcm_order(const Graph &G);
std::vector<int> mindeg_order(const Graph &G);
std::vector<int> nd_order(const Graph &G, int leaf = 64);

// Result of the host analysis: B = A[P, Q] = L U with L unit lower (strict
// part stored) and U upper; combined CSR pattern/values of L\U by rows of B.
struct HostLU {
  int n = 0;
  std::vector<int> P, Q;               // B[i][j] = A[P[i]][Q[j]]
  std::vector<int> frow, fcol;         // combined L\U pattern (sorted rows)
  std::vector<double> fval;
  int zero_pivot = -1;                 // first pivot with |u_kk| <= tol (row of B), or -1
};

// Threshold-pivoted left-looking sparse LU (Gilbert & Peierls 1988) of
// A[Qs, Qs] with partial pivoting (threshold 1 = largest magnitude, as the
// reference's pivot_threshold = 1.0), Qs the fill-reducing ordering.
HostLU host_lu(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
               const std::vector<double> &val, const std::vector<int> &Qs, double tol,
               double pivot_threshold = 1.0);

}  // namespace bpf
