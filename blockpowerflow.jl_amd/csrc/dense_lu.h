// Dense tail of the fixed-pivot LU refactorization (see dense_lu.hip).
#pragma once

#include "bpf_internal.h"

namespace bpf {

// Column width of the blocked dense LU and of its triangular-solve blocks.
constexpr int kDluNb = 128;

// Moves the cnt tail entries between the sparse factor values (fac[pos[e]])
// and the dense m x m matrix (S[didx[e]]).
void dense_tail_move(hipStream_t s, int cnt, const int *pos, const int *didx, double *fac, double *S,
                     bool to_dense);
// In-place LU without pivoting of the row-major m x m matrix S.  Linv/Uinv
// receive the inverses of the diagonal kDluNb blocks (ceil(m/kDluNb) blocks
// of kDluNb^2 doubles each); P (m x kDluNb) and Q (kDluNb x m) are scratch.
// A zero pivot at row r sets *status = min(*status, base + r).
void dense_lu_factor(hipStream_t s, int m, double *S, double *Linv, double *Uinv, double *P, double *Q,
                     int *status, int base);
// X (m x p row-major, overwritten) -> solution in X; Y is m x p scratch.
void dense_lu_solve(hipStream_t s, int m, int p, const double *S, const double *Linv, const double *Uinv,
                    double *X, double *Y);

}  // namespace bpf
