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
// Explicit inverses of the factors of S (after dense_lu_factor, m a multiple
// of kDluNb): Zl = L^{-1} (unit lower), Zu = U^{-1} (upper), row-major m x m,
// by recursive doubling over the diagonal-block inverses Linv / Uinv.  The
// strictly upper blocks of Zl and strictly lower blocks of Zu are never
// written (the caller zeroes them once); T is m x m scratch.
void dense_lu_invert(hipStream_t s, int m, const double *S, const double *Linv, const double *Uinv,
                     double *Zl, double *Zu, double *T);
// X (m x p row-major, overwritten) -> U^{-1} L^{-1} X by two block-triangular
// products; Y is m x p scratch.  With build, Zl and Zu are first rebuilt from
// the factor S and its diagonal-block inverses (dense_lu_invert).
void dense_lu_solve_inv(hipStream_t s, int m, int p, const double *S, const double *Linv, const double *Uinv,
                        bool build, double *Zl, double *Zu, double *T, double *X, double *Y);
// X (m x p row-major, overwritten) -> solution in X; Y is m x p scratch.
void dense_lu_solve(hipStream_t s, int m, int p, const double *S, const double *Linv, const double *Uinv,
                    double *X, double *Y);

}  // namespace bpf
