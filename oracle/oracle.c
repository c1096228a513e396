/*
 * CPU oracle for the multi-RHS solve path.  TEST INFRASTRUCTURE ONLY: this
 * library is loaded by tests/, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py, as the checker.  The product never links or calls
 * it.
 *
 * Plain C restatements of the sparse kernels the reference reaches through its
 * vendor libraries, written from the published algorithms:
 *   - CSR SpMM               mul!(W, A, P)           src/block_gmres.jl:299
 *   - ILU(0), IKJ variant    ilu02(A) / ilu0(A,'O')  src/preconditioners_nvidia.jl:43,
 *                            src/preconditioners_amd.jl:44   (Saad, Alg. 10.4)
 *   - IC(0)                  ic02(A) / ic0(A,'O')    src/preconditioners_nvidia.jl:7,
 *                            src/preconditioners_amd.jl:6
 *   - triangular solves on the combined factor
 *                            ldiv!(y, ilu, x)        src/preconditioners_nvidia.jl:55-58
 *                            ldiv!(y, ic, x)         src/preconditioners_nvidia.jl:19-22
 *   - fixed-pattern LU refactorization with fixed row/column permutations
 *                            cusolverRfResetValues + cusolverRfRefactor
 *                                                    src/cusolverRF.jl:304-314
 *
 * Dense blocks are column-major n x p with a leading dimension, as in the
 * reference (Julia Matrix).  Indices are 0-based int32; column indices within a
 * CSR row are sorted ascending.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   spmm      - scipy CSR product; bitwise against the row-sequential sum.
 *   ilu0      - (L U)_ij = a_ij on pattern(A) for all (i,j); equals the exact LU
 *               when pattern(A) is closed under fill (tridiagonal, dense).
 *   ic0       - (L L^T)_ij = a_ij on the lower pattern; equals Cholesky on
 *               dense/tridiagonal SPD matrices (numpy.linalg.cholesky).
 *   trsm_*    - scipy.sparse.linalg.spsolve_triangular.
 *   lu_refactor - L U = P A Q entrywise; solution against scipy spsolve; the
 *               reference's own known answers (values x2 -> solution x0.5).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Y = A X, X and Y column-major with leading dimensions ldx/ldy. */
int orc_spmm(int n, const int *rowptr, const int *colidx, const double *val,
             int p, const double *X, int ldx, double *Y, int ldy) {
  for (int c = 0; c < p; ++c) {
    const double *x = X + (size_t)c * ldx;
    double *y = Y + (size_t)c * ldy;
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += val[k] * x[colidx[k]];
      y[i] = acc;
    }
  }
  return 0;
}

/* Position of the diagonal in each row, or -1. */
static int find_diag(int n, const int *rowptr, const int *colidx, int *diag) {
  int missing = 0;
  for (int i = 0; i < n; ++i) {
    diag[i] = -1;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colidx[k] == i) { diag[i] = k; break; }
    if (diag[i] < 0) missing = 1;
  }
  return missing;
}

/*
 * ILU(0) on the pattern of A (IKJ, Saad Alg. 10.4).  out receives the combined
 * factor: strict lower part = L (unit diagonal implied), upper part with
 * diagonal = U, on A's pattern -- the storage convention of the reference's
 * ILU0 object (src/preconditioners_nvidia.jl:50-51).
 * Returns 0, or -(i+1) if row i has a missing or zero pivot.
 */
int orc_ilu0(int n, const int *rowptr, const int *colidx, const double *val, double *out) {
  int nnz = rowptr[n];
  memcpy(out, val, sizeof(double) * (size_t)nnz);
  int *diag = (int *)malloc(sizeof(int) * (size_t)n);
  int *pos = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) pos[i] = -1;
  if (find_diag(n, rowptr, colidx, diag)) {
    for (int i = 0; i < n; ++i)
      if (diag[i] < 0) { free(diag); free(pos); return -(i + 1); }
  }
  int rc = 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) pos[colidx[k]] = k;
    for (int kk = rowptr[i]; kk < rowptr[i + 1]; ++kk) {
      int k = colidx[kk];
      if (k >= i) break;
      double piv = out[diag[k]];
      if (piv == 0.0) { rc = -(k + 1); break; }
      double lik = out[kk] / piv;
      out[kk] = lik;
      for (int jj = diag[k] + 1; jj < rowptr[k + 1]; ++jj) {
        int p = pos[colidx[jj]];
        if (p >= 0) out[p] -= lik * out[jj];
      }
    }
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) pos[colidx[k]] = -1;
    if (rc == 0 && out[diag[i]] == 0.0) rc = -(i + 1);
  }
  free(diag);
  free(pos);
  return rc;
}

/*
 * IC(0) on the lower pattern of a symmetric A (row-oriented, left-looking).
 * out has A's pattern; its lower part with diagonal receives L (A = L L^T on
 * the lower pattern); its strict upper part is copied from A unchanged.
 * Returns 0, or -(i+1) if the pivot of row i is not positive.
 */
int orc_ic0(int n, const int *rowptr, const int *colidx, const double *val, double *out) {
  int nnz = rowptr[n];
  memcpy(out, val, sizeof(double) * (size_t)nnz);
  int *diag = (int *)malloc(sizeof(int) * (size_t)n);
  double *w = (double *)calloc((size_t)n, sizeof(double));
  char *mark = (char *)calloc((size_t)n, 1);
  if (find_diag(n, rowptr, colidx, diag)) {
    for (int i = 0; i < n; ++i)
      if (diag[i] < 0) { free(diag); free(w); free(mark); return -(i + 1); }
  }
  int rc = 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    /* scatter the (already final) entries l_ij, j < i, as they are computed */
    for (int kk = rowptr[i]; kk < diag[i]; ++kk) {
      int k = colidx[kk];
      double s = out[kk];
      /* s -= sum_{j<k} l_ij l_kj over the pattern */
      for (int jj = rowptr[k]; jj < diag[k]; ++jj) {
        int j = colidx[jj];
        if (mark[j]) s -= w[j] * out[jj];
      }
      double lik = s / out[diag[k]];
      out[kk] = lik;
      w[k] = lik;
      mark[k] = 1;
    }
    double d = out[diag[i]];
    for (int kk = rowptr[i]; kk < diag[i]; ++kk) d -= out[kk] * out[kk];
    if (!(d > 0.0)) rc = -(i + 1);
    else out[diag[i]] = sqrt(d);
    for (int kk = rowptr[i]; kk < diag[i]; ++kk) { w[colidx[kk]] = 0.0; mark[colidx[kk]] = 0; }
  }
  free(diag);
  free(w);
  free(mark);
  return rc;
}

/*
 * Triangular solves on a combined CSR factor F (sorted columns), p columns.
 * mode 0: unit lower     (strict lower part of F, implicit unit diagonal)
 * mode 1: upper          (upper part of F with diagonal)
 * mode 2: lower          (lower part of F with diagonal)
 * mode 3: lower^T        (solve L^T y = x using the lower part of F)
 * Y may alias X.
 */
int orc_trsm(int mode, int n, const int *rowptr, const int *colidx, const double *val,
             int p, const double *X, int ldx, double *Y, int ldy) {
  for (int c = 0; c < p; ++c) {
    const double *x = X + (size_t)c * ldx;
    double *y = Y + (size_t)c * ldy;
    if (y != x) memcpy(y, x, sizeof(double) * (size_t)n);
    if (mode == 0 || mode == 2) {
      for (int i = 0; i < n; ++i) {
        double acc = y[i], d = 1.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          int j = colidx[k];
          if (j < i) acc -= val[k] * y[j];
          else if (j == i) d = val[k];
        }
        y[i] = mode == 0 ? acc : acc / d;
      }
    } else if (mode == 1) {
      for (int i = n - 1; i >= 0; --i) {
        double acc = y[i], d = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          int j = colidx[k];
          if (j > i) acc -= val[k] * y[j];
          else if (j == i) d = val[k];
        }
        if (d == 0.0) return -(i + 1);
        y[i] = acc / d;
      }
    } else if (mode == 3) {
      /* column-oriented backward sweep over rows of L (= columns of L^T) */
      for (int i = n - 1; i >= 0; --i) {
        double d = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
          if (colidx[k] == i) d = val[k];
        if (d == 0.0) return -(i + 1);
        double yi = y[i] / d;
        y[i] = yi;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          int j = colidx[k];
          if (j < i) y[j] -= val[k] * yi;
        }
      }
    } else {
      return -1000000;
    }
  }
  return 0;
}

/*
 * Fixed-pattern LU refactorization (what cusolverRfResetValues + Refactor do,
 * src/cusolverRF.jl:304-314): given B = P A Q with P = rowperm, Q = colperm
 * (B[i][j] = A[P[i]][Q[j]]) and the combined pattern F of L (strict lower,
 * unit diagonal) + U (upper with diagonal) from the symbolic analysis, compute
 * the factor values with the same pivot sequence.  The pattern F must contain
 * the full fill of B; the computation is then row-wise Doolittle/IKJ on F,
 * which is exact LU.
 * A: CSR (rowptr, colidx, aval).  F: CSR (frow, fcol) -> fval (output).
 * Returns 0, -(i+1) on a zero pivot, or -1000000-(i+1) if an entry of B falls
 * outside F.
 */
int orc_lu_refactor(int n, const int *rowptr, const int *colidx, const double *aval,
                    const int *P, const int *Q,
                    const int *frow, const int *fcol, double *fval) {
  int *qinv = (int *)malloc(sizeof(int) * (size_t)n);
  int *pos = (int *)malloc(sizeof(int) * (size_t)n);
  int *fdiag = (int *)malloc(sizeof(int) * (size_t)n);
  for (int j = 0; j < n; ++j) { qinv[Q[j]] = j; pos[j] = -1; }
  int rc = find_diag(n, frow, fcol, fdiag) ? -1 : 0;
  for (int i = 0; i < n && rc == 0; ++i) {
    for (int k = frow[i]; k < frow[i + 1]; ++k) { pos[fcol[k]] = k; fval[k] = 0.0; }
    int r = P[i];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      int j = qinv[colidx[k]];
      if (pos[j] < 0) { rc = -1000000 - (i + 1); break; }
      fval[pos[j]] += aval[k];
    }
    for (int kk = frow[i]; kk < fdiag[i] && rc == 0; ++kk) {
      int k = fcol[kk];
      double piv = fval[fdiag[k]];
      if (piv == 0.0) { rc = -(k + 1); break; }
      double lik = fval[kk] / piv;
      fval[kk] = lik;
      for (int jj = fdiag[k] + 1; jj < frow[k + 1]; ++jj) {
        int p = pos[fcol[jj]];
        if (p < 0) { rc = -1000000 - (i + 1); break; }
        fval[p] -= lik * fval[jj];
      }
    }
    for (int k = frow[i]; k < frow[i + 1]; ++k) pos[fcol[k]] = -1;
    if (rc == 0 && fval[fdiag[i]] == 0.0) rc = -(i + 1);
  }
  free(qinv);
  free(pos);
  free(fdiag);
  return rc;
}
