// Fixed-pattern LU refactorization and multi-RHS direct solves: the
// CUSOLVERRF replacement (src/cusolverRF.jl).
//
//   bpf_rf_create    CusolverRfLU(A; ordering, tol)       :222-301
//                      host: ordering + pivoted sparse LU (glu_analysis_host,
//                      :56-220), then device setup and first refactorization
//   bpf_rf_refactor  update!(rflu, A) / lu!(rflu, A)      :304-314, :334
//                      device: scatter A's values into the L\U pattern of
//                      B = A[P, Q] and run the level-scheduled IKJ elimination
//                      with the pivots fixed
//   bpf_rf_solve     ldiv!(x, rflu, b)                    :317-331 (any nrhs)
//   bpf_rf_batch_*   CusolverRfLUBatch, rf_batch_refactor!, ldiv!   :336-481
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dense_lu.h"
#include "krylov.h"
#include "ordering.h"

using namespace bpf;

namespace bpf {
void launch_scatter_values(hipStream_t s, int nnz, const int *amap, const double *aval, double *fac);
template <typename TD>
void launch_perm_cm_to_rm(hipStream_t s, int n, int p, const double *src, int lds, const int *rowmap,
                          TD *dst, int ldd);
template <typename TS>
void launch_perm_rm_to_cm(hipStream_t s, int n, int p, const TS *src, int lds, const int *rowmap,
                          double *dst, int ldd);
// Row-major rows by a row map: dst[k] = src[map[k]] (gather), dst[map[k]] =
// src[k] (scatter) or, with add, dst[map[k]] += src[k].
template <typename TS, typename TD>
void launch_perm_rows_rm(hipStream_t s, int n, int p, const TS *src, const int *rowmap, TD *dst,
                         bool scatter, bool add = false);
}  // namespace bpf

struct bpf_rf_s {
  bpf_ctx ctx = nullptr;
  int n = 0, nnzA = 0, nnzF = 0;
  int batch = 0;                       // 0: single CusolverRfLU
  std::vector<int> h_rowptr, h_colidx; // pattern of A
  HostLU lu;                           // P, Q, pattern (and first values) of L\U
  DevBuf dP, dQ, frow, fcol, fdiag, fac, amap;
  ElimPlan plan;                       // elimination plan of the fill pattern
  std::vector<int> h_fdiag;
  int max_row = 0;
  TriOp lo, up;
  DevBuf facb;                         // batch: nnzF x batch factor values
  DevBuf w1, w2, w3;                   // work blocks
  DevBuf status, stat;
  // Dense tail (single handles with n >= kDenseMinN): rows/columns [t, n)
  // of B = A[P, Q] are refactored and solved as a dense m x m block.
  int t = 0, m = 0, ntail = 0;
  TriOp lo_h, up_h;                    // head operators (rows and columns < t)
  TriOp l21, u12;                      // couplings: tail rows x head columns, head rows x tail columns
  DevBuf S, Linv, Uinv, P, Q;          // dense tail factor, diagonal-block inverses, scratch
  DevBuf tpos, tdidx;                  // tail entries: factor position, dense index
  DevBuf Xd, Yd;                       // FP64 m x p solve blocks
  DevBuf Zl, Zu, Zt;                   // explicit L^{-1}, U^{-1} of the tail (m x m) and their scratch
  bool zinv = false;                   // Zl, Zu hold the inverses of the current factor values
  bool dense() const { return m > 0; }
  TriOp &lop() { return dense() ? lo_h : lo; }
  TriOp &uop() { return dense() ? up_h : up; }
};

namespace {

void check_pattern(bpf_rf rf, bpf_csr A) {
  BPF_REQUIRE(A->n == rf->n && A->nnz == rf->nnzA && A->h_rowptr == rf->h_rowptr &&
                  A->h_colidx == rf->h_colidx,
              "matrix pattern differs from the analysed pattern");
}

// fac <- values of A scattered into the L\U pattern, then IKJ elimination
// (with a dense tail: IKJ on the head rows, the tail rows eliminated against
// the head only, then a dense LU of their Schur complement).
void refactor_into(bpf_rf rf, const double *d_aval, double *fac) {
  hipStream_t s = rf->ctx->stream;
  fill_double(s, fac, 0.0, rf->nnzF);
  launch_scatter_values(s, rf->nnzA, rf->amap.as<int>(), d_aval, fac);
  rf->status.ensure(16);
  fill_int(s, rf->status.as<int>(), rf->n, 1);
  const int head = rf->dense() ? rf->t : rf->n;
  const std::vector<int> &frow = rf->lu.frow;
  launch_ilu_factor(s, rf->lop(), rf->n, rf->frow.as<int>(), rf->fcol.as<int>(), rf->fdiag.as<int>(),
                    rf->plan, fac,
                    rf->max_row, rf->status.as<int>(), factor_bytes(head, frow[head] - frow[0]));
  if (rf->dense()) {
    launch_partial_factor(s, rf->t, rf->n, rf->t, rf->frow.as<int>(), rf->fcol.as<int>(), rf->fdiag.as<int>(),
                          rf->plan, fac, rf->max_row, rf->status.as<int>(),
                          factor_bytes(rf->n - rf->t, frow[rf->n] - frow[rf->t]));
    double *S = rf->S.as<double>();
    fill_double(s, S, 0.0, (size_t)rf->m * rf->m);
    dense_tail_move(s, rf->ntail, rf->tpos.as<int>(), rf->tdidx.as<int>(), fac, S, true);
    rf->zinv = false;   // rebuilt by the first solve that uses them
    dense_lu_factor(s, rf->m, S, rf->Linv.as<double>(), rf->Uinv.as<double>(), rf->P.as<double>(),
                    rf->Q.as<double>(), rf->status.as<int>(), rf->t);
    dense_tail_move(s, rf->ntail, rf->tpos.as<int>(), rf->tdidx.as<int>(), fac, S, false);
  }
  int st = 0;
  BPF_CHECK_HIP(hipMemcpyAsync(&st, rf->status.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  triop_check(rf->lop(), s);
  if (st < rf->n) BPF_THROW(BPF_E_SINGULAR, "refactorization: zero pivot at step %d", st);
}

// Refreshes the triangular operators' value copies from the factor values.
void refresh_ops(bpf_rf rf, const double *fac) {
  hipStream_t s = rf->ctx->stream;
  triop_refresh(rf->lop(), fac, s);
  triop_refresh(rf->uop(), fac, s);
  if (rf->dense()) {
    triop_refresh(rf->l21, fac, s);
    triop_refresh(rf->u12, fac, s);
  }
}

void check_ops(bpf_rf rf) {
  hipStream_t s = rf->ctx->stream;
  triop_check(rf->lop(), s);
  triop_check(rf->uop(), s);
  if (rf->dense()) {
    triop_check(rf->l21, s);
    triop_check(rf->u12, s);
  }
}

constexpr int kDenseMinN = 16384;

}  // namespace

std::atomic<int> bpf::g_dense_inverse{1};
std::atomic<int> bpf::g_dense_blocks{12};

namespace {

void setup(bpf_ctx ctx, bpf_csr A, int ordering, double tol, bpf_rf rf, bool allow_dense) {
  rf->ctx = ctx;
  rf->n = A->n;
  rf->nnzA = A->nnz;
  rf->h_rowptr = A->h_rowptr;
  rf->h_colidx = A->h_colidx;
  const int n = A->n;
  std::vector<double> hv(A->nnz);
  BPF_CHECK_HIP(hipMemcpyAsync(hv.data(), A->val.ptr, A->nnz * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  // Fill-reducing ordering (src/cusolverRF.jl:70-95).
  std::vector<int> Qs;
  if (ordering == BPF_ORDER_NATURAL) {
    Qs.resize(n);
    for (int i = 0; i < n; ++i) Qs[i] = i;
  } else {
    Graph G = symmetric_graph(n, A->h_rowptr, A->h_colidx);
    if (ordering == BPF_ORDER_RCM) Qs = rcm_order(G);
    else if (ordering == BPF_ORDER_METIS) Qs = nd_order(G);
    else if (ordering == BPF_ORDER_AMD || ordering == BPF_ORDER_MDQ) Qs = mindeg_order(G);
    else BPF_THROW(BPF_E_INVALID, "unknown ordering %d", ordering);
  }
  // Host LU with partial pivoting (threshold 1.0, src/cusolverRF.jl:148).
  rf->lu = host_lu(n, A->h_rowptr, A->h_colidx, hv, Qs, tol, 1.0);
  if (rf->lu.zero_pivot >= 0)
    BPF_THROW(BPF_E_SINGULAR, "singular matrix: pivot %d below tol = %g", rf->lu.zero_pivot, tol);
  const HostLU &lu = rf->lu;
  rf->nnzF = lu.frow[n];
  rf->h_fdiag.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    for (int k = lu.frow[i]; k < lu.frow[i + 1]; ++k)
      if (lu.fcol[k] == i) rf->h_fdiag[i] = k;
    rf->max_row = std::max(rf->max_row, lu.frow[i + 1] - lu.frow[i]);
  }
  // A entry k (row r, col c) -> F position of (pinvP[r], qinv[c]).
  std::vector<int> pinv(n), qinv(n), amap(A->nnz);
  for (int i = 0; i < n; ++i) { pinv[lu.P[i]] = i; qinv[lu.Q[i]] = i; }
  for (int r = 0; r < n; ++r)
    for (int k = A->h_rowptr[r]; k < A->h_rowptr[r + 1]; ++k) {
      const int i = pinv[r], j = qinv[A->h_colidx[k]];
      const int *b = lu.fcol.data() + lu.frow[i], *e = lu.fcol.data() + lu.frow[i + 1];
      const int *f = std::lower_bound(b, e, j);
      if (f == e || *f != j) BPF_THROW(BPF_E_INVALID, "factor pattern misses an entry of A");
      amap[k] = lu.frow[i] + (int)(f - b);
    }
  hipStream_t s = ctx->stream;
  upload(rf->dP, lu.P, s);
  upload(rf->dQ, lu.Q, s);
  upload(rf->frow, lu.frow, s);
  upload(rf->fcol, lu.fcol, s);
  upload(rf->fdiag, rf->h_fdiag, s);
  upload(rf->amap, amap, s);
  build_elim_plan(n, lu.frow, lu.fcol, rf->h_fdiag, s, rf->plan);
  rf->fac.ensure(std::max(1, rf->nnzF) * sizeof(double));
  // Dense tail for large single systems: the last m rows (m a multiple of
  // the dense block width, at most 12 blocks and about n/16).  On the
  // 70,000-bus pattern 12 blocks (1,536 rows) balance the single-workgroup
  // sparse tail against the serial diagonal-block LUs best (8, 16, 24 and 32
  // blocks measured slower).
  if (allow_dense && n >= kDenseMinN) {
    const int nb = kDluNb;
    rf->m = std::min(g_dense_blocks * nb, (n / 16) / nb * nb);
  }
  if (!rf->dense()) {
    build_triop(rf->lo, n, lu.frow, lu.fcol, 0, s);
    build_triop(rf->up, n, lu.frow, lu.fcol, 1, s);
    return;
  }
  const int m = rf->m, t = n - m;
  rf->t = t;
  TriWindow head;
  head.row_hi = t;
  head.col_hi = t;
  build_triop(rf->lo_h, n, lu.frow, lu.fcol, 0, s, head);
  build_triop(rf->up_h, n, lu.frow, lu.fcol, 1, s, head);
  TriWindow w21;
  w21.row_lo = t;
  w21.col_hi = t;
  w21.unit = true;
  w21.skip_empty = true;
  build_triop(rf->l21, n, lu.frow, lu.fcol, 0, s, w21);
  TriWindow w12;
  w12.row_hi = t;
  w12.col_lo = t;
  w12.unit = true;
  w12.skip_empty = true;
  build_triop(rf->u12, n, lu.frow, lu.fcol, 1, s, w12);
  std::vector<int> tpos, tdidx;
  for (int i = t; i < n; ++i)
    for (int k = lu.frow[i]; k < lu.frow[i + 1]; ++k)
      if (lu.fcol[k] >= t) {
        tpos.push_back(k);
        tdidx.push_back((i - t) * m + (lu.fcol[k] - t));
      }
  rf->ntail = (int)tpos.size();
  upload(rf->tpos, tpos, s);
  upload(rf->tdidx, tdidx, s);
  const int nblk = (m + kDluNb - 1) / kDluNb;
  rf->S.ensure((size_t)m * m * sizeof(double));
  rf->Linv.ensure((size_t)nblk * kDluNb * kDluNb * sizeof(double));
  rf->Uinv.ensure((size_t)nblk * kDluNb * kDluNb * sizeof(double));
  rf->P.ensure((size_t)m * kDluNb * sizeof(double));
  rf->Q.ensure((size_t)m * kDluNb * sizeof(double));
}

// Row-major solve of B z = y with the current triangular operators.  With a
// dense tail: head forward sweep, tail rows minus their head couplings, dense
// tail solve (FP64), head rows minus their tail couplings, head backward sweep.
template <typename TV>
void tri_solve(bpf_rf rf, int p, TV *x) {
  hipStream_t s = rf->ctx->stream;
  if (!rf->dense()) {
    trisolve<TV>(s, rf->lo, p, x, p, x, p);
    trisolve<TV>(s, rf->up, p, x, p, x, p);
    return;
  }
  const int t = rf->t, m = rf->m;
  trisolve<TV>(s, rf->lo_h, p, x, p, x, p);
  trisolve<TV>(s, rf->l21, p, x, p, x, p);
  rf->Xd.ensure((size_t)m * p * sizeof(double));
  rf->Yd.ensure((size_t)m * p * sizeof(double));
  TV *xt = x + (size_t)t * p;
  copy_cast<TV, double>(s, m, p, xt, p, rf->Xd.as<double>(), p, 1.0);
  // FP32 solves (refined afterwards) by the explicit inverse factors, FP64
  // solves too with the option at 2; else blocked substitution.  The
  // inverses are built by the first such solve after a refactorization.
  const bool inv = g_dense_inverse == 2 || (g_dense_inverse == 1 && std::is_same<TV, float>::value);
  if (inv) {
    const size_t mm = (size_t)m * m;
    if (!rf->Zl.ptr) {
      rf->Zl.ensure(mm * sizeof(double));
      rf->Zu.ensure(mm * sizeof(double));
      rf->Zt.ensure(mm * sizeof(double));
      fill_double(s, rf->Zl.as<double>(), 0.0, mm);   // strictly upper blocks stay zero
      fill_double(s, rf->Zu.as<double>(), 0.0, mm);   // strictly lower blocks stay zero
    }
    dense_lu_solve_inv(s, m, p, rf->S.as<double>(), rf->Linv.as<double>(), rf->Uinv.as<double>(), !rf->zinv,
                       rf->Zl.as<double>(), rf->Zu.as<double>(), rf->Zt.as<double>(), rf->Xd.as<double>(),
                       rf->Yd.as<double>());
    rf->zinv = true;
  } else
    dense_lu_solve(s, m, p, rf->S.as<double>(), rf->Linv.as<double>(), rf->Uinv.as<double>(), rf->Xd.as<double>(),
                   rf->Yd.as<double>());
  copy_cast<double, TV>(s, m, p, rf->Xd.as<double>(), p, xt, p, 1.0);
  trisolve<TV>(s, rf->u12, p, x, p, x, p);
  trisolve<TV>(s, rf->up_h, p, x, p, x, p);
}

}  // namespace

extern "C" {

int bpf_rf_create(bpf_ctx ctx, bpf_csr A, int ordering, double tol, bpf_rf *out) {
  return guard([&] {
    BPF_REQUIRE(ctx && A && out, "null argument");
    auto *rf = new bpf_rf_s();
    try {
      setup(ctx, A, ordering, tol, rf, true);
      refactor_into(rf, A->val.as<double>(), rf->fac.as<double>());
      refresh_ops(rf, rf->fac.as<double>());
    } catch (...) {
      delete rf;
      throw;
    }
    *out = rf;
  });
}

int bpf_rf_refactor(bpf_rf rf, bpf_csr A) {
  return guard([&] {
    BPF_REQUIRE(rf && A, "null handle");
    BPF_REQUIRE(rf->batch == 0, "use bpf_rf_batch_refactor for a batch handle");
    check_pattern(rf, A);
    refactor_into(rf, A->val.as<double>(), rf->fac.as<double>());
    refresh_ops(rf, rf->fac.as<double>());
  });
}

int bpf_rf_solve(bpf_rf rf, int dtype, int nrhs, const double *d_B, int ldb, double *d_X, int ldx) {
  return guard([&] {
    BPF_REQUIRE(rf && d_B && d_X, "null argument");
    BPF_REQUIRE(nrhs >= 0 && ldb >= rf->n && ldx >= rf->n, "bad sizes");
    BPF_REQUIRE(dtype == BPF_F64 || dtype == BPF_F32, "bad dtype");
    if (nrhs == 0 || rf->n == 0) return;
    hipStream_t s = rf->ctx->stream;
    const int n = rf->n;
    const size_t es = dtype == BPF_F32 ? sizeof(float) : sizeof(double);
    rf->w1.ensure((size_t)n * nrhs * es);
    // z = B[P, :] (row-major), solve, X[Q, :] = z.
    if (dtype == BPF_F64) {
      double *z = rf->w1.as<double>();
      launch_perm_cm_to_rm<double>(s, n, nrhs, d_B, ldb, rf->dP.as<int>(), z, nrhs);
      tri_solve<double>(rf, nrhs, z);
      launch_perm_rm_to_cm<double>(s, n, nrhs, z, nrhs, rf->dQ.as<int>(), d_X, ldx);
    } else {
      float *z = rf->w1.as<float>();
      launch_perm_cm_to_rm<float>(s, n, nrhs, d_B, ldb, rf->dP.as<int>(), z, nrhs);
      tri_solve<float>(rf, nrhs, z);
      launch_perm_rm_to_cm<float>(s, n, nrhs, z, nrhs, rf->dQ.as<int>(), d_X, ldx);
    }
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    check_ops(rf);
  });
}

int bpf_rf_solve_refined(bpf_rf rf, bpf_csr A, int nrhs, const double *d_B, int ldb, double *d_X,
                         int ldx, int refine, double berr_tol, double *h_berr, int *h_steps) {
  return guard([&] {
    BPF_REQUIRE(rf && A && d_B && d_X, "null argument");
    check_pattern(rf, A);
    BPF_REQUIRE(nrhs >= 0 && ldb >= rf->n && ldx >= rf->n && refine >= 0, "bad sizes");
    BPF_REQUIRE(berr_tol >= 0.0, "berr_tol must be >= 0");
    if (h_steps) *h_steps = 0;
    if (nrhs == 0 || rf->n == 0) return;
    hipStream_t s = rf->ctx->stream;
    const int n = rf->n, p = nrhs;
    const size_t m = (size_t)n * p;
    rf->w1.ensure(m * sizeof(double));   // B row-major
    rf->w2.ensure(m * sizeof(double));   // X row-major
    rf->w3.ensure(m * (sizeof(double) + sizeof(float)));   // R (FP64) + z (FP32)
    double *B = rf->w1.as<double>(), *X = rf->w2.as<double>(), *R = rf->w3.as<double>();
    float *z = reinterpret_cast<float *>(R + m);
    launch_cm_to_rm(s, n, p, d_B, ldb, B, p, false);
    // X0 = FP32 solve of B; then up to `refine` steps X += solve32(B - A X).
    // With berr_tol > 0 the backward error of every column is checked on each
    // residual, and the loop stops once all of them are <= berr_tol.
    launch_perm_rows_rm<double, float>(s, n, p, B, rf->dP.as<int>(), z, false);
    tri_solve<float>(rf, p, z);
    launch_perm_rows_rm<float, double>(s, n, p, z, rf->dQ.as<int>(), X, true);
    const bool adaptive = berr_tol > 0.0;
    std::vector<double> be(p, 0.0);
    bool bnorms = false;
    int steps = 0;
    for (int it = 0; it <= refine; ++it) {
      const bool last = it == refine;
      const bool want_berr = adaptive || (last && h_berr);
      if (last && !want_berr) break;
      residual_rm(s, A, p, B, X, R);
      if (want_berr) {
        backward_error_rm(rf->ctx, A, p, R, X, B, be.data(), rf->stat, bnorms);
        bnorms = true;   // B's column norms stay in rf->stat for the later residuals
        if (adaptive && std::all_of(be.begin(), be.end(), [&](double e) { return e <= berr_tol; })) break;
      }
      if (last) break;
      launch_perm_rows_rm<double, float>(s, n, p, R, rf->dP.as<int>(), z, false);
      tri_solve<float>(rf, p, z);
      launch_perm_rows_rm<float, double>(s, n, p, z, rf->dQ.as<int>(), X, true, true);   // X[Q[k]] += z[k]
      ++steps;
    }
    if (h_berr) std::copy(be.begin(), be.end(), h_berr);
    if (h_steps) *h_steps = steps;
    launch_rm_to_cm(s, n, p, X, p, false, d_X, ldx);
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    check_ops(rf);
  });
}

int bpf_rf_info(bpf_rf rf, int *n, int *nnz_factor, int *nl, int *nu) {
  return guard([&] {
    BPF_REQUIRE(rf, "null handle");
    if (n) *n = rf->n;
    if (nnz_factor) *nnz_factor = rf->nnzF;
    if (nl) *nl = rf->lop().nlevels;
    if (nu) *nu = rf->uop().nlevels;
  });
}

int bpf_rf_get_factor(bpf_rf rf, int32_t *h_P, int32_t *h_Q, int32_t *h_rowptr, int32_t *h_colidx,
                      double *h_val) {
  return guard([&] {
    BPF_REQUIRE(rf, "null handle");
    if (h_P) std::memcpy(h_P, rf->lu.P.data(), rf->n * sizeof(int));
    if (h_Q) std::memcpy(h_Q, rf->lu.Q.data(), rf->n * sizeof(int));
    if (h_rowptr) std::memcpy(h_rowptr, rf->lu.frow.data(), (rf->n + 1) * sizeof(int));
    if (h_colidx) std::memcpy(h_colidx, rf->lu.fcol.data(), rf->nnzF * sizeof(int));
    if (h_val) {
      BPF_CHECK_HIP(hipMemcpyAsync(h_val, rf->fac.ptr, rf->nnzF * sizeof(double), hipMemcpyDeviceToHost,
                                   rf->ctx->stream));
      BPF_CHECK_HIP(hipStreamSynchronize(rf->ctx->stream));
    }
  });
}

int bpf_rf_destroy(bpf_rf rf) {
  return guard([&] {
    if (!rf) return;
    (void)hipStreamSynchronize(rf->ctx->stream);
    delete rf;
  });
}

int bpf_rf_batch_create(bpf_ctx ctx, bpf_csr A, int batchsize, int ordering, double tol, bpf_rf *out) {
  return guard([&] {
    BPF_REQUIRE(ctx && A && out, "null argument");
    BPF_REQUIRE(batchsize > 1,
                "Please specify a number of batch greater than 1 for `CuSolverRfLUBatch`");
    auto *rf = new bpf_rf_s();
    try {
      setup(ctx, A, ordering, tol, rf, false);
      rf->batch = batchsize;
      rf->facb.ensure((size_t)rf->nnzF * batchsize * sizeof(double));
      // every batch entry starts from A's values (src/cusolverRF.jl:399-413)
      for (int b = 0; b < batchsize; ++b)
        refactor_into(rf, A->val.as<double>(), rf->facb.as<double>() + (size_t)b * rf->nnzF);
    } catch (...) {
      delete rf;
      throw;
    }
    *out = rf;
  });
}

int bpf_rf_batch_refactor(bpf_rf rf, const double *d_nzval, int ld) {
  return guard([&] {
    BPF_REQUIRE(rf && d_nzval, "null argument");
    BPF_REQUIRE(rf->batch > 1, "not a batch handle");
    BPF_REQUIRE(ld >= rf->nnzA, "leading dimension < nnz");
    for (int b = 0; b < rf->batch; ++b)
      refactor_into(rf, d_nzval + (size_t)b * ld, rf->facb.as<double>() + (size_t)b * rf->nnzF);
  });
}

int bpf_rf_batch_solve(bpf_rf rf, const double *d_B, int ldb, double *d_X, int ldx) {
  return guard([&] {
    BPF_REQUIRE(rf && d_B && d_X, "null argument");
    BPF_REQUIRE(rf->batch > 1, "not a batch handle");
    BPF_REQUIRE(ldb >= rf->n && ldx >= rf->n, "bad sizes");
    hipStream_t s = rf->ctx->stream;
    const int n = rf->n;
    rf->w1.ensure((size_t)n * sizeof(double));
    double *z = rf->w1.as<double>();
    for (int b = 0; b < rf->batch; ++b) {
      const double *f = rf->facb.as<double>() + (size_t)b * rf->nnzF;
      triop_refresh(rf->lo, f, s);
      triop_refresh(rf->up, f, s);
      launch_perm_cm_to_rm<double>(s, n, 1, d_B + (size_t)b * ldb, ldb, rf->dP.as<int>(), z, 1);
      tri_solve<double>(rf, 1, z);
      launch_perm_rm_to_cm<double>(s, n, 1, z, 1, rf->dQ.as<int>(), d_X + (size_t)b * ldx, ldx);
    }
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    check_ops(rf);
  });
}

}  // extern "C"
