// Block GMRES with one Krylov block of any width p, as the reference runs it
// (src/block_gmres.jl:46-56,282-385): the n x p blocks V_i live
// column-major on the device; the sparse operator and the preconditioners act
// on row-major copies through the library's own kernels; the dense algebra
// over p x p and 2p x p blocks (block modified Gram-Schmidt products,
// Householder QR of the basis block and of the Hessenberg blocks, the
// reflector applications, the block back-substitution) runs on rocBLAS and
// rocSOLVER, with LAPACK's Householder semantics as in householder!
// (src/block_utils.jl:46-53: geqrf, the upper triangle copied out, orgqr).
// Block BiCGSTAB for p > 128 uses the same libraries for its p x p
// products and LU solves.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <cmath>
#include <cstring>

#include "krylov.h"

namespace bpf {

std::atomic<int> g_wide_min{129};

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

#define BPF_BLAS(expr)                                                                              \
  do {                                                                                              \
    rocblas_status st_ = (expr);                                                                    \
    if (st_ != rocblas_status_success)                                                              \
      BPF_THROW(st_ == rocblas_status_memory_error ? BPF_E_NOMEM : BPF_E_HIP, "%s failed: %s (%s:%d)", \
                #expr, rocblas_status_to_string(st_), __FILE__, __LINE__);                         \
  } while (0)

// dst(i, j) for i < rows, j < cols (column-major, leading dimensions lds,
// ldd): mode 0 copies, 1 copies the upper triangle and zeroes the rest,
// 2 zeroes, 3 adds src.
__global__ __launch_bounds__(256) void sub_kernel(int rows, int cols, const double *src, int lds, double *dst,
                                                  int ldd, int mode) {
  const size_t total = (size_t)rows * cols;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % rows), j = (int)(t / rows);
    double *d = dst + (size_t)j * ldd + i;
    if (mode == 0) *d = src[(size_t)j * lds + i];
    else if (mode == 1) *d = i <= j ? src[(size_t)j * lds + i] : 0.0;
    else if (mode == 2) *d = 0.0;
    else *d += src[(size_t)j * lds + i];
  }
}

enum { SUB_COPY = 0, SUB_UPPER = 1, SUB_ZERO = 2, SUB_ADD = 3 };

// Sets *flag when the block is rank-deficient: a diagonal entry of its R at
// rounding level against the norm of its column (r_jj^2 <= 1e-10 ||q_j||^2)
// or a zero column.
__global__ __launch_bounds__(256) void rank_flag_kernel(int p, const double *R, int ldr, const double *cn,
                                                        int *flag) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= p) return;
  const double r = R[(size_t)j * ldr + j];
  if (!(r * r > 1e-10 * cn[j] * cn[j])) atomicOr(flag, 1);
}

void sub(hipStream_t s, int rows, int cols, const double *src, int lds, double *dst, int ldd, int mode) {
  if (rows <= 0 || cols <= 0) return;
  const size_t total = (size_t)rows * cols;
  const int grid = (int)std::min<size_t>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(sub_kernel, dim3(grid), dim3(256), 0, s, rows, cols, src, lds, dst, ldd, mode);
  BPF_CHECK_HIP(hipGetLastError());
}

// C = op(A) op(B) * alpha + beta C, column-major.
void gemm(rocblas_handle h, bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int lda,
          const double *B, int ldb, double beta, double *C, int ldc) {
  BPF_BLAS(rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                         tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb,
                         &beta, C, ldc));
}

double nrm2(rocblas_handle h, size_t count, const double *x) {
  double r = 0.0;
  BPF_REQUIRE(count <= (size_t)INT32_MAX, "block of %zu entries too large for one norm", count);
  BPF_BLAS(rocblas_dnrm2(h, (int)count, x, 1, &r));   // host pointer mode: returns when r is set
  return r;
}

// householder!(Q, R, tau) of src/block_utils.jl:46-53 on an m x p
// column-major block: geqrf, R = its upper triangle, then orgqr unless
// compact.
void householder(rocblas_handle h, hipStream_t s, int m, int p, double *Q, int ldq, double *R, int ldr, double *tau,
                 bool compact) {
  BPF_BLAS(rocsolver_dgeqrf(h, m, p, Q, ldq, tau));
  sub(s, p, p, Q, ldq, R, ldr, SUB_UPPER);
  if (!compact) BPF_BLAS(rocsolver_dorgqr(h, m, p, p, Q, ldq, tau));
}

// r_jj of R1 against the column norms sqrt(G_jj) of the Gram matrix, which
// potrf has overwritten above the diagonal only: d[j] = sqrt(G_jj) first.
__global__ __launch_bounds__(256) void gram_diag_kernel(int p, const double *G, int ldg, double *d) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < p) d[j] = sqrt(G[(size_t)j * ldg + j]);
}

// Orthonormal basis of an n x p column-major block Q (in place) and its
// p x p upper R (ld ldr), with Q_in = Q_out R.  CholQR2 (Gram matrix on
// rocBLAS, Cholesky on rocSOLVER, triangular solve; twice, R = R2 R1) when
// the block has full numerical rank: r_jj^2 > 1e-10 G_jj for every column
// of the first Cholesky factor, as on the CholQR2 path of narrower blocks.
// Otherwise householder! (src/block_utils.jl:46-53: geqrf, orgqr), the
// reference's route, which completes a rank-deficient block with
// orthonormal directions, from a copy of Q; then returns true.  Either way
// span(Q_out) = span(Q_in) where Q_in has full rank, so the Krylov spaces and
// the residual norms are the reference's; the factors may differ from
// Householder's by column signs.
struct BasisQr {
  DevBuf save, G, G2, T, cn, tau, dinfo;
  void init(BufPool *pool, int n, int p) {
    for (DevBuf *b : {&save, &G, &G2, &T, &cn, &tau, &dinfo}) b->pool = pool;
    save.ensure((size_t)n * p * sizeof(double));
    for (DevBuf *b : {&G, &G2, &T}) b->ensure((size_t)p * p * sizeof(double));
    cn.ensure(p * sizeof(double));
    tau.ensure(p * sizeof(double));
    dinfo.ensure(4 * sizeof(int));
  }
};

bool basis_qr(rocblas_handle h, hipStream_t s, int n, int p, double *Q, double *R, int ldr, BasisQr &w) {
  const double one = 1.0;
  int *di = w.dinfo.as<int>();
  int hi[2] = {0, 0};
  BPF_CHECK_HIP(hipMemcpyAsync(w.save.ptr, Q, (size_t)n * p * sizeof(double), hipMemcpyDeviceToDevice, s));
  BPF_CHECK_HIP(hipMemsetAsync(di, 0, 4 * sizeof(int), s));
  gemm(h, true, false, p, p, n, 1.0, Q, n, Q, n, 0.0, w.G.as<double>(), p);
  hipLaunchKernelGGL(gram_diag_kernel, dim3((p + 255) / 256), dim3(256), 0, s, p, w.G.as<double>(), p,
                     w.cn.as<double>());
  BPF_CHECK_HIP(hipGetLastError());
  BPF_BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, p, w.G.as<double>(), p, di));
  hipLaunchKernelGGL(rank_flag_kernel, dim3((p + 255) / 256), dim3(256), 0, s, p, w.G.as<double>(), p,
                     w.cn.as<double>(), di + 1);
  BPF_CHECK_HIP(hipGetLastError());
  BPF_CHECK_HIP(hipMemcpyAsync(hi, di, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  if (hi[0] == 0 && hi[1] == 0) {
    BPF_BLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, n, p, &one, w.G.as<double>(), p, Q, n));
    gemm(h, true, false, p, p, n, 1.0, Q, n, Q, n, 0.0, w.G2.as<double>(), p);
    BPF_BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, p, w.G2.as<double>(), p, di + 2));
    BPF_CHECK_HIP(hipMemcpyAsync(hi, di + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    if (hi[0] == 0) {
      BPF_BLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                             rocblas_diagonal_non_unit, n, p, &one, w.G2.as<double>(), p, Q, n));
      sub(s, p, p, w.G.as<double>(), p, w.T.as<double>(), p, SUB_UPPER);
      BPF_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                             rocblas_diagonal_non_unit, p, p, &one, w.G2.as<double>(), p, w.T.as<double>(), p, R,
                             ldr));
      return false;
    }
    BPF_CHECK_HIP(hipMemcpyAsync(Q, w.save.ptr, (size_t)n * p * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  householder(h, s, n, p, Q, n, R, ldr, w.tau.as<double>(), false);
  return true;
}

struct HistoryWriter {
  double *h;
  int cap;
  int n = 0;
  void push(double v) {
    if (h && n < cap) h[n] = v;
    ++n;
  }
};

}  // namespace

rocblas_handle blas_handle(bpf_ctx ctx) {
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    BPF_BLAS(rocblas_create_handle(&h));
    ctx->blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(ctx->blas);
  BPF_BLAS(rocblas_set_stream(h, ctx->stream));
  return h;
}

void blas_release(bpf_ctx ctx) {
  if (ctx->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
  ctx->blas = nullptr;
}

// Block GMRES over one n x p Krylov block (src/block_gmres.jl:170-411), FP64.
// Stops on ||M (B - A X)||_F <= atol + rtol ||M (B - A X0)||_F as the
// reference does.  Two additions decide at the end of a pass instead of the
// estimate: with berr_tol > 0 the per-RHS backward error, and after a
// rank-deficient basis block in the pass (linearly dependent right-hand
// sides: Householder QR completes them with directions that rounding leaves
// short of orthogonal to the basis, so the estimate is no longer the
// residual) the true residual.  An unmet target starts another pass from the
// true residual, restart or not.
void gmres_wide(bpf_ctx ctx, bpf_csr A, bpf_prec M, bpf_prec N, int p, const double *dB, int ldb,
                const double *dX0, int ldx0, double *dX, int ldx, const bpf_solver_params &prm,
                bpf_solver_stats &st, double *h_hist, int hist_cap) {
  const double t0 = now_s();
  const int n = A->n;
  BPF_REQUIRE(p <= n, "block_gmres: %d right-hand sides exceed the %d unknowns (one n x p Krylov block)", p, n);
  hipStream_t s = ctx->stream;
  rocblas_handle h = blas_handle(ctx);
  BufPool *pool = &ctx->pool;
  const size_t nb = (size_t)n * p, bb = nb * sizeof(double), pb = (size_t)p * p * sizeof(double);
  DevBuf Brm(pool), Xrm(pool), Rrm(pool), T1(pool), T2(pool), Q(pool), W0(pool), stat(pool);
  DevBuf tauH(pool), C(pool), D(pool), Cb(pool), Psi(pool);
  BasisQr bq;
  bq.init(pool, n, p);
  for (DevBuf *b : {&Brm, &Xrm, &Rrm, &T1, &T2, &Q, &W0}) b->ensure(bb);
  for (DevBuf *b : {&C, &Cb, &Psi}) b->ensure(pb);
  D.ensure(2 * pb);
  tauH.ensure(p * sizeof(double));
  launch_cm_to_rm(s, n, p, dB, ldb, Brm.ptr, p, false);
  if (dX0) launch_cm_to_rm(s, n, p, dX0, ldx0, Xrm.ptr, p, false);
  else fill_double(s, Xrm.as<double>(), 0.0, nb);
  // W0 = M (B - A X), column-major (:218-226, :268-270)
  auto initial_residual = [&]() {
    residual_rm(s, A, p, Brm.as<double>(), Xrm.as<double>(), Rrm.as<double>());
    const double *src = Rrm.as<double>();
    if (M) {
      prec_apply_rm<double>(s, M, p, Rrm.as<double>(), p, T1.as<double>(), p);
      src = T1.as<double>();
    }
    launch_rm_to_cm(s, n, p, src, p, false, W0.as<double>(), n);
  };
  std::vector<double> berr(p, INFINITY);
  auto berr_max = [&]() {
    residual_rm(s, A, p, Brm.as<double>(), Xrm.as<double>(), Rrm.as<double>());
    backward_error_rm(ctx, A, p, Rrm.as<double>(), Xrm.as<double>(), Brm.as<double>(), berr.data(), stat);
    double m = 0.0;
    for (double v : berr) m = std::max(m, v);
    return m;
  };
  HistoryWriter hist{h_hist, hist_cap};
  initial_residual();
  double RNorm = nrm2(h, nb, W0.as<double>());                             // :227
  if (prm.history) hist.push(RNorm);
  const double eps = prm.atol + prm.rtol * RNorm;                          // :230
  const int itmax = prm.itmax > 0 ? prm.itmax : 2 * (n / p);              // :238
  const int mem = std::max(1, std::min(n / p, prm.memory));              // :42
  const bool use_berr = prm.berr_tol > 0.0;
  if (prm.verbose > 0) printf("BLOCK-GMRES: system of size %d with %d right-hand sides (one Krylov block)\n", n, p);
  int iter = 0, inner_itmax = itmax, npass = 0;
  bool solved = use_berr ? berr_max() <= prm.berr_tol : RNorm <= eps;
  bool tired = iter >= itmax, overtimed = false;
  // V: basis blocks; Rc: block columns of R; Om: the orthogonal 2p x 2p
  // factors Omega_k of the Hessenberg QR, formed explicitly so that each
  // later application is one matrix product; Z: blocks of Z.
  std::vector<DevBuf> V, Rc, Om, Z;
  bool have_w0 = true;
  auto fresh = [&](std::vector<DevBuf> &v, size_t bytes) {
    v.emplace_back(pool);
    v.back().ensure(bytes);
    return v.back().as<double>();
  };
  while (!(solved || tired || overtimed)) {
    V.clear(); Rc.clear(); Om.clear(); Z.clear();
    if (!have_w0) initial_residual();                                      // :266-271
    have_w0 = false;
    // V_1 and Gamma = Z_1: householder!(V[1], Z[1], tau[1]) (:274-276)
    double *V0 = fresh(V, bb);
    BPF_CHECK_HIP(hipMemcpyAsync(V0, W0.ptr, bb, hipMemcpyDeviceToDevice, s));
    double *Z0 = fresh(Z, pb);
    bool rank_deficient = basis_qr(h, s, n, p, V0, Z0, p, bq);
    ++npass;
    int inner = 0;
    bool inner_tired = false, pass_solved = false;
    while (!(pass_solved || inner_tired || overtimed)) {                   // :282
      const int k = inner++;
      // P = N V_k, W = A P, Q = M W (:297-300) on row-major copies
      launch_cm_to_rm(s, n, p, V[k].as<double>(), n, T1.ptr, p, false);
      double *src = T1.as<double>(), *oth = T2.as<double>();
      if (N) {
        prec_apply_rm<double>(s, N, p, src, p, oth, p);
        std::swap(src, oth);
      }
      launch_spmm<double>(s, n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.ptr, false, p, src, p, oth,
                          p, 1.0, 0.0);
      std::swap(src, oth);
      if (M) {
        prec_apply_rm<double>(s, M, p, src, p, oth, p);
        std::swap(src, oth);
      }
      double *Qp = Q.as<double>();
      launch_rm_to_cm(s, n, p, src, p, false, Qp, n);
      // Block column k of R: Psi_ik for i = 0..k, (k+1)p x p column-major.
      const int ldr = (k + 1) * p;
      double *Rk = fresh(Rc, (size_t)ldr * pb / p);
      for (int i = 0; i <= k; ++i) {                                        // :301-304
        double *Rik = Rk + (size_t)i * p;
        gemm(h, true, false, p, p, n, 1.0, V[i].as<double>(), n, Qp, n, 0.0, Rik, ldr);
        gemm(h, false, false, n, p, p, -1.0, V[i].as<double>(), n, Rik, ldr, 1.0, Qp, n);
      }
      if (prm.reorthogonalization) {                                       // :307-313
        for (int i = 0; i <= k; ++i) {
          double *Rik = Rk + (size_t)i * p;
          gemm(h, true, false, p, p, n, 1.0, V[i].as<double>(), n, Qp, n, 0.0, Psi.as<double>(), p);
          gemm(h, false, false, n, p, p, -1.0, V[i].as<double>(), n, Psi.as<double>(), p, 1.0, Qp, n);
          sub(s, p, p, Psi.as<double>(), p, Rik, ldr, SUB_ADD);
        }
      }
      const bool hh = basis_qr(h, s, n, p, Qp, C.as<double>(), p, bq);      // :316
      rank_deficient |= hh;
      if (hh) {
        // A rank-deficient block (the Krylov space exhausted to rounding, or
        // dependent directions): Householder QR completed it with directions
        // that are orthonormal among themselves but not to the earlier basis,
        // and blocks built on them would break the Arnoldi relation (the
        // least-squares solution then grows without bound).  The block is
        // projected off V_0..V_k once more, Phi_i = V_i^T Q, Q -= V_i Phi_i,
        // and orthonormalized again, Q = Q2 C2; then A P = sum_i V_i
        // (Psi_ik + Phi_i C) + Q2 (C2 C), so Psi_ik += Phi_i C and C <- C2 C.
        double *Cp = C.as<double>(), *Ph = Psi.as<double>(), *T2 = D.as<double>();
        for (int i = 0; i <= k; ++i) {
          gemm(h, true, false, p, p, n, 1.0, V[i].as<double>(), n, Qp, n, 0.0, Ph, p);
          gemm(h, false, false, n, p, p, -1.0, V[i].as<double>(), n, Ph, p, 1.0, Qp, n);
          gemm(h, false, false, p, p, p, 1.0, Ph, p, Cp, p, 1.0, Rk + (size_t)i * p, ldr);
        }
        basis_qr(h, s, n, p, Qp, T2, p, bq);
        gemm(h, false, false, p, p, p, 1.0, T2, p, Cp, p, 0.0, T2 + (size_t)p * p, p);
        sub(s, p, p, T2 + (size_t)p * p, p, Cp, p, SUB_COPY);
      }
      // previous reflections: [Psi_ik; Psi_i+1,k] = Omega_i^T [...] (:320-326)
      double *Dp = D.as<double>();
      for (int i = 0; i + 1 <= k; ++i) {
        gemm(h, true, false, 2 * p, p, 2 * p, 1.0, Om[i].as<double>(), 2 * p, Rk + (size_t)i * p, ldr, 0.0, Dp,
             2 * p);
        sub(s, 2 * p, p, Dp, 2 * p, Rk + (size_t)i * p, ldr, SUB_COPY);
      }
      // Omega_k: Householder QR of [Psi_kk; C] (:329-331), then Omega_k
      // itself (2p x 2p) from its reflectors
      double *Omk = fresh(Om, 4 * pb);
      double *Rkk = Rk + (size_t)k * p;
      sub(s, p, p, Rkk, ldr, Omk, 2 * p, SUB_COPY);
      sub(s, p, p, C.as<double>(), p, Omk + p, 2 * p, SUB_COPY);
      householder(h, s, 2 * p, p, Omk, 2 * p, Rkk, ldr, tauH.as<double>(), true);
      BPF_BLAS(rocsolver_dorgqr(h, 2 * p, 2 * p, p, Omk, 2 * p, tauH.as<double>()));
      // [Z_k; Lambda] = Omega_k^T [Z_k; 0] (:334-337)
      gemm(h, true, false, 2 * p, p, p, 1.0, Omk, 2 * p, Z[k].as<double>(), p, 0.0, Dp, 2 * p);
      sub(s, p, p, Dp, 2 * p, Z[k].as<double>(), p, SUB_COPY);
      sub(s, p, p, Dp + p, 2 * p, Cb.as<double>(), p, SUB_COPY);
      RNorm = nrm2(h, (size_t)p * p, Cb.as<double>());                    // :341-342
      if (prm.history) hist.push(RNorm);
      pass_solved = RNorm <= eps;                                          // :349
      inner_tired = prm.restart ? inner >= std::min(mem, inner_itmax) : inner >= inner_itmax;
      if (prm.timemax > 0 && now_s() - t0 > prm.timemax) overtimed = true;
      if (prm.verbose > 0 && (iter + inner) % prm.verbose == 0)
        printf("%5d  %5d  %7.1e  %.2fs%s\n", npass, iter + inner, RNorm, now_s() - t0,
               hh ? "  (rank-deficient block: Householder QR)" : "");
      if (!(pass_solved || inner_tired || overtimed)) {                    // :356-363
        V.emplace_back(std::move(Q));
        Q = DevBuf(pool);
        Q.ensure(bb);
        Z.emplace_back(std::move(Cb));
        Cb = DevBuf(pool);
        Cb.ensure(pb);
      }
    }
    // Y: R Y = Z by block back-substitution (:367-375), in place in Z.
    for (int i = inner - 1; i >= 0; --i) {
      for (int j = inner - 1; j > i; --j)
        gemm(h, false, false, p, p, p, -1.0, Rc[j].as<double>() + (size_t)i * p, (j + 1) * p, Z[j].as<double>(), p,
             1.0, Z[i].as<double>(), p);
      const double one = 1.0;
      BPF_BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                             rocblas_diagonal_non_unit, p, p, &one, Rc[i].as<double>() + (size_t)i * p, (i + 1) * p,
                             Z[i].as<double>(), p));
    }
    // X += N (sum_i V_i Y_i) (:377-385)
    double *Dx = Q.as<double>();
    for (int i = 0; i < inner; ++i)
      gemm(h, false, false, n, p, p, 1.0, V[i].as<double>(), n, Z[i].as<double>(), p, i == 0 ? 0.0 : 1.0, Dx, n);
    launch_cm_to_rm(s, n, p, Dx, n, T1.ptr, p, false);
    const double *add = T1.as<double>();
    if (N) {
      prec_apply_rm<double>(s, N, p, T1.as<double>(), p, T2.as<double>(), p);
      add = T2.as<double>();
    }
    axpby(s, nb, 1.0, add, 1.0, Xrm.as<double>());
    inner_itmax -= inner;
    iter += inner;
    tired = iter >= itmax;
    if (prm.timemax > 0 && now_s() - t0 > prm.timemax) overtimed = true;
    if (use_berr) {
      st.max_berr = berr_max();
      solved = st.max_berr <= prm.berr_tol;
    } else if (rank_deficient) {
      initial_residual();
      have_w0 = true;
      solved = nrm2(h, nb, W0.as<double>()) <= eps;
    } else {
      solved = pass_solved;
    }
  }
  launch_rm_to_cm(s, n, p, Xrm.ptr, p, false, dX, ldx);
  if (!use_berr || npass == 0) st.max_berr = berr_max();
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  if (N) prec_check(N, s);
  if (M) prec_check(M, s);
  st.niter = iter;
  st.solved = solved;
  st.status = overtimed ? BPF_STATUS_OVERTIMED : solved ? BPF_STATUS_SOLVED
                                                        : tired ? BPF_STATUS_TIRED : BPF_STATUS_UNKNOWN;
  st.refine_steps = std::max(0, npass - 1);
  st.nresiduals = hist.n;
  st.timer = now_s() - t0;
}

// Dense operations of block BiCGSTAB for p > 128 (src/block_bicgstab.jl:
// 253-274): n x p blocks row-major (a row-major n x p block is a
// column-major p x n one), p x p matrices column-major.
void WideDense::init(bpf_ctx ctx, int p_) {
  h = blas_handle(ctx);
  s = ctx->stream;
  p = p_;
  ipiv.pool = &ctx->pool;
  ipiv.ensure(std::max(1, p) * sizeof(rocblas_int));
  info.pool = &ctx->pool;
  info.ensure(sizeof(rocblas_int));
}

void WideDense::proj(int n, const double *Vm, const double *Wm, double *out) {
  // out = Vm^T Wm: with Vt = Vm^T, Wt = Wm^T (p x n column-major), Vt Wt^T
  gemm(static_cast<rocblas_handle>(h), false, true, p, p, n, 1.0, Vm, p, Wm, p, 0.0, out, p);
}

void WideDense::update(int n, const double *Vm, const double *psi, double *Y, double alpha, double beta) {
  // Y = alpha Vm psi + beta Y: Yt = alpha psi^T Vt + beta Yt
  gemm(static_cast<rocblas_handle>(h), true, false, p, n, p, alpha, psi, p, Vm, p, beta, Y, p);
}

void WideDense::lu(double *K) {
  BPF_BLAS(rocsolver_dgetrf(static_cast<rocblas_handle>(h), p, p, K, p, ipiv.as<rocblas_int>(),
                            info.as<rocblas_int>()));
}

void WideDense::solve(const double *K, double *rhs) {
  BPF_BLAS(rocsolver_dgetrs(static_cast<rocblas_handle>(h), rocblas_operation_none, p, p, const_cast<double *>(K), p,
                            ipiv.as<rocblas_int>(), rhs, p));
}

}  // namespace bpf
