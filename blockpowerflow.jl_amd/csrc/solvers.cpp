// C ABI of the block-Krylov solvers: block_gmres (plain FP64 drop-in and the
// mixed-precision form with FP64 iterative refinement), block_bicgstab, and
// the per-RHS backward error.
#include <chrono>
#include <cmath>
#include <cstring>

#include "krylov.h"

using namespace bpf;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
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

// Row-major FP64 copies of the caller's column-major blocks plus scratch.
struct Blocks {
  DevBuf B, X, R, Xt, scale, stat;
  void init(int n, int p, BufPool *pool) {
    for (DevBuf *b : {&B, &X, &R, &Xt, &scale, &stat}) b->pool = pool;
    const size_t m = (size_t)n * p * sizeof(double);
    B.ensure(m); X.ensure(m); R.ensure(m); Xt.ensure(m);
    scale.ensure(std::max(1, p) * sizeof(double));
  }
};

void col_norms2(bpf_ctx ctx, int n, int p, const double *R, DevBuf &stat, std::vector<double> &out) {
  const size_t pp = std::max(1, p);
  stat.ensure((3 + (size_t)colstat_blocks(n)) * pp * sizeof(double));
  colstat_rowmajor<double>(ctx->stream, n, p, R, p, 1, stat.as<double>(), stat.as<double>() + 3 * pp);
  out.resize(p);
  BPF_CHECK_HIP(hipMemcpyAsync(out.data(), stat.ptr, p * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto &v : out) v = std::sqrt(v);
}

void col_inf(bpf_ctx ctx, int n, int p, const double *X, DevBuf &stat, std::vector<double> &out) {
  stat.ensure(3 * (size_t)std::max(1, p) * sizeof(double));
  colstat_rowmajor<double>(ctx->stream, n, p, X, p, 0, stat.as<double>());
  out.resize(p);
  BPF_CHECK_HIP(hipMemcpyAsync(out.data(), stat.ptr, p * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
}

// Scale columns of R to unit 2-norm (zero columns keep scale 1); returns the
// scales (host) and leaves them on the device in blk.scale.
void scale_rhs(bpf_ctx ctx, int n, int p, double *R, Blocks &blk, const std::vector<double> &nrm,
               std::vector<double> &sc) {
  sc.resize(p);
  for (int c = 0; c < p; ++c) sc[c] = nrm[c] > 0.0 ? nrm[c] : 1.0;
  BPF_CHECK_HIP(hipMemcpyAsync(blk.scale.ptr, sc.data(), p * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
  launch_scale_cols(ctx->stream, n, p, R, p, blk.scale.as<double>(), true);
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
}

void check_common(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p, const void *B, int ldb, const void *X,
                  int ldx, const bpf_solver_params *prm) {
  BPF_REQUIRE(ctx && A && prm, "null argument");
  BPF_REQUIRE(p > 0, "p must be positive");
  BPF_REQUIRE(B && X, "null block");
  BPF_REQUIRE(ldb >= A->n && ldx >= A->n, "leading dimension < n");
  if (N) BPF_REQUIRE(N->n == A->n, "preconditioner size %d differs from n = %d", N->n, A->n);
}

const char *status_str(int st) {
  switch (st) {
    case BPF_STATUS_SOLVED: return "solution good enough given atol and rtol";
    case BPF_STATUS_TIRED: return "maximum number of iterations exceeded";
    case BPF_STATUS_OVERTIMED: return "time limit exceeded";
    default: return "unknown";
  }
}

// Block GMRES (src/block_gmres.jl:170-411) with an FP64 basis on the device.
// mixed = false: everything FP64 (the drop-in).  mixed = true: flexible form
// with the preconditioner (ILU(0) solves, optional GMRES polynomial) in FP32.
// M (may be null) is the left preconditioner: the Krylov run sees M (B - A X)
// and M A N^{-1} (:226,:270,:300), and without a backward-error target it
// stops on ||M (B - A X)||_F as the reference does.
void gmres_plain(bpf_ctx ctx, bpf_csr A, bpf_prec M, bpf_prec N, int p, const double *dB, int ldb,
                 const double *dX0, int ldx0, double *dX, int ldx, const bpf_solver_params &prm,
                 bpf_solver_stats &st, double *h_hist, int hist_cap, bool mixed = false) {
  const double t0 = now_s();
  const int n = A->n;
  hipStream_t s = ctx->stream;
  // Large systems run the mixed form in A's locality relabelling (the
  // products and the preconditioner's triangular solves on relabelled
  // copies; B, X0 and X permuted at the layout conversion).  The FP64 form
  // keeps A's row order, so its Gram sums run in the reference's order: the
  // single-pass block MGS amplifies a change of summation order (70,000 buses:
  // the first cycle stops with the true backward error just above its target
  // in the relabelled order, and a second cycle doubles the iterations).
  const int *order = nullptr;
  if (mixed && g_relabel_min > 0 && n >= g_relabel_min && !M &&
      (!N || N->kind == BPF_PREC_ILU0 || N->kind == BPF_PREC_IC0)) {
    CsrRelabel *rl = relabel_csr(A, s);
    if (N) N = relabel_prec(N, rl, s);
    A = rl->Ar;
    order = rl->order.as<int>();
  }
  auto to_rm = [&](const double *src, int ld, double *dst) {
    if (order) launch_perm_cm_to_rm<double>(s, n, p, src, ld, order, dst, p);
    else launch_cm_to_rm(s, n, p, src, ld, dst, p, false);
  };
  Blocks blk;
  blk.init(n, p, &ctx->pool);
  to_rm(dB, ldb, blk.B.as<double>());
  if (dX0) {
    to_rm(dX0, ldx0, blk.X.as<double>());
    residual_rm(s, A, p, blk.B.as<double>(), blk.X.as<double>(), blk.R.as<double>());
  } else {
    fill_double(s, blk.X.as<double>(), 0.0, (size_t)n * p);
    BPF_CHECK_HIP(hipMemcpyAsync(blk.R.ptr, blk.B.ptr, (size_t)n * p * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  }
  HistoryWriter hist{h_hist, hist_cap};
  std::vector<double> rn, sc, bnorm, xnorm;
  // R0 = M (B - A X) with a left preconditioner; blk.R stays the true residual.
  DevBuf MR(&ctx->pool);
  double *R0 = blk.R.as<double>();
  auto left = [&]() {
    if (!M) return;
    prec_apply_rm<double>(s, M, p, blk.R.as<double>(), p, R0, p);
  };
  if (M) {
    MR.ensure((size_t)n * p * sizeof(double));
    R0 = MR.as<double>();
  }
  left();
  col_norms2(ctx, n, p, R0, blk.stat, rn);
  double RNorm = 0.0;
  for (double v : rn) RNorm += v * v;
  RNorm = std::sqrt(RNorm);
  if (prm.history) hist.push(RNorm);
  const double eps = prm.atol + prm.rtol * RNorm;                       // :230
  const int itmax = prm.itmax > 0 ? prm.itmax : 2 * (n / p);          // :238
  const int mem = std::max(1, std::min(n / p, prm.memory));           // :42
  const bool use_berr = prm.berr_tol > 0.0;
  col_inf(ctx, n, p, blk.B.as<double>(), blk.stat, bnorm);
  std::vector<double> berr(p, INFINITY);
  auto berr_ok = [&]() {
    backward_error_rm(ctx, A, p, blk.R.as<double>(), blk.X.as<double>(), blk.B.as<double>(), berr.data(),
                      blk.stat);
    double m = 0.0;
    for (double v : berr) m = std::max(m, v);
    st.max_berr = m;
    return m <= prm.berr_tol;
  };
  bool solved = use_berr ? berr_ok() : RNorm <= eps;
  int iter = 0;
  bool tired = iter >= itmax, overtimed = false;
  if (prm.verbose > 0) printf("BLOCK-GMRES: system of size %d with %d right-hand sides\n", n, p);
  ArnoldiConfig cfg;
  cfg.n = n; cfg.p = p; cfg.s = prm.block_width; cfg.f32 = false; cfg.flexible = mixed;
  cfg.prec32 = mixed;
  cfg.M = M;
  if (mixed && N) {
    cfg.inner_its = std::max(0, prm.inner_its);
    cfg.inner_s = prm.inner_width;
  }
  // Block MGS as in the reference (src/block_gmres.jl:301-304); with
  // reorthogonalization=true two classical passes.  (A single classical pass
  // loses orthogonality on these Jacobians and stalls the solve.)
  cfg.orth = prm.reorthogonalization ? ORTH_BCGS2 : ORTH_BMGS;
  cfg.bmgs_blocks = std::max(1, prm.bmgs_blocks);
  BlockArnoldi arn(ctx, A, N, cfg);
  arn.set_scale(blk.scale.as<double>());
  Poly poly;
  if (mixed && N && prm.poly_degree > 0 && cfg.inner_its == 0) {
    poly.build(ctx, A, N, prm.poly_degree);
    arn.set_poly(&poly);
  }
  int cycles = 0;
  while (!(solved || tired || overtimed)) {
    ++cycles;
    scale_rhs(ctx, n, p, R0, blk, rn, sc);
    const int maxit = prm.restart ? std::min(mem, itmax - iter) : itmax - iter;
    // Per-column targets: peek at the solution once the residual has dropped,
    // to turn the backward-error tolerance into a residual target.
    bool have_x = false, est_ok = false;
    std::vector<double> target(p, 0.0);
    auto stop = [&](const double *r, int it) {
      double F = 0.0;
      double worst = 0.0;
      for (int c = 0; c < p; ++c) {
        const double rc = r[c] * sc[c];
        F += rc * rc;
        worst = std::max(worst, r[c]);
      }
      F = std::sqrt(F);
      if (prm.history) hist.push(F);
      if (prm.verbose > 0 && (iter + it) % prm.verbose == 0)
        printf("%5d  %7.1e  %.2fs\n", iter + it, F, now_s() - t0);
      bool ok;
      if (use_berr) {
        if (!have_x && worst < 1e-3) {
          arn.peek(blk.X.as<double>(), blk.Xt.as<double>());
          col_inf(ctx, n, p, blk.Xt.as<double>(), blk.stat, xnorm);
          for (int c = 0; c < p; ++c) target[c] = prm.berr_tol * (A->norm_inf * xnorm[c] + bnorm[c]);
          have_x = true;
        }
        ok = have_x;
        for (int c = 0; c < p && ok; ++c) ok = r[c] * sc[c] <= target[c];
      } else {
        ok = F <= eps;
      }
      est_ok = ok;
      if (prm.timemax > 0 && now_s() - t0 > prm.timemax) overtimed = true;
      return ok || overtimed;
    };
    const int its = arn.run(R0, blk.X.as<double>(), maxit, stop, nullptr);
    iter += its;
    residual_rm(s, A, p, blk.B.as<double>(), blk.X.as<double>(), blk.R.as<double>());
    left();
    col_norms2(ctx, n, p, R0, blk.stat, rn);
    // With a backward-error target the true residual decides; otherwise the
    // reference's rule: the residual estimate ||Lambda||_F <= eps (:349).
    // After a breakdown in CholQR2 (a rank-deficient block, as with linearly
    // dependent right-hand sides: columns completed with new directions, or a
    // shifted factor) the estimate is not to be trusted: then the true
    // residual decides, and a further cycle restarts from it.
    if (use_berr) {
      solved = berr_ok();
    } else if (arn.shifted()) {
      double F = 0.0;
      for (double v : rn) F += v * v;
      solved = std::sqrt(F) <= eps;
    } else {
      solved = est_ok;
    }
    tired = iter >= itmax;
    if (prm.timemax > 0 && now_s() - t0 > prm.timemax) overtimed = true;
    if (its == 0) break;
  }
  if (order) launch_perm_rm_to_cm<double>(s, n, p, blk.X.as<double>(), p, order, dX, ldx);
  else launch_rm_to_cm(s, n, p, blk.X.ptr, p, false, dX, ldx);
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  st.niter = iter;
  st.solved = solved;
  st.status = overtimed ? BPF_STATUS_OVERTIMED : solved ? BPF_STATUS_SOLVED
                                                        : tired ? BPF_STATUS_TIRED : BPF_STATUS_UNKNOWN;
  st.refine_steps = cycles - 1;
  st.nresiduals = hist.n;
  st.timer = now_s() - t0;
  if (!use_berr) {
    backward_error_rm(ctx, A, p, blk.R.as<double>(), blk.X.as<double>(), blk.B.as<double>(), berr.data(),
                      blk.stat);
    double m = 0.0;
    for (double v : berr) m = std::max(m, v);
    st.max_berr = m;
  }
  if (prm.verbose > 0) printf("%s\n", status_str(st.status));
}

// MIXED32: FP32 flexible block GMRES (FP32 SpMM, FP32 ILU solves, FP32
// basis; FP64 Gram/Cholesky/Hessenberg) inside FP64 iterative refinement with
// the true residual B - A X.
void gmres_mixed(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p, const double *dB, int ldb,
                 const double *dX0, int ldx0, double *dX, int ldx, const bpf_solver_params &prm,
                 bpf_solver_stats &st, double *h_hist, int hist_cap) {
  const double t0 = now_s();
  const int n = A->n;
  hipStream_t s = ctx->stream;
  Blocks blk;
  blk.init(n, p, &ctx->pool);
  launch_cm_to_rm(s, n, p, dB, ldb, blk.B.ptr, p, false);
  if (dX0) launch_cm_to_rm(s, n, p, dX0, ldx0, blk.X.ptr, p, false);
  else fill_double(s, blk.X.as<double>(), 0.0, (size_t)n * p);
  const double tol = prm.berr_tol > 0.0 ? prm.berr_tol : 1e-10;
  const double inner = prm.inner_rtol > 0.0 ? prm.inner_rtol : 1e-6;
  const int itmax = prm.itmax > 0 ? prm.itmax : 2 * (n / p);
  const int maxref = prm.max_refine > 0 ? prm.max_refine : 10;
  HistoryWriter hist{h_hist, hist_cap};
  std::vector<double> rn, sc, bnorm, xnorm, berr(p);
  col_inf(ctx, n, p, blk.B.as<double>(), blk.stat, bnorm);
  ArnoldiConfig cfg;
  cfg.n = n; cfg.p = p; cfg.s = prm.block_width; cfg.flexible = true;
  // FP32 basis: two classical passes keep it orthonormal to working
  // precision (bmgs_blocks > 1: block MGS over steps of that many blocks).
  cfg.f32 = true;
  cfg.orth = prm.bmgs_blocks > 1 ? ORTH_BMGS : ORTH_BCGS2;
  cfg.bmgs_blocks = std::max(1, prm.bmgs_blocks);
  BlockArnoldi arn(ctx, A, N, cfg);
  arn.set_scale(blk.scale.as<double>());
  Poly poly;
  if (N && prm.poly_degree > 0) {
    poly.build(ctx, A, N, prm.poly_degree);
    arn.set_poly(&poly);
    if (prm.verbose > 0) {
      printf("polynomial preconditioner: degree %d, roots", poly.degree());
      for (auto &r : poly.roots) printf(" (%.3g,%.3g)", r.real(), r.imag());
      printf("\n");
    }
  }
  int iter = 0, step = 0;
  bool solved = false, overtimed = false;
  for (;;) {
    residual_rm(s, A, p, blk.B.as<double>(), blk.X.as<double>(), blk.R.as<double>());
    backward_error_rm(ctx, A, p, blk.R.as<double>(), blk.X.as<double>(), blk.B.as<double>(), berr.data(),
                      blk.stat);
    double m = 0.0;
    for (double v : berr) m = std::max(m, v);
    st.max_berr = m;
    col_norms2(ctx, n, p, blk.R.as<double>(), blk.stat, rn);
    if (prm.history) {
      double F = 0.0;
      for (double v : rn) F += v * v;
      hist.push(std::sqrt(F));
    }
    if (prm.verbose > 0) printf("refinement %d: max backward error %.3e\n", step, m);
    solved = m <= tol;
    if (solved || step >= maxref || iter >= itmax) break;
    if (prm.timemax > 0 && now_s() - t0 > prm.timemax) { overtimed = true; break; }
    col_inf(ctx, n, p, blk.X.as<double>(), blk.stat, xnorm);
    std::vector<double> target(p);
    for (int c = 0; c < p; ++c)
      target[c] = std::max(inner * rn[c], 0.5 * tol * (A->norm_inf * xnorm[c] + bnorm[c]));
    scale_rhs(ctx, n, p, blk.R.as<double>(), blk, rn, sc);
    auto stop = [&](const double *r, int it) {
      bool ok = true;
      double worst = 0.0;
      for (int c = 0; c < p; ++c) {
        ok = ok && r[c] * sc[c] <= target[c];
        worst = std::max(worst, r[c] * sc[c] / std::max(target[c], 1e-300));
      }
      if (prm.verbose > 0 && it % prm.verbose == 0)
        printf("  step %d inner %4d  max(estimate/target) %.3e\n", step, it, worst);
      if (prm.timemax > 0 && now_s() - t0 > prm.timemax) overtimed = true;
      (void)it;
      return ok || overtimed;
    };
    iter += arn.run(blk.R.as<double>(), blk.X.as<double>(), itmax - iter, stop, nullptr);
    ++step;
  }
  launch_rm_to_cm(s, n, p, blk.X.ptr, p, false, dX, ldx);
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  st.niter = iter;
  st.solved = solved;
  st.refine_steps = step;
  st.status = overtimed ? BPF_STATUS_OVERTIMED : solved ? BPF_STATUS_SOLVED : BPF_STATUS_TIRED;
  st.nresiduals = hist.n;
  st.timer = now_s() - t0;
  if (prm.verbose > 0) printf("%s\n", status_str(st.status));
}

}  // namespace

extern "C" {

void bpf_solver_params_default(bpf_solver_params *prm) {
  if (!prm) return;
  std::memset(prm, 0, sizeof(*prm));
  prm->memory = 20;
  prm->restart = 0;
  prm->itmax = 0;
  prm->atol = 1.4901161193847656e-08;   // sqrt(eps(Float64))
  prm->rtol = 1.4901161193847656e-08;
  prm->timemax = 0.0;
  prm->berr_tol = 0.0;
  prm->precision = BPF_PRECISION_FP64;
  prm->block_width = 0;
  prm->max_refine = 10;
  prm->inner_rtol = 1e-6;
  prm->reorthogonalization = 0;
  prm->poly_degree = 0;
  prm->bmgs_blocks = 1;
  prm->inner_its = 0;
  prm->inner_width = 0;
}

int bpf_block_gmres_x0(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p, const double *d_B, int ldb,
                       const double *d_X0, int ldx0, double *d_X, int ldx,
                       const bpf_solver_params *prm, bpf_solver_stats *stats, double *h_history,
                       int hist_cap) {
  return bpf_block_gmres_ex(ctx, A, nullptr, N, BPF_F64, p, d_B, ldb, d_X0, ldx0, d_X, ldx, prm, stats,
                            h_history, hist_cap);
}

int bpf_block_gmres(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p, const double *d_B, int ldb,
                    double *d_X, int ldx, const bpf_solver_params *prm, bpf_solver_stats *stats,
                    double *h_history, int hist_cap) {
  return bpf_block_gmres_x0(ctx, A, N, p, d_B, ldb, nullptr, 0, d_X, ldx, prm, stats, h_history,
                            hist_cap);
}

// Left and right preconditioning and the element type of the blocks (the
// reference's solvers are generic over it, src/block_gmres.jl:170-173):
// Float32 B and X0 are widened to FP64 column-major copies, the solve runs
// in FP64 and X is rounded back to Float32.
int bpf_block_gmres_ex(bpf_ctx ctx, bpf_csr A, bpf_prec M, bpf_prec N, int dtype, int p, const void *d_B,
                       int ldb, const void *d_X0, int ldx0, void *d_X, int ldx, const bpf_solver_params *prm,
                       bpf_solver_stats *stats, double *h_history, int hist_cap) {
  return guard([&] {
    BPF_REQUIRE(dtype == BPF_F64 || dtype == BPF_F32, "bad dtype %d", dtype);
    check_common(ctx, A, N, p, d_B, ldb, d_X, ldx, prm);
    if (M) BPF_REQUIRE(M->n == A->n, "left preconditioner size %d differs from n = %d", M->n, A->n);
    if (d_X0) BPF_REQUIRE(ldx0 >= A->n, "leading dimension of X0 < n");
    const int n = A->n;
    hipStream_t s = ctx->stream;
    const double *B = static_cast<const double *>(d_B), *X0 = static_cast<const double *>(d_X0);
    double *X = static_cast<double *>(d_X);
    int lb = ldb, lx0 = ldx0, lx = ldx;
    DevBuf Bw(&ctx->pool), X0w(&ctx->pool), Xw(&ctx->pool);
    if (dtype == BPF_F32) {
      const size_t mb = (size_t)std::max(1, n) * p * sizeof(double);
      Bw.ensure(mb);
      Xw.ensure(mb);
      copy_cast<float, double>(s, p, n, static_cast<const float *>(d_B), ldb, Bw.as<double>(), n, 1.0);
      B = Bw.as<double>();
      X = Xw.as<double>();
      lb = lx = n;
      if (d_X0) {
        X0w.ensure(mb);
        copy_cast<float, double>(s, p, n, static_cast<const float *>(d_X0), ldx0, X0w.as<double>(), n, 1.0);
        X0 = X0w.as<double>();
        lx0 = n;
      }
    }
    bpf_solver_stats st{};
    if (n == 0) {
      st.solved = 1;
      st.status = BPF_STATUS_SOLVED;
    } else if (prm->precision == BPF_PRECISION_MIXED) {
      gmres_plain(ctx, A, M, N, p, B, lb, X0, lx0, X, lx, *prm, st, h_history, hist_cap, true);
    } else if (prm->precision == BPF_PRECISION_MIXED32) {
      if (M) BPF_THROW(BPF_E_UNSUPPORTED, "left preconditioning is not available with precision MIXED32");
      gmres_mixed(ctx, A, N, p, B, lb, X0, lx0, X, lx, *prm, st, h_history, hist_cap);
    } else if (p >= g_wide_min && (prm->block_width <= 0 || prm->block_width >= p)) {
      // one n x p Krylov block, as the reference runs every width
      gmres_wide(ctx, A, M, N, p, B, lb, X0, lx0, X, lx, *prm, st, h_history, hist_cap);
    } else {
      gmres_plain(ctx, A, M, N, p, B, lb, X0, lx0, X, lx, *prm, st, h_history, hist_cap);
    }
    if (dtype == BPF_F32 && n > 0) {
      copy_cast<double, float>(s, p, n, X, n, static_cast<float *>(d_X), ldx, 1.0);
      BPF_CHECK_HIP(hipStreamSynchronize(s));
    }
    if (stats) *stats = st;
  });
}

int bpf_backward_error(bpf_ctx ctx, bpf_csr A, int p, const double *d_X, int ldx, const double *d_B,
                       int ldb, double *h_berr) {
  return guard([&] {
    BPF_REQUIRE(ctx && A && h_berr, "null argument");
    if (p <= 0) return;
    BPF_REQUIRE(ldx >= A->n && ldb >= A->n, "leading dimension < n");
    const int n = A->n;
    Blocks blk;
    blk.init(n, p, &ctx->pool);
    launch_cm_to_rm(ctx->stream, n, p, d_B, ldb, blk.B.ptr, p, false);
    launch_cm_to_rm(ctx->stream, n, p, d_X, ldx, blk.X.ptr, p, false);
    residual_rm(ctx->stream, A, p, blk.B.as<double>(), blk.X.as<double>(), blk.R.as<double>());
    backward_error_rm(ctx, A, p, blk.R.as<double>(), blk.X.as<double>(), blk.B.as<double>(), h_berr,
                      blk.stat);
  });
}

}  // extern "C"
