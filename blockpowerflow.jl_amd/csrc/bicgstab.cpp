// Block BiCGSTAB on the device (El Guennouni, Jbilou & Sadok 2003), following
// block_bicgstab! (src/block_bicgstab.jl:70-178 CPU, :187-299 GPU): p x p
// systems K alpha = D solved by LU with partial pivoting (the CPU variant's
// lu!), omega = |<Q,U>| / |<Q,Q>| with absolute values as at :146 / :265.
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

}  // namespace

extern "C" int bpf_block_bicgstab(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p, const double *d_B,
                                  int ldb, const double *d_C, int ldc, double *d_X, int ldx,
                                  const bpf_solver_params *prm, bpf_solver_stats *stats,
                                  double *h_hist, int hist_cap) {
  return guard([&] {
    BPF_REQUIRE(ctx && A && prm, "null argument");
    BPF_REQUIRE(p > 0 && d_B && d_X, "empty block");
    BPF_REQUIRE(ldb >= A->n && ldx >= A->n, "leading dimension < n");
    if (d_C) BPF_REQUIRE(ldc >= A->n, "leading dimension of C < n");
    if (N) BPF_REQUIRE(N->n == A->n, "preconditioner size differs from n");
    // p > 128: the p x p algebra on rocBLAS / rocSOLVER, p x p matrices
    // column-major
    const bool wide = p > 128;
    const int S = wide ? p : pad_width(p);
    WideDense wd;
    if (wide) wd.init(ctx, p);
    const double t0 = now_s();
    const int n = A->n;
    hipStream_t s = ctx->stream;
    const size_t m = (size_t)n * p, mb = m * sizeof(double);
    DevBuf B, X, P, R, V, Q, U, Y, Cm, K, D, al, be, piv, dots, part, grp;
    for (DevBuf *b : {&B, &X, &P, &R, &V, &Q, &U, &Y}) b->ensure(mb);
    const size_t SS = (size_t)S * S;
    for (DevBuf *b : {&K, &D, &al, &be}) b->ensure(SS * sizeof(double));
    piv.ensure(S * sizeof(int));
    BPF_REQUIRE(!wide || p <= n, "block_bicgstab: %d right-hand sides exceed the %d unknowns", p, n);
    dots.ensure((4 + kDotBlocks) * sizeof(double));
    std::vector<GroupDesc> hg{{0, p, 0, n, 0, 1 << 30, 1, 0}};
    upload(grp, hg, s);
    const GroupDesc *dg = grp.as<GroupDesc>();
    launch_cm_to_rm(s, n, p, d_B, ldb, B.ptr, p, false);
    if (d_C) {
      Cm.ensure(mb);
      launch_cm_to_rm(s, n, p, d_C, ldc, Cm.ptr, p, false);
    }
    const double *Cp = d_C ? Cm.as<double>() : B.as<double>();
    fill_double(s, X.as<double>(), 0.0, m);                                   // :213
    BPF_CHECK_HIP(hipMemcpyAsync(P.ptr, B.ptr, mb, hipMemcpyDeviceToDevice, s));   // :214
    BPF_CHECK_HIP(hipMemcpyAsync(R.ptr, B.ptr, mb, hipMemcpyDeviceToDevice, s));   // :215
    double hd[2];
    auto dot = [&](const double *x, const double *y) {
      launch_dot<double>(s, m, x, y, dots.as<double>(), dots.as<double>() + 4);
      BPF_CHECK_HIP(hipMemcpyAsync(hd, dots.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
      BPF_CHECK_HIP(hipStreamSynchronize(s));
      return hd[0];
    };
    auto proj = [&](const double *Vm, const double *Wm, double *out) {   // out = Vm^T Wm
      if (wide) return wd.proj(n, Vm, Wm, out);
      ProjArgs pa{nullptr, Vm, 0, 1, p, Wm, p, n, dg, 1, 0, nullptr};
      launch_proj<double, double>(s, pa, S, out, 0, SS, false, part);
    };
    auto gemm = [&](const double *Vm, const double *psi, double *Yout, double alpha, double beta) {
      if (wide) return wd.update(n, Vm, psi, Yout, alpha, beta);
      GemmArgs ga{nullptr, Vm, 0, 1, p, psi, 0, SS, S, Yout, p, n, alpha, beta, dg};
      launch_block_gemm<double, double>(s, ga, S, 1);
    };
    auto lu = [&](double *Km) {
      if (wide) return wd.lu(Km);
      launch_lu_small(s, Km, piv.as<int>(), S, dg, 1);
    };
    auto lu_solve = [&](const double *Km, double *rhs) {
      if (wide) return wd.solve(Km, rhs);
      launch_lu_small_solve(s, Km, piv.as<int>(), rhs, S, dg, 1);
    };
    auto precond = [&](const double *in, double *out) {
      prec_apply_rm<double>(s, N, p, in, p, out, p);
    };
    int nh = 0;
    auto push = [&](double v) {
      if (prm->history) {
        if (h_hist && nh < hist_cap) h_hist[nh] = v;
        ++nh;
      }
    };
    double rNorm = std::sqrt(dot(R.as<double>(), R.as<double>()));          // :231
    const int itmax = prm->itmax > 0 ? prm->itmax : 2 * n;                 // :234
    push(rNorm);
    const double eps = prm->atol + prm->rtol * rNorm;                       // :239
    if (prm->verbose > 0) printf("BLOCK-BICGSTAB: system of size %d with %d right-hand sides\n", n, p);
    int iter = 0;
    bool solved = rNorm <= eps, tired = iter >= itmax, overtimed = false;
    while (!(solved || tired || overtimed)) {
      double *Yp = N ? Y.as<double>() : P.as<double>();
      if (N) precond(P.as<double>(), Yp);                                   // :251
      launch_spmm<double>(s, n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.ptr, false, p, Yp, p,
                          V.as<double>(), p, 1.0, 0.0);                      // :252
      proj(Cp, V.as<double>(), K.as<double>());                             // :253
      proj(Cp, R.as<double>(), al.as<double>());                            // :254-255
      lu(K.as<double>());                                                   // :256
      lu_solve(K.as<double>(), al.as<double>());                            // :257-258
      BPF_CHECK_HIP(hipMemcpyAsync(U.ptr, R.ptr, mb, hipMemcpyDeviceToDevice, s));           // :259
      gemm(V.as<double>(), al.as<double>(), U.as<double>(), -1.0, 1.0);     // :260
      gemm(Yp, al.as<double>(), X.as<double>(), 1.0, 1.0);                  // :261-262
      double *Yu = N ? Y.as<double>() : U.as<double>();
      if (N) precond(U.as<double>(), Yu);                                   // :263
      launch_spmm<double>(s, n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.ptr, false, p, Yu, p,
                          Q.as<double>(), p, 1.0, 0.0);                      // :264
      const double qu = dot(Q.as<double>(), U.as<double>());
      const double qq = dot(Q.as<double>(), Q.as<double>());
      const double omega = std::fabs(qu) / std::fabs(qq);                   // :265
      axpby(s, m, omega, Yu, 1.0, X.as<double>());                          // :266
      BPF_CHECK_HIP(hipMemcpyAsync(R.ptr, U.ptr, mb, hipMemcpyDeviceToDevice, s));
      axpby(s, m, -omega, Q.as<double>(), 1.0, R.as<double>());             // :267
      proj(Cp, Q.as<double>(), be.as<double>());                            // :268-269
      lu_solve(K.as<double>(), be.as<double>());                            // :270-271
      BPF_CHECK_HIP(hipMemcpyAsync(U.ptr, P.ptr, mb, hipMemcpyDeviceToDevice, s));
      axpby(s, m, -omega, V.as<double>(), 1.0, U.as<double>());             // :272
      BPF_CHECK_HIP(hipMemcpyAsync(P.ptr, R.ptr, mb, hipMemcpyDeviceToDevice, s));
      gemm(U.as<double>(), be.as<double>(), P.as<double>(), -1.0, 1.0);     // :273-274
      ++iter;
      rNorm = std::sqrt(dot(R.as<double>(), R.as<double>()));              // :277
      push(rNorm);
      solved = rNorm <= eps;
      tired = iter >= itmax;
      if (prm->timemax > 0 && now_s() - t0 > prm->timemax) overtimed = true;
      if (prm->verbose > 0 && iter % prm->verbose == 0) printf("%5d  %7.1e  %.2fs\n", iter, rNorm, now_s() - t0);
      if (!std::isfinite(rNorm)) break;
    }
    launch_rm_to_cm(s, n, p, X.ptr, p, false, d_X, ldx);
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    if (N) {
      prec_check(N, s);
    }
    bpf_solver_stats st{};
    st.niter = iter;
    st.solved = solved;
    st.status = overtimed ? BPF_STATUS_OVERTIMED : solved ? BPF_STATUS_SOLVED
                                                          : tired ? BPF_STATUS_TIRED : BPF_STATUS_UNKNOWN;
    st.nresiduals = nh;
    st.timer = now_s() - t0;
    st.max_berr = NAN;
    if (stats) *stats = st;
  });
}

extern "C" int bpf_block_bicgstab_t(bpf_ctx ctx, bpf_csr A, bpf_prec N, int dtype, int p, const void *d_B,
                                    int ldb, const void *d_C, int ldc, void *d_X, int ldx,
                                    const bpf_solver_params *prm, bpf_solver_stats *stats,
                                    double *h_hist, int hist_cap) {
  if (dtype == BPF_F64)
    return bpf_block_bicgstab(ctx, A, N, p, static_cast<const double *>(d_B), ldb,
                              static_cast<const double *>(d_C), ldc, static_cast<double *>(d_X), ldx, prm, stats,
                              h_hist, hist_cap);
  return guard([&] {
    BPF_REQUIRE(dtype == BPF_F32, "bad dtype %d", dtype);
    BPF_REQUIRE(ctx && A && prm, "null argument");
    BPF_REQUIRE(p > 0 && d_B && d_X, "empty block");
    BPF_REQUIRE(ldb >= A->n && ldx >= A->n, "leading dimension < n");
    if (d_C) BPF_REQUIRE(ldc >= A->n, "leading dimension of C < n");
    // Float32 blocks: widened to FP64 column-major copies, solved in FP64,
    // X rounded back to Float32.
    const int n = A->n;
    hipStream_t s = ctx->stream;
    DevBuf B(&ctx->pool), Cw(&ctx->pool), X(&ctx->pool);
    const size_t mb = (size_t)std::max(1, n) * p * sizeof(double);
    B.ensure(mb);
    X.ensure(mb);
    copy_cast<float, double>(s, p, n, static_cast<const float *>(d_B), ldb, B.as<double>(), n, 1.0);
    if (d_C) {
      Cw.ensure(mb);
      copy_cast<float, double>(s, p, n, static_cast<const float *>(d_C), ldc, Cw.as<double>(), n, 1.0);
    }
    const int rc = bpf_block_bicgstab(ctx, A, N, p, B.as<double>(), n, d_C ? Cw.as<double>() : nullptr, n,
                                      X.as<double>(), n, prm, stats, h_hist, hist_cap);
    if (rc != BPF_OK) throw Error{rc};
    copy_cast<double, float>(s, p, n, X.as<double>(), n, static_cast<float *>(d_X), ldx, 1.0);
    BPF_CHECK_HIP(hipStreamSynchronize(s));
  });
}
