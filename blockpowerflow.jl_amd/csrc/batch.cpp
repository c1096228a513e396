// Many independent sparse systems of uneven size solved together.
//
// Systems with the same RHS count p form a bin.  A bin stacks its systems by
// rows into one block-diagonal CSR operator with one ILU(0) factor, so SpMM,
// the factorization and the level-scheduled triangular solves each cover the
// whole bin in one launch (levels merge across systems).  Each system is one
// Krylov group of the block-Arnoldi engine (its own row range, Hessenberg,
// back substitution and convergence flag): the dense kernels also run once
// per iteration for every system still iterating, and a converged system
// drops out of the recurrence.  The reference solves one system per call
// (src/block_gmres.jl:170-411); this is an addition with per-system results
// identical in meaning to solving each system alone.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>

#include "krylov.h"

using namespace bpf;

namespace bpf {
std::atomic<int> g_batch_threads{4};
}

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Bin {
  bpf_ctx cctx = nullptr;               // the bin's own streams, buffer pool and scratch
  int p = 0;
  std::vector<int> sys;                 // member systems, in group order
  std::vector<GroupDesc> groups;        // row range and slot of each member
  int rows = 0, gmax = 0, nslots = 0;
  bpf_csr A = nullptr;
  bpf_prec N = nullptr;
  std::vector<double> anorm;            // ||A_j||_inf per member
  DevBuf d_groups, B, X, R, Xt, scale, stat, cmtab, ldtab;
  std::unique_ptr<BlockArnoldi> arn;
  long long arn_mode = -1;
  // Inner-solve width chosen by the autotuner (0 = not tuned yet) for
  // tuned_its inner iterations, and the width the last solve used.
  int tuned_width = 0, tuned_its = -1, last_width = 0;
  ~Bin() {
    arn.reset();
    if (N) bpf_prec_destroy(N);
    if (A) bpf_csr_destroy(A);
    for (DevBuf *x : {&d_groups, &B, &X, &R, &Xt, &scale, &stat, &cmtab, &ldtab}) x->release();
    if (cctx) bpf_ctx_destroy(cctx);
  }
};

}  // namespace

struct bpf_batch_s {
  bpf_ctx ctx = nullptr;
  int nsys = 0;
  std::vector<int> n, p;
  std::vector<std::unique_ptr<Bin>> bins;
  std::vector<int> bin_of, slot_of;     // system -> bin, group index in bin
};

namespace {

void rethrow(int rc) {
  if (rc != BPF_OK) throw Error{rc};
}

// Per-slot host copies of segmented column statistics.
void group_stats(bpf_ctx ctx, Bin &b, const double *X, int mode, std::vector<double> &out) {
  const size_t ns = std::max(1, b.nslots);
  b.stat.ensure((1 + (size_t)colstat_groups_blocks(b.gmax)) * ns * sizeof(double));
  colstat_groups<double>(ctx->stream, b.d_groups.as<GroupDesc>(), (int)b.groups.size(), b.gmax, b.p, X,
                         b.p, mode, b.stat.as<double>(), b.nslots, b.stat.as<double>() + ns);
  out.resize(b.nslots);
  BPF_CHECK_HIP(hipMemcpyAsync(out.data(), b.stat.ptr, b.nslots * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
}

struct BinResult {
  int iters = 0, cycles = 0;
  bool overtimed = false;
  double max_berr = 0.0;
};

// Block GMRES over one bin: FP64 basis; mixed = FP32 ILU(0) solves in the
// flexible form (the single-system BPF_PRECISION_MIXED path).
BinResult solve_bin(bpf_ctx ctx, Bin &b, const double *const *d_B, double *const *d_X,
                    const bpf_solver_params &prm, double t0, std::vector<char> &done_sys,
                    std::vector<double> &berr_sys) {
  BinResult res;
  hipStream_t s = ctx->stream;
  const int m = (int)b.groups.size(), p = b.p, rows = b.rows;
  const size_t blk = (size_t)rows * p * sizeof(double);
  for (DevBuf *x : {&b.B, &b.X, &b.R, &b.Xt}) x->ensure(blk);
  b.scale.ensure((size_t)b.nslots * sizeof(double));
  // Pack the members' column-major right-hand sides into the stacked block.
  {
    std::vector<double *> tab(m);
    std::vector<int> ld(m);
    for (int j = 0; j < m; ++j) {
      tab[j] = const_cast<double *>(d_B[b.sys[j]]);
      ld[j] = b.groups[j].nr;
    }
    upload(b.cmtab, tab, s);
    upload(b.ldtab, ld, s);
    launch_pack_groups(s, b.d_groups.as<GroupDesc>(), m, b.gmax, p, b.cmtab.as<double *>(),
                       b.ldtab.as<int>(), b.B.as<double>(), p, true);
  }
  fill_double(s, b.X.as<double>(), 0.0, (size_t)rows * p);
  BPF_CHECK_HIP(hipMemcpyAsync(b.R.ptr, b.B.ptr, blk, hipMemcpyDeviceToDevice, s));
  const bool mixed = prm.precision == BPF_PRECISION_MIXED;
  const bool use_berr = prm.berr_tol > 0.0;
  const int inner_its = mixed ? std::max(0, prm.inner_its) : 0;
  // columns per inner Krylov block (the whole system's p when not split)
  const int iw = inner_its > 0 && prm.inner_width > 0 && prm.inner_width < p ? prm.inner_width : p;
  const long long mode = mixed ? 1 + 2LL * inner_its + 4096LL * iw : 0;
  b.last_width = inner_its > 0 ? iw : 0;
  if (!b.arn || b.arn_mode != mode) {
    ArnoldiConfig cfg;
    cfg.n = rows;
    cfg.p = p;
    cfg.s = p;
    cfg.f32 = false;
    cfg.flexible = mixed;
    cfg.prec32 = mixed;
    cfg.orth = prm.reorthogonalization ? ORTH_BCGS2 : ORTH_BMGS;
    cfg.bmgs_blocks = std::max(1, prm.bmgs_blocks);
    cfg.groups = b.groups;
    cfg.inner_its = inner_its;
    cfg.inner_s = iw < p ? iw : 0;
    b.arn.reset();
    b.arn.reset(new BlockArnoldi(ctx, b.A, b.N, cfg));
    b.arn_mode = mode;
  }
  BlockArnoldi &arn = *b.arn;
  arn.set_scale(b.scale.as<double>());
  Poly poly;
  if (mixed && prm.poly_degree > 0 && inner_its == 0) {
    poly.build(ctx, b.A, b.N, prm.poly_degree);
    arn.set_poly(&poly);
  } else {
    arn.set_poly(nullptr);
  }
  std::vector<double> rn, bnorm, xnorm, rstat, xstat;
  group_stats(ctx, b, b.B.as<double>(), 0, bnorm);
  group_stats(ctx, b, b.R.as<double>(), 1, rn);
  for (auto &v : rn) v = std::sqrt(v);
  // eps_j = atol + rtol ||R0_j||_F per system (src/block_gmres.jl:230).
  std::vector<double> eps(m);
  for (int j = 0; j < m; ++j) {
    double f = 0.0;
    for (int c = 0; c < p; ++c) f += rn[b.groups[j].slot + c] * rn[b.groups[j].slot + c];
    eps[j] = prm.atol + prm.rtol * std::sqrt(f);
  }
  const int nmax = b.gmax;
  const int itmax = prm.itmax > 0 ? prm.itmax : 2 * (nmax / p) + 2;
  const int mem = std::max(1, std::min(nmax / p, prm.memory));
  std::vector<char> done(m, 0);
  std::vector<double> berr(b.nslots, INFINITY);
  auto check_berr = [&]() {
    residual_rm(s, b.A, p, b.B.as<double>(), b.X.as<double>(), b.R.as<double>());
    group_stats(ctx, b, b.R.as<double>(), 0, rstat);
    group_stats(ctx, b, b.X.as<double>(), 0, xstat);
    for (int j = 0; j < m; ++j) {
      double w = 0.0;
      for (int c = 0; c < p; ++c) {
        const int q = b.groups[j].slot + c;
        const double den = b.anorm[j] * xstat[q] + bnorm[q];
        berr[q] = den > 0.0 ? rstat[q] / den : (rstat[q] > 0.0 ? INFINITY : 0.0);
        w = std::max(w, berr[q]);
      }
      berr_sys[b.sys[j]] = w;
      if (use_berr) done[j] = w <= prm.berr_tol;
    }
  };
  if (use_berr) check_berr();
  else
    for (int j = 0; j < m; ++j) done[j] = std::sqrt([&] {
      double f = 0.0;
      for (int c = 0; c < p; ++c) f += rn[b.groups[j].slot + c] * rn[b.groups[j].slot + c];
      return f;
    }()) <= eps[j];
  int iter = 0;
  std::vector<double> sc(b.nslots, 1.0), target(b.nslots, 0.0);
  std::vector<char> active(m), have_x(m), est_ok(m);
  auto all_done = [&] {
    for (int j = 0; j < m; ++j)
      if (!done[j]) return false;
    return true;
  };
  while (!all_done() && iter < itmax && !res.overtimed) {
    ++res.cycles;
    // Converged systems sit this cycle out with no correction (K = 0).
    for (int j = 0; j < m; ++j) {
      active[j] = !done[j];
      have_x[j] = 0;
      est_ok[j] = 0;
      arn.set_active(j, active[j], active[j] ? (1 << 30) : 0);
    }
    arn.upload_groups();
    for (int q = 0; q < b.nslots; ++q) sc[q] = rn[q] > 0.0 ? rn[q] : 1.0;
    BPF_CHECK_HIP(hipMemcpyAsync(b.scale.ptr, sc.data(), b.nslots * sizeof(double), hipMemcpyHostToDevice, s));
    launch_scale_groups(s, b.d_groups.as<GroupDesc>(), m, b.gmax, p, b.R.as<double>(), p,
                        b.scale.as<double>(), true);
    const int maxit = prm.restart ? std::min(mem, itmax - iter) : itmax - iter;
    auto stop = [&](const double *r, int it) {
      bool changed = false, need_peek = false;
      std::vector<double> worst(m, 0.0);
      for (int j = 0; j < m; ++j) {
        if (!active[j]) continue;
        for (int c = 0; c < p; ++c) worst[j] = std::max(worst[j], r[b.groups[j].slot + c]);
        if (use_berr && !have_x[j] && worst[j] < 1e-3) need_peek = true;
      }
      if (need_peek) {
        // Turn the backward-error tolerance into residual targets once the
        // solution estimate is meaningful: ||r_c|| <= tol (||A|| ||x_c|| + ||b_c||).
        arn.peek(b.X.as<double>(), b.Xt.as<double>());
        group_stats(ctx, b, b.Xt.as<double>(), 0, xnorm);
        for (int j = 0; j < m; ++j) {
          if (!active[j] || have_x[j] || worst[j] >= 1e-2) continue;
          for (int c = 0; c < p; ++c) {
            const int q = b.groups[j].slot + c;
            target[q] = prm.berr_tol * (b.anorm[j] * xnorm[q] + bnorm[q]);
          }
          have_x[j] = 1;
        }
      }
      int left = 0;
      for (int j = 0; j < m; ++j) {
        if (!active[j]) continue;
        bool ok;
        if (use_berr) {
          ok = have_x[j];
          for (int c = 0; c < p && ok; ++c) {
            const int q = b.groups[j].slot + c;
            ok = r[q] * sc[q] <= target[q];
          }
        } else {
          double f = 0.0;
          for (int c = 0; c < p; ++c) {
            const int q = b.groups[j].slot + c;
            f += r[q] * sc[q] * r[q] * sc[q];
          }
          ok = std::sqrt(f) <= eps[j];
        }
        // A system whose Krylov space is exhausted (n_j / p block steps, the
        // memory cap of a single-system solve) also stops; the next cycle
        // restarts it from the true residual.
        const bool full = it >= std::max(1, b.groups[j].nr / p);
        if (ok || full) {
          est_ok[j] = ok;
          active[j] = 0;
          arn.set_active(j, false, it);
          changed = true;
        } else {
          ++left;
        }
      }
      if (changed) arn.upload_groups();
      if (prm.timemax > 0 && now_s() - t0 > prm.timemax) res.overtimed = true;
      return left == 0 || res.overtimed;
    };
    const int its = arn.run(b.R.as<double>(), b.X.as<double>(), maxit, stop, nullptr);
    iter += its;
    if (use_berr) {
      check_berr();
    } else {
      residual_rm(s, b.A, p, b.B.as<double>(), b.X.as<double>(), b.R.as<double>());
    }
    group_stats(ctx, b, b.R.as<double>(), 1, rn);
    for (auto &v : rn) v = std::sqrt(v);
    if (!use_berr) {
      // After a Cholesky breakdown in CholQR2 the residual estimates are not
      // reliable: the true residual decides (src/block_gmres.jl:349's rule
      // applied to it), and unsolved systems restart from it.
      const bool trust = !arn.shifted();
      for (int j = 0; j < m; ++j) {
        if (done[j]) continue;
        if (trust) {
          done[j] = est_ok[j];
        } else {
          double f = 0.0;
          for (int c = 0; c < p; ++c) f += rn[b.groups[j].slot + c] * rn[b.groups[j].slot + c];
          done[j] = std::sqrt(f) <= eps[j];
        }
      }
    }
    if (prm.timemax > 0 && now_s() - t0 > prm.timemax) res.overtimed = true;
    if (its == 0) break;
  }
  if (!use_berr) check_berr();
  // Unpack the solutions.
  {
    std::vector<double *> tab(m);
    for (int j = 0; j < m; ++j) tab[j] = d_X[b.sys[j]];
    upload(b.cmtab, tab, s);
    launch_pack_groups(s, b.d_groups.as<GroupDesc>(), m, b.gmax, p, b.cmtab.as<double *>(),
                       b.ldtab.as<int>(), b.X.as<double>(), p, false);
    BPF_CHECK_HIP(hipStreamSynchronize(s));
  }
  arn.set_poly(nullptr);
  for (int j = 0; j < m; ++j) {
    done_sys[b.sys[j]] = done[j];
    res.max_berr = std::max(res.max_berr, berr_sys[b.sys[j]]);
  }
  res.iters = iter;
  return res;
}

// solve_bin with the inner-solve width chosen by timing (inner_width < 0):
// on a bin's first such solve each candidate width (8, 16, 32, 64 up to the
// bin's RHS count) solves the bin completely, the fastest to solve every
// system is kept for this bin and number of inner iterations, and the
// bin is solved once more with it if it was not the last candidate.  Every
// candidate run is a full solve (the right-hand sides are read again), so
// the answer is the chosen width's.
BinResult solve_bin_tuned(bpf_ctx ctx, Bin &b, const double *const *d_B, double *const *d_X,
                          const bpf_solver_params &prm, double t0, std::vector<char> &done_sys,
                          std::vector<double> &berr_sys) {
  const bool mixed = prm.precision == BPF_PRECISION_MIXED;
  if (prm.inner_width >= 0 || !mixed || prm.inner_its <= 0) {
    bpf_solver_params q = prm;
    q.inner_width = std::max(0, prm.inner_width);
    return solve_bin(ctx, b, d_B, d_X, q, t0, done_sys, berr_sys);
  }
  bpf_solver_params q = prm;
  if (b.tuned_width > 0 && b.tuned_its == prm.inner_its) {
    q.inner_width = b.tuned_width;
    return solve_bin(ctx, b, d_B, d_X, q, t0, done_sys, berr_sys);
  }
  std::vector<int> cand;
  for (int w : {8, 16, 32, 64})
    if (w <= b.p) cand.push_back(w);
  if (cand.empty() || cand.back() < b.p) cand.push_back(b.p);
  int best = cand.back();
  double best_t = INFINITY;
  BinResult res;
  for (int w : cand) {
    q.inner_width = w;
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    const double ts = now_s();
    res = solve_bin(ctx, b, d_B, d_X, q, t0, done_sys, berr_sys);
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    const double dt = now_s() - ts;
    bool all = true;
    for (int i : b.sys) all = all && done_sys[i];
    if (all && dt < best_t) {
      best_t = dt;
      best = w;
    }
  }
  b.tuned_width = best;
  b.tuned_its = prm.inner_its;
  if (best != cand.back()) {
    q.inner_width = best;
    res = solve_bin(ctx, b, d_B, d_X, q, t0, done_sys, berr_sys);
  }
  return res;
}

}  // namespace

extern "C" {

int bpf_batch_create(bpf_ctx ctx, int nsys, const int32_t *n, const int32_t *nnz,
                     const int32_t *const *h_rowptr, const int32_t *const *h_colidx,
                     const double *const *h_val, const int32_t *p, bpf_batch *out) {
  return guard([&] {
    BPF_REQUIRE(ctx && out && n && nnz && h_rowptr && h_colidx && h_val && p, "null argument");
    BPF_REQUIRE(nsys > 0, "empty batch");
    auto bt = std::make_unique<bpf_batch_s>();
    bt->ctx = ctx;
    bt->nsys = nsys;
    bt->n.assign(n, n + nsys);
    bt->p.assign(p, p + nsys);
    bt->bin_of.assign(nsys, -1);
    bt->slot_of.assign(nsys, -1);
    std::map<int, std::vector<int>> by_p;
    for (int i = 0; i < nsys; ++i) {
      BPF_REQUIRE(n[i] > 0, "system %d: n must be positive", i);
      BPF_REQUIRE(p[i] > 0 && p[i] <= 128, "system %d: p = %d outside 1..128", i, p[i]);
      BPF_REQUIRE(h_rowptr[i] && (nnz[i] == 0 || (h_colidx[i] && h_val[i])), "system %d: null CSR array", i);
      BPF_REQUIRE(h_rowptr[i][0] == 0 && h_rowptr[i][n[i]] == nnz[i],
                  "system %d: rowptr must start at 0 and end at nnz", i);
      for (int k = 0; k < nnz[i]; ++k)
        BPF_REQUIRE(h_colidx[i][k] >= 0 && h_colidx[i][k] < n[i], "system %d: column index out of range", i);
      by_p[p[i]].push_back(i);
    }
    for (auto &kv : by_p) {
      auto b = std::make_unique<Bin>();
      b->p = kv.first;
      b->sys = kv.second;
      // Largest systems first: their groups lead the grids.
      std::stable_sort(b->sys.begin(), b->sys.end(), [&](int x, int y) { return n[x] > n[y]; });
      size_t tot_nnz = 0;
      for (int i : b->sys) tot_nnz += (size_t)nnz[i];
      for (int i : b->sys) BPF_REQUIRE((size_t)b->rows + n[i] < (size_t)INT32_MAX, "bin too large");
      std::vector<int32_t> rp, ci;
      std::vector<double> va;
      rp.reserve(1);
      rp.push_back(0);
      ci.reserve(tot_nnz);
      va.reserve(tot_nnz);
      BPF_REQUIRE(tot_nnz < (size_t)INT32_MAX, "bin has too many nonzeros");
      int r0 = 0;
      for (size_t j = 0; j < b->sys.size(); ++j) {
        const int i = b->sys[j];
        double an = 0.0;
        for (int r = 0; r < n[i]; ++r) {
          double rs = 0.0;
          for (int k = h_rowptr[i][r]; k < h_rowptr[i][r + 1]; ++k) {
            ci.push_back(h_colidx[i][k] + r0);
            va.push_back(h_val[i][k]);
            rs += std::fabs(h_val[i][k]);
          }
          an = std::max(an, rs);
          rp.push_back((int32_t)ci.size());
        }
        b->anorm.push_back(an);
        GroupDesc g{0, b->p, r0, n[i], (int)j * b->p, 1 << 30, 1, 0};
        b->groups.push_back(g);
        bt->bin_of[i] = (int)bt->bins.size();
        bt->slot_of[i] = (int)j;
        b->gmax = std::max(b->gmax, n[i]);
        r0 += n[i];
      }
      b->rows = r0;
      b->nslots = (int)b->sys.size() * b->p;
      // Each bin runs on a context of its own, so bins can be solved at the
      // same time from separate host threads.
      rethrow(bpf_ctx_create(ctx->device, &b->cctx));
      rethrow(bpf_csr_create(b->cctx, b->rows, (int)ci.size(), rp.data(), ci.data(), va.data(), &b->A));
      rethrow(bpf_prec_create(b->cctx, b->A, BPF_PREC_ILU0, &b->N));
      upload(b->d_groups, b->groups, b->cctx->stream);
      BPF_CHECK_HIP(hipStreamSynchronize(b->cctx->stream));
      bt->bins.push_back(std::move(b));
    }
    *out = bt.release();
  });
}

int bpf_batch_solve(bpf_batch bt, const double *const *d_B, double *const *d_X,
                    const bpf_solver_params *prm, bpf_solver_stats *stats, int32_t *h_solved) {
  return guard([&] {
    BPF_REQUIRE(bt && d_B && d_X && prm, "null argument");
    BPF_REQUIRE(prm->precision == BPF_PRECISION_FP64 || prm->precision == BPF_PRECISION_MIXED,
                "batched solve supports precision FP64 or MIXED");
    for (int i = 0; i < bt->nsys; ++i) BPF_REQUIRE(d_B[i] && d_X[i], "system %d: null block", i);
    const double t0 = now_s();
    std::vector<char> done(bt->nsys, 0);
    std::vector<double> berr(bt->nsys, INFINITY);
    bpf_solver_stats st{};
    bool overtimed = false;
    int max_cycles = 0;
    // The bins' streams start after the work queued on the caller's stream
    // (which may have produced the right-hand sides).
    hipEvent_t ready = nullptr;
    BPF_CHECK_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    struct EventGuard {
      hipEvent_t e;
      ~EventGuard() { (void)hipEventDestroy(e); }
    } ready_guard{ready};
    BPF_CHECK_HIP(hipEventRecord(ready, bt->ctx->stream));
    const int nbin = (int)bt->bins.size();
    std::vector<BinResult> res(nbin);
    auto run = [&](int i) {
      Bin &b = *bt->bins[i];
      BPF_CHECK_HIP(hipStreamWaitEvent(b.cctx->stream, ready, 0));
      // Refresh the factor: the caller may have changed nothing, but a bin is
      // factored once per solve so repeated solves time the same work.
      rethrow(bpf_prec_factor(b.N, b.A));
      res[i] = solve_bin_tuned(b.cctx, b, d_B, d_X, *prm, t0, done, berr);
      BPF_CHECK_HIP(hipStreamSynchronize(b.cctx->stream));
      if (b.cctx->stream2) BPF_CHECK_HIP(hipStreamSynchronize(b.cctx->stream2));
    };
    // Bins write disjoint systems' entries of done / berr and their own
    // blocks, so they run on up to g_batch_threads host threads at once;
    // a solve in which some bin times its inner-width candidates runs the
    // bins one after another, so that no bin's timings include another's.
    bool tuning = false;
    if (prm->precision == BPF_PRECISION_MIXED && prm->inner_width < 0 && prm->inner_its > 0)
      for (const auto &b : bt->bins) tuning = tuning || !(b->tuned_width > 0 && b->tuned_its == prm->inner_its);
    const int nthr = tuning ? 1 : std::max(1, std::min<int>(nbin, g_batch_threads));
    const int dev = bt->ctx->device;
    run_pool(nbin, nthr, run, [dev] { BPF_CHECK_HIP(hipSetDevice(dev)); });
    for (const BinResult &r : res) {
      st.niter += r.iters;
      max_cycles = std::max(max_cycles, r.cycles);
      overtimed = overtimed || r.overtimed;
      st.max_berr = std::max(st.max_berr, r.max_berr);
    }
    int nsolved = 0;
    for (int i = 0; i < bt->nsys; ++i) {
      if (h_solved) h_solved[i] = done[i];
      nsolved += done[i];
    }
    st.solved = nsolved == bt->nsys;
    st.status = st.solved ? BPF_STATUS_SOLVED : overtimed ? BPF_STATUS_OVERTIMED : BPF_STATUS_TIRED;
    st.refine_steps = std::max(0, max_cycles - 1);
    st.timer = now_s() - t0;
    if (stats) *stats = st;
  });
}

int bpf_batch_inner_widths(bpf_batch bt, int32_t *h_width) {
  return guard([&] {
    BPF_REQUIRE(bt && h_width, "null argument");
    for (int i = 0; i < bt->nsys; ++i) h_width[i] = bt->bins[bt->bin_of[i]]->last_width;
  });
}

int bpf_batch_destroy(bpf_batch bt) {
  delete bt;
  return BPF_OK;
}

}  // extern "C"
