// Block-Arnoldi engine: grouped block GMRES iterations on the device.
//
// Follows the structure of block_gmres! (src/block_gmres.jl:251-393):
//   P = N^{-1} V_k                      (:297-298)
//   W = A P                             (:299)
//   block Gram-Schmidt against V_0..V_k (:301-313; classical block form,
//                                        optional second pass = the
//                                        reference's `reorthogonalization`)
//   V_{k+1} C = W  (thin QR)            (:316; CholQR2 in place of geqrf/orgqr)
//   apply the stored reflectors and form the new one on [Psi_kk; C],
//   rotate the right-hand side          (:318-337)
//   residual estimate ||Lambda||        (:341-342, per column here)
//   block back substitution, X += Z Y   (:366-385)
#include "krylov.h"
#include "profile.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bpf {

std::atomic<bool> g_overlap_hess{true};
std::atomic<bool> g_graph{true};
std::atomic<long long> g_graph_replays{0};
constexpr int kOverlapMaxIt = 1 << 16;

void *BlockArena::block(int i) {
  while (i >= (int)ptrs.size()) {
    DevBuf b(pool);
    b.ensure(block_elems * esize * chunk);
    for (int q = 0; q < chunk; ++q) ptrs.push_back(static_cast<char *>(b.ptr) + q * block_elems * esize);
    chunks.push_back(std::move(b));
  }
  return ptrs[i];
}

PtrTable::~PtrTable() {
  if (h) (void)hipHostFree(h);
}

void PtrTable::set(int i, void *ptr, hipStream_t s) {
  if (i >= cap) {
    int nc = std::max(64, cap * 2);
    while (nc <= i) nc *= 2;
    void **nh = nullptr;
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    BPF_CHECK_HIP(hipHostMalloc((void **)&nh, nc * sizeof(void *), hipHostMallocDefault));
    std::memset(nh, 0, nc * sizeof(void *));
    if (h) {
      std::memcpy(nh, h, cap * sizeof(void *));
      (void)hipHostFree(h);
    }
    h = nh;
    DevBuf nd;
    nd.ensure(nc * sizeof(void *));
    if (cap > 0)
      BPF_CHECK_HIP(hipMemcpyAsync(nd.ptr, h, cap * sizeof(void *), hipMemcpyHostToDevice, s));
    d = std::move(nd);
    cap = nc;
  }
  if (h[i] == ptr) return;
  h[i] = ptr;
  BPF_CHECK_HIP(hipMemcpyAsync(d.as<void *>() + i, h + i, sizeof(void *), hipMemcpyHostToDevice, s));
}

void PtrTable::reserve(int n, hipStream_t s) {
  if (n > cap) set(n - 1, nullptr, s);   // grows the table; the new entry stays null
}

BlockArnoldi::BlockArnoldi(bpf_ctx ctx, bpf_csr A, bpf_prec N, const ArnoldiConfig &cfg)
    : ctx_(ctx), A_(A), N_(N), cfg_(cfg), st_(ctx->stream) {
  BPF_REQUIRE(cfg.p > 0 && cfg.n > 0, "empty block");
  if (!cfg.groups.empty()) {
    h_groups_ = cfg.groups;
    int wmax = 0;
    for (const auto &g : h_groups_) {
      wmax = std::max(wmax, g.w);
      nslots_ = std::max(nslots_, g.slot + g.w);
      gmax_rows_ = std::max(gmax_rows_, g.nr);
      rows_sum_ += g.nr;
    }
    S_ = pad_width(wmax);
    BPF_REQUIRE(S_ > 0, "block width %d unsupported", wmax);
  } else {
    int s = cfg.s > 0 ? std::min(cfg.s, cfg.p) : cfg.p;
    if (s > 128) {
      // Split into near-equal groups of at most 128 columns.
      const int ng = (cfg.p + 127) / 128;
      s = (cfg.p + ng - 1) / ng;
    }
    S_ = pad_width(s);
    BPF_REQUIRE(S_ > 0, "block width %d unsupported", s);
    for (int c0 = 0; c0 < cfg.p; c0 += s)
      h_groups_.push_back({c0, std::min(s, cfg.p - c0), 0, cfg.n, c0, 1 << 30, 1, 0});
    nslots_ = cfg.p;
    gmax_rows_ = cfg.n;
  }
  ng_ = (int)h_groups_.size();
  // Vector loads of four columns: every group's offset and width and the
  // row length are multiples of 4 (blocks are then 16-byte aligned: each
  // starts n p elements after an aligned allocation).
  vec4_ = cfg.p % 4 == 0;
  for (const auto &g : h_groups_) vec4_ = vec4_ && g.c0 % 4 == 0 && g.w % 4 == 0;
  // per-solve work buffers come from the context's pool
  for (DevBuf *b : {&d_groups_, &W_, &P_, &E_, &rnorm_, &Y_, &Zb_, &Rinvb_, &Xtmp_, &Vf_, &Zf_, &tmpcol_,
                    &cq_.G, &cq_.R1, &cq_.R1i, &cq_.R2, &cq_.R2i, &cq_.R3, &cq_.R3i, &cq_.flags, &cq_.part,
                    &cq_.gflags, &cq_.gsets})
    b->pool = &ctx->pool;
  V_.pool = Z_.pool = &ctx->pool;
  upload(d_groups_, h_groups_, st_);
  const size_t es = cfg.f32 ? sizeof(float) : sizeof(double);
  const size_t blk = (size_t)cfg.n * cfg.p;
  V_.init(blk, es, 4);
  if (cfg.flexible && N) Z_.init(blk, es, 4);
  W_.ensure(blk * es);
  P_.ensure(blk * es);
  const size_t ss = (size_t)ng_ * S_ * S_ * sizeof(double);
  cq_.init(ng_, S_);
  E_.ensure(2 * ss);
  rnorm_.ensure(std::max(1, nslots_) * sizeof(double));
  flags_.ensure(16);
  ctx->ensure_small(std::max(nslots_, 16));
  if (cfg.nested) {
    BPF_CHECK_HIP(hipEventCreateWithFlags(&own_ev_, hipEventDisableTiming));
    BPF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void **>(&own_rn_), std::max(1, nslots_) * sizeof(double)));
    ev_hess_ = own_ev_;
  } else {
    ctx->ensure_stream2(std::max(1, nslots_));
    hs_ = ctx->stream2;
    ev_main_ = ctx->ev_a;
    ev_hess_ = ctx->ev_b;
  }
  if (cfg.inner_its > 0 && cfg.flexible && N && !cfg.f32) {
    ArnoldiConfig ic;
    ic.n = cfg.n;
    ic.p = cfg.p;
    ic.s = cfg.inner_s;
    ic.f32 = true;
    ic.flexible = false;
    // one classical Gram-Schmidt pass over all inner blocks and one CholQR
    // pass: the inner basis only has to be near-orthonormal (measured at
    // 70,000 buses: the same outer iterations as two passes of each, 0.67 s
    // against 0.73 s per solve)
    ic.orth = ORTH_BMGS;
    ic.bmgs_blocks = std::max(8, cfg.inner_its);
    ic.cholqr_passes = 1;
    ic.nested = true;
    if (!cfg.groups.empty() && cfg.inner_s > 0) {
      // stacked systems: each system's columns split into inner groups of
      // at most inner_s columns (independent Krylov blocks over its rows)
      for (const auto &g : cfg.groups)
        for (int c = 0; c < g.w; c += cfg.inner_s)
          ic.groups.push_back({g.c0 + c, std::min(cfg.inner_s, g.w - c), g.r0, g.nr, g.slot + c, g.K, g.active, 0});
    } else {
      ic.groups = cfg.groups;
    }
    inner_ = std::make_unique<BlockArnoldi>(ctx, A, N, ic);
  }
}

void BlockArnoldi::set_active(int g, bool active, int K) {
  h_groups_[g].active = active ? 1 : 0;
  h_groups_[g].K = K;
}

void BlockArnoldi::upload_groups() { upload(d_groups_, h_groups_, st_); }

BlockArnoldi::~BlockArnoldi() {
  // the second stream, its events and h_rn_ belong to the context
  if (gexec_) {
    (void)hipStreamSynchronize(st_);
    (void)hipGraphExecDestroy(gexec_);
  }
  if (hs_) (void)hipStreamSynchronize(hs_);
  if (h_fl_ || own_rn_ || own_ev_) (void)hipStreamSynchronize(st_);
  if (h_fl_) (void)hipHostFree(h_fl_);
  if (own_rn_) (void)hipHostFree(own_rn_);
  if (own_ev_) (void)hipEventDestroy(own_ev_);
}

double *BlockArnoldi::col_region(int k) {
  while ((int)cols_.size() <= k) {
    const int kk = (int)cols_.size();
    DevBuf b(&ctx_->pool);
    b.ensure(hcol_elems(kk, ng_, S_) * sizeof(double));
    coltab_.set(kk, b.ptr, st_);
    cols_.push_back(std::move(b));
  }
  return cols_[k].as<double>();
}

void CholQrWork::init(int ng, int S) {
  const size_t ss = (size_t)ng * S * S * sizeof(double);
  G.ensure(ss); R1.ensure(ss); R1i.ensure(ss); R2.ensure(ss); R2i.ensure(ss); R3.ensure(ss); R3i.ensure(ss);
  flags.ensure(16);
  gflags.ensure(std::max(1, ng) * sizeof(int));
  gsets.ensure(2 * (size_t)std::max(1, ng) * sizeof(GroupDesc));
}

template <typename TV>
void cholqr2_groups(hipStream_t s, const GroupDesc *dg, int ng, int S, int p, int gmax_rows, int rows_sum,
                    int n_shift, void *W, void *Vout, double *Cout, size_t cstride, CholQrWork &w,
                    const ReorthFn &reorth, bool host_check, int passes, int vec4) {
  // Two CholQR passes; C = R2 R1 (upper), Vout = W R1^{-1} R2^{-1}.
  const size_t SS = (size_t)S * S;
  int *fl = w.flags.as<int>();
  fill_int(s, fl + 1, 0, 1);
  w.gflags.ensure(std::max(1, ng) * sizeof(int));
  w.gsets.ensure(2 * (size_t)std::max(1, ng) * sizeof(GroupDesc));
  DevBuf *Rs[3] = {&w.R1, &w.R2, &w.R3};
  DevBuf *Ris[3] = {&w.R1i, &w.R2i, &w.R3i};
  // cells < 0: every group takes part (kernel-timer accounting); the passes
  // over a subset of groups chosen on the device are counted as no traffic
  auto pass = [&](int i, void *src, void *dst, bool deflate, const GroupDesc *grp, double cells) {
    ProjArgs pa{nullptr, src, 0, 1, p, src, p, gmax_rows, grp, ng, 0, nullptr, rows_sum};
    pa.cells = cells;
    launch_proj<TV, TV>(s, pa, S, w.G.as<double>(), 0, SS, false, w.part);
    CholArgs ca{w.G.as<double>(), Rs[i]->as<double>(), Ris[i]->as<double>(), SS, S, n_shift, grp, fl};
    ca.deflate = deflate;
    ca.gflags = w.gflags.as<int>();
    launch_chol(s, ca, ng);
    GemmArgs ga{nullptr, src, 0, 1, p, Ris[i]->as<double>(), 0, SS, S, dst, p, gmax_rows, 1.0, 0.0, grp};
    ga.vec4 = vec4;
    ga.rows_sum = rows_sum;
    ga.cells = cells;
    ga.keep_skipped = grp != dg;   // a subset pass leaves the other groups alone
    launch_block_gemm<TV, TV>(s, ga, S, ng);
  };
  // The first pass drops columns whose Cholesky pivot is not positive (they
  // lie in the span of the columns before them: linearly dependent
  // right-hand sides, an exhausted Krylov space) and fills them with
  // pseudo-random directions, as Householder QR completes a rank-deficient
  // block with arbitrary orthonormal ones; W = Q1 R1 holds with zero rows of
  // R1 there.
  if (passes == 1) {
    // one pass: Vout = W R1^{-1}, C = R1 (deflated columns completed in Vout)
    pass(0, W, Vout, true, dg, -1.0);
    launch_fill_deflated<TV>(s, static_cast<TV *>(Vout), p, w.R1.as<double>(), S, SS, dg, ng, gmax_rows, fl + 1);
    copy_cast<double, double>(s, ng, (int)SS, w.R1.as<double>(), (int)SS, Cout, (int)cstride, 1.0);
    return;
  }
  pass(0, W, W, true, dg, -1.0);
  launch_fill_deflated<TV>(s, static_cast<TV *>(W), p, w.R1.as<double>(), S, SS, dg, ng, gmax_rows, fl + 1);
  // The host learns whether any group broke down (one 4-byte read); the
  // common case, none, is the plain second pass over all groups.
  int h = 0;
  if (host_check) {
    BPF_CHECK_HIP(hipMemcpyAsync(&h, fl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
  }
  if (!(h & 6)) {
    pass(1, W, Vout, false, dg, -1.0);
    launch_trmm_upper(s, w.R2.as<double>(), w.R1.as<double>(), Cout, S, SS, cstride, dg, ng);
    return;
  }
  // Otherwise the groups are split on the device by their own bits, so no
  // group's arithmetic depends on another's.
  GroupDesc *brk = w.gsets.as<GroupDesc>(), *clean = brk + ng;
  launch_split_groups(s, dg, w.gflags.as<int>(), ng, brk);
  // Groups without a breakdown: the second pass.
  pass(1, W, Vout, false, clean, 0.0);
  launch_trmm_upper(s, w.R2.as<double>(), w.R1.as<double>(), Cout, S, SS, cstride, clean, ng);
  // Groups with one: the new directions are orthogonal to nothing yet, and
  // amplified noise may lean on the earlier basis.  The block is projected
  // off the earlier basis (the other columns only move at rounding level, W
  // having been projected already), then two passes make it orthonormal;
  // C = R3 R2 R1, the product staged in G.
  if (reorth) reorth(W, brk);
  pass(1, W, W, false, brk, 0.0);
  pass(2, W, Vout, false, brk, 0.0);
  launch_trmm_upper(s, w.R3.as<double>(), w.R2.as<double>(), w.G.as<double>(), S, SS, SS, brk, ng);
  launch_trmm_upper(s, w.G.as<double>(), w.R1.as<double>(), Cout, S, SS, cstride, brk, ng);
}
template void cholqr2_groups<double>(hipStream_t, const GroupDesc *, int, int, int, int, int, int, void *,
                                     void *, double *, size_t, CholQrWork &, const ReorthFn &, bool, int, int);
template void cholqr2_groups<float>(hipStream_t, const GroupDesc *, int, int, int, int, int, int, void *,
                                    void *, double *, size_t, CholQrWork &, const ReorthFn &, bool, int, int);

template <typename TV>
void BlockArnoldi::cholqr2(void *W, void *Vout, double *Cout, size_t cstride, int kprev) {
  ReorthFn reorth;
  if (kprev >= 0) {
    // one Gram-Schmidt pass of the block against V_0..V_kprev for the groups
    // given, in steps of at most 8 basis blocks (bounded partial sums);
    // traffic not counted (the subset is chosen on the device)
    reorth = [this, kprev](void *Wp, const GroupDesc *dg) {
      constexpr int kStep = 8;
      const size_t SS = (size_t)S_ * S_, gstr = (size_t)kStep * SS;
      const void *const *vt = vtab_.dev<const void>();
      const size_t need = hcol_elems(kStep - 2, ng_, S_) * sizeof(double);
      tmpcol_.ensure(need);
      double *t = tmpcol_.as<double>();
      for (int i = 0; i <= kprev; i += kStep) {
        const int nb = std::min(kStep, kprev + 1 - i);
        ProjArgs pa{vt, nullptr, i, nb, cfg_.p, Wp, cfg_.p, gmax_rows_, dg, ng_, 0, nullptr, rows_sum_};
        pa.cells = 0.0;
        launch_proj<TV, TV>(st_, pa, S_, t, SS, gstr, false, part_);
        GemmArgs ga{vt, nullptr, i, nb, cfg_.p, t, SS, gstr, S_, Wp, cfg_.p, gmax_rows_, -1.0, 1.0, dg};
        ga.vec4 = vec4_;
        ga.rows_sum = rows_sum_;
        ga.cells = 0.0;
        launch_block_gemm<TV, TV>(st_, ga, S_, ng_);
      }
    };
  }
  // a nested run (a preconditioner step) never waits on the host
  cholqr2_groups<TV>(st_, d_groups_.as<GroupDesc>(), ng_, S_, cfg_.p, gmax_rows_, rows_sum_, cfg_.n, W, Vout,
                     Cout, cstride, cq_, reorth, !cfg_.nested, cfg_.cholqr_passes, vec4_);
}

// Psi_{i,k} = V_i^T W and W -= V_i Psi_{i,k}, i = 0..k, into block column k.
// BMGS (the reference's loop, src/block_gmres.jl:301-304) projects against
// one block at a time; BCGS projects against all blocks at once (one read of
// the basis per pass); BCGS2 repeats the pass and accumulates (:307-313).
template <typename TV>
void BlockArnoldi::orthogonalize(int k, double *col) {
  const GroupDesc *dg = d_groups_.as<GroupDesc>();
  const int n = cfg_.n, p = cfg_.p;
  const size_t SS = (size_t)S_ * S_;
  const size_t gstr = (size_t)(k + 2) * SS;
  const void *const *vt = vtab_.dev<const void>();
  // work of the groups still iterating (kernel traffic accounting)
  double cells = 0.0;
  for (const auto &g : h_groups_)
    if (g.active) cells += (double)(rows_sum_ > 0 ? g.nr : n) * S_;
  if (cfg_.orth == ORTH_BMGS) {
    // Block MGS over steps of g consecutive basis blocks (g = 1: the
    // reference's loop); within a step the blocks are projected together.
    const int g = std::max(1, cfg_.bmgs_blocks);
    for (int i = 0; i <= k; i += g) {
      const int nb = std::min(g, k + 1 - i);
      ProjArgs pa{vt, nullptr, i, nb, p, W_.ptr, p, gmax_rows_, dg, ng_, 0, nullptr, rows_sum_};
      pa.cells = cells;
      // an inner (preconditioner) run's projection coefficients: FP32 sums
      // per row chunk (its CholQR Gram matrices stay FP64)
      pa.f32acc = cfg_.nested ? 1 : 0;
      launch_proj<TV, TV>(st_, pa, S_, col + (size_t)i * SS, SS, gstr, false, part_);
      GemmArgs ga{vt, nullptr, i, nb, p, col + (size_t)i * SS, SS, gstr, S_, W_.ptr, p, gmax_rows_, -1.0, 1.0, dg};
      ga.vec4 = vec4_;
      ga.rows_sum = rows_sum_;
      ga.cells = cells;
      launch_block_gemm<TV, TV>(st_, ga, S_, ng_);
    }
    return;
  }
  ProjArgs pa{vt, nullptr, 0, k + 1, p, W_.ptr, p, gmax_rows_, dg, ng_, 0, nullptr, rows_sum_};
  pa.cells = cells;
  launch_proj<TV, TV>(st_, pa, S_, col, SS, gstr, false, part_);
  GemmArgs ga{vt, nullptr, 0, k + 1, p, col, SS, gstr, S_, W_.ptr, p, gmax_rows_, -1.0, 1.0, dg};
  ga.vec4 = vec4_;
  ga.rows_sum = rows_sum_;
  ga.cells = cells;
  launch_block_gemm<TV, TV>(st_, ga, S_, ng_);
  if (cfg_.orth == ORTH_BCGS2) {
    tmpcol_.ensure(hcol_elems(k, ng_, S_) * sizeof(double));
    double *t = tmpcol_.as<double>();
    launch_proj<TV, TV>(st_, pa, S_, t, SS, gstr, false, part_);
    GemmArgs gb = ga;
    gb.psi = t;
    gb.rows_sum = rows_sum_;
    launch_block_gemm<TV, TV>(st_, gb, S_, ng_);
    axpby_groups(st_, ng_, (size_t)(k + 1) * SS, gstr, t, col);
  }
}

template <typename TV>
void BlockArnoldi::form(double *X, double alpha) {
  // Y = R^{-1} Z (block back substitution), then X += alpha * sum_i Z_i Y_i
  // (flexible) or X += alpha * N^{-1} sum_i V_i Y_i.
  const int K = k_;
  if (K <= 0) return;
  const GroupDesc *dg = d_groups_.as<GroupDesc>();
  const int n = cfg_.n, p = cfg_.p;
  const size_t SS = (size_t)S_ * S_;
  Y_.ensure((size_t)K * ng_ * SS * sizeof(double));
  // groups still iterating use all K block columns
  bool changed = false;
  for (auto &g : h_groups_)
    if (g.active && g.K != K) { g.K = K; changed = true; }
  if (changed) upload_groups();
  Zb_.ensure((size_t)K * ng_ * SS * sizeof(double));
  Rinvb_.ensure((size_t)K * ng_ * SS * sizeof(double));
  BacksubArgs ba{coltab_.dev<double>(), K, S_, ng_, Y_.as<double>(), dg, Zb_.as<double>(), Rinvb_.as<double>()};
  launch_backsub(st_, ba, ng_);
  if (scale_) launch_scale_ycols(st_, Y_.as<double>(), K, ng_, S_, dg, scale_);
  const bool flex = cfg_.flexible && N_;
  const void *const *tab = flex ? ztab_.dev<const void>() : vtab_.dev<const void>();
  if (flex || !N_) {
    GemmArgs ga{tab, nullptr, 0, K, p, Y_.as<double>(), (size_t)ng_ * SS, SS,
                S_, X, p, gmax_rows_, alpha, 1.0, dg, 1};
    ga.rows_sum = rows_sum_;
    ga.vec4 = vec4_;
    launch_block_gemm<TV, double>(st_, ga, S_, ng_);
  } else {
    // P = sum_i V_i Y_i (TV), then X += alpha * N^{-1} P.
    GemmArgs ga{tab, nullptr, 0, K, p, Y_.as<double>(), (size_t)ng_ * SS, SS,
                S_, P_.ptr, p, gmax_rows_, 1.0, 0.0, dg, 1};
    ga.rows_sum = rows_sum_;
    ga.vec4 = vec4_;
    launch_block_gemm<TV, TV>(st_, ga, S_, ng_);
    TV *P = P_.as<TV>();
    prec_apply_rm<TV>(st_, N_, p, P, p, P, p);
    Xtmp_.ensure((size_t)n * p * sizeof(double));
    copy_cast<TV, double>(st_, n, p, P, p, Xtmp_.as<double>(), p, 1.0);
    if (!staged_) axpby(st_, (size_t)n * p, alpha, Xtmp_.as<double>(), 1.0, X);
  }
}

void BlockArnoldi::peek(const double *Xbase, double *Xout) {
  BPF_CHECK_HIP(hipMemcpyAsync(Xout, Xbase, (size_t)cfg_.n * cfg_.p * sizeof(double),
                               hipMemcpyDeviceToDevice, st_));
  if (cfg_.f32) form<float>(Xout, 1.0);
  else form<double>(Xout, 1.0);
}

int BlockArnoldi::run(const double *R0, double *X, int maxit, const StopFn &stop, double *rnorm_out) {
  return cfg_.f32 ? run_t<float>(R0, X, maxit, stop, rnorm_out)
                  : run_t<double>(R0, X, maxit, stop, rnorm_out);
}

void BlockArnoldi::precondition(const double *V, double *Z, int maxit) {
  BPF_REQUIRE(cfg_.nested, "precondition: not a nested run");
  const int n = cfg_.n, p = cfg_.p;
  const StopFn never = [](const double *, int) { return false; };
  // The kernel timers record events around launches: no capture then.
  if (!g_graph || gfail_ || prof::enabled()) {
    fill_double(st_, Z, 0.0, (size_t)n * p);
    run(V, Z, maxit, never, nullptr);
    return;
  }
  if (cfg_.f32) copy_cast<double, float>(st_, n, p, V, p, W_.as<float>(), p, 1.0);
  else copy_cast<double, double>(st_, n, p, V, p, W_.as<double>(), p, 1.0);
  struct Staged {
    bool &f;
    explicit Staged(bool &x) : f(x) { f = true; }
    ~Staged() { f = false; }
  } staged(staged_);
  if (gexec_ && gits_ == maxit) {
    BPF_CHECK_HIP(hipGraphLaunch(gexec_, st_));
    ++g_graph_replays;
  } else {
    if (gexec_) {
      BPF_CHECK_HIP(hipStreamSynchronize(st_));
      (void)hipGraphExecDestroy(gexec_);
      gexec_ = nullptr;
    }
    // One eager run computes this call's correction and sizes every buffer,
    // pointer table and SpMM launch shape, so the run captured next records
    // kernel launches only.
    run(nullptr, nullptr, maxit, never, nullptr);
    hipGraph_t g = nullptr;
    BPF_CHECK_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    try {
      run(nullptr, nullptr, maxit, never, nullptr);
    } catch (...) {
      e = hipErrorUnknown;
    }
    const hipError_t ee = hipStreamEndCapture(st_, &g);
    if (e == hipSuccess) e = ee;
    if (e == hipSuccess) e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      // Not capturable here: eager runs from now on (this call's correction
      // is already in Xtmp_).
      (void)hipGetLastError();
      gexec_ = nullptr;
      gfail_ = true;
    } else {
      gits_ = maxit;
    }
  }
  BPF_CHECK_HIP(hipMemcpyAsync(Z, Xtmp_.ptr, (size_t)n * p * sizeof(double), hipMemcpyDeviceToDevice, st_));
}

template <typename TV>
int BlockArnoldi::run_t(const double *R0, double *X, int maxit, const StopFn &stop,
                        double *rnorm_out) {
  const GroupDesc *dg = d_groups_.as<GroupDesc>();
  const int n = cfg_.n, p = cfg_.p;
  const size_t SS = (size_t)S_ * S_;
  const bool flex = cfg_.flexible && N_;
  const bool f32 = std::is_same<TV, float>::value;
  k_ = 0;
  shifted_ = false;
  fill_int(st_, flags_.as<int>(), 0, 1);
  // V_0 Gamma = R0 (CholQR2); E = [Gamma; 0].  (Staged: R0 is already in W_.)
  if (!staged_) copy_cast<double, TV>(st_, n, p, R0, p, W_.as<TV>(), p, 1.0);
  fill_double(st_, E_.as<double>(), 0.0, 2 * (size_t)ng_ * SS);
  vtab_.set(0, V_.block(0), st_);
  // Gamma lands in rows 0..S of each group's E = [Gamma; 0].
  cholqr2<TV>(W_.ptr, V_.block(0), E_.as<double>(), 2 * SS, -1);
  // Snapshots of the CholQR flag word: entry j + 1 after iteration j's
  // CholQR2, entry 0 after V_0's, so the run reports the breakdowns of the
  // iterations it keeps only (an overlapped run queues one it drops).  A
  // nested run reports none: its queue holds kernels only, no copies to the
  // host and no events (so that it can be captured into a graph).
  const bool nested = cfg_.nested;
  if (!nested && h_fl_cap_ < maxit + 2) {
    if (h_fl_) {
      BPF_CHECK_HIP(hipStreamSynchronize(st_));
      BPF_CHECK_HIP(hipHostFree(h_fl_));
      h_fl_ = nullptr;
    }
    BPF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_fl_), (size_t)(maxit + 2) * sizeof(int)));
    h_fl_cap_ = maxit + 2;
  }
  if (!nested) BPF_CHECK_HIP(hipMemcpyAsync(h_fl_, flags_.ptr, sizeof(int), hipMemcpyDeviceToHost, st_));
  // (the context's pinned buffer may have been regrown for another solver
  // since this one was built)
  double *h_rn = own_rn_;
  if (!cfg_.nested) {
    ctx_->ensure_stream2(std::max(1, nslots_));
    h_rn = ctx_->h_rn;
  }
  // (runs longer than kOverlapMaxIt iterations keep the serial order: the
  // pointer tables are sized for the whole run up front)
  const bool ovl = g_overlap_hess && maxit <= kOverlapMaxIt && !cfg_.nested;
  if (ovl) {
    // no table growth (and device-table replacement) while the second stream runs
    coltab_.reserve(maxit + 2, st_);
    vtab_.reserve(maxit + 2, st_);
    if (flex) ztab_.reserve(maxit + 2, st_);
  }
  // Preconditioner, SpMM, Gram-Schmidt and CholQR2 of iteration k (st_).
  auto main_work = [&](int k) {
    // P = N^{-1} V_k
    TV *Vk = static_cast<TV *>(V_.block(k));
    TV *Pk = Vk;
    if (N_) {
      Pk = flex ? static_cast<TV *>(Z_.block(k)) : P_.as<TV>();
      if (flex) ztab_.set(k, Pk, st_);
      if (inner_ && flex && !f32) {
        // Z_k from inner FP32 block GMRES iterations on A z = V_k (the
        // flexible form takes any preconditioner step)
        inner_->precondition(reinterpret_cast<const double *>(Vk), reinterpret_cast<double *>(Pk),
                             cfg_.inner_its);
      } else if (cfg_.prec32 && !f32) {
        // FP32 preconditioner on an FP64 basis: round V_k, apply, widen Z_k.
        const size_t fb = (size_t)n * p * sizeof(float);
        Vf_.ensure(fb);
        Zf_.ensure(fb);
        float *vf = Vf_.as<float>(), *zf = Zf_.as<float>();
        copy_cast<TV, float>(st_, n, p, Vk, p, vf, p, 1.0);
        if (flex && poly_ && poly_->degree() > 0) {
          poly_->apply<float>(st_, A_, N_, n, p, vf, zf);
        } else {
          prec_apply_rm<float>(st_, N_, p, vf, p, zf, p);
        }
        copy_cast<float, TV>(st_, n, p, zf, p, Pk, p, 1.0);
      } else if (flex && poly_ && poly_->degree() > 0) {
        poly_->apply<TV>(st_, A_, N_, n, p, Vk, Pk);
      } else {
        prec_apply_rm<TV>(st_, N_, p, Vk, p, Pk, p);
      }
    }
    // W = A P
    launch_spmm<TV>(st_, n, A_->nnz, A_->rowptr.as<int>(), A_->colidx.as<int>(),
                    f32 ? A_->val32.ptr : A_->val.ptr, f32, p, Pk, p, W_.as<TV>(), p, 1.0, 0.0);
    // Q = M W with a left preconditioner (src/block_gmres.jl:300)
    if (cfg_.M) prec_apply_rm<TV>(st_, cfg_.M, p, W_.as<TV>(), p, W_.as<TV>(), p);
    // Block Gram-Schmidt against V_0..V_k.
    double *col = col_region(k);
    orthogonalize<TV>(k, col);
    const size_t gstr = (size_t)(k + 2) * SS;
    // V_{k+1} C = W; C into block k+1 of each group's column.
    cholqr2<TV>(W_.ptr, V_.block(k + 1), col + (size_t)(k + 1) * SS, gstr, k);
    if (!nested)
      BPF_CHECK_HIP(hipMemcpyAsync(h_fl_ + k + 1, flags_.ptr, sizeof(int), hipMemcpyDeviceToHost, st_));
    vtab_.set(k + 1, V_.block(k + 1), st_);
  };
  // QR update of the block Hessenberg matrix for iteration k and the copy
  // of the residual estimates; ev_hess_ marks its end.
  auto hess_work = [&](int k) {
    hipStream_t hs = st_;
    if (ovl) {
      hs = hs_;
      BPF_CHECK_HIP(hipEventRecord(ev_main_, st_));   // after everything queued so far
      BPF_CHECK_HIP(hipStreamWaitEvent(hs, ev_main_, 0));
    }
    HessArgs ha{coltab_.dev<double>(), k, S_, ng_, dg};
    launch_hess_apply(hs, ha, ng_);
    HessQrArgs qa{coltab_.dev<double>(), k, S_, ng_, E_.as<double>(), dg, rnorm_.as<double>()};
    launch_hess_qr(hs, qa, ng_);
    if (nested) return;   // no host reads: the chain runs on st_ alone
    BPF_CHECK_HIP(hipMemcpyAsync(h_rn, rnorm_.ptr, (size_t)nslots_ * sizeof(double), hipMemcpyDeviceToHost, hs));
    BPF_CHECK_HIP(hipEventRecord(ev_hess_, hs));
  };
  int it = 0;
  for (;;) {
    const bool more = it < maxit;
    if (ovl) {
      // Iteration `it` is queued before iteration it-1 is tested, so the
      // update of it-1 (second stream) runs beside it; if it-1 stops the
      // solve, its work is dropped (iterations past k_ are never used).
      if (more) main_work(it);
      if (it > 0) {
        BPF_CHECK_HIP(hipEventSynchronize(ev_hess_));
        k_ = it;
        if (stop(h_rn, it)) break;
      }
      if (!more) break;
      hess_work(it);
      ++it;
    } else if (cfg_.nested) {
      // a fixed number of iterations queued without waiting on the host
      if (!more) break;
      main_work(it);
      hess_work(it);
      ++it;
      k_ = it;
    } else {
      if (!more) break;
      main_work(it);
      hess_work(it);
      BPF_CHECK_HIP(hipEventSynchronize(ev_hess_));
      ++it;
      k_ = it;
      if (stop(h_rn, it)) break;
    }
  }
  // the solution assembly below reads the last update's results
  if (!nested) BPF_CHECK_HIP(hipStreamWaitEvent(st_, ev_hess_, 0));
  if (!nested) {
    if (rnorm_out) std::memcpy(rnorm_out, h_rn, (size_t)nslots_ * sizeof(double));
    BPF_CHECK_HIP(hipStreamSynchronize(st_));
    shifted_ = h_fl_[k_] != 0;   // the flags after the last kept iteration
  }
  form<TV>(X, 1.0);
  if (N_) prec_check(N_, st_);
  if (cfg_.M) prec_check(cfg_.M, st_);
  return it;
}

template int BlockArnoldi::run_t<double>(const double *, double *, int, const StopFn &, double *);
template int BlockArnoldi::run_t<float>(const double *, double *, int, const StopFn &, double *);

// ------------------------------------------------------------- utilities
void residual_rm(hipStream_t s, bpf_csr A, int p, const double *B, const double *X, double *R) {
  BPF_CHECK_HIP(hipMemcpyAsync(R, B, (size_t)A->n * p * sizeof(double), hipMemcpyDeviceToDevice, s));
  launch_spmm<double>(s, A->n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.ptr, false, p, X, p,
                      R, p, -1.0, 1.0);
}

void backward_error_rm(bpf_ctx ctx, bpf_csr A, int p, const double *R, const double *X,
                       const double *B, double *h_berr, DevBuf &stat, bool b_norms_known) {
  BPF_REQUIRE(!b_norms_known || stat.bytes >= 3 * (size_t)p * sizeof(double), "B's norms are not in stat");
  stat.ensure(3 * (size_t)p * sizeof(double));
  double *d = stat.as<double>();
  colstat_rowmajor<double>(ctx->stream, A->n, p, R, p, 0, d);
  colstat_rowmajor<double>(ctx->stream, A->n, p, X, p, 0, d + p);
  if (!b_norms_known) colstat_rowmajor<double>(ctx->stream, A->n, p, B, p, 0, d + 2 * p);
  std::vector<double> h(3 * (size_t)p);
  BPF_CHECK_HIP(hipMemcpyAsync(h.data(), d, 3 * (size_t)p * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < p; ++c) {
    const double den = A->norm_inf * h[p + c] + h[2 * p + c];
    h_berr[c] = den > 0.0 ? h[c] / den : (h[c] > 0.0 ? INFINITY : 0.0);
  }
}

}  // namespace bpf
