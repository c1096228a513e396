// C ABI: context, device memory, CSR matrices, SpMM, ILU(0)/IC(0).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <numeric>

#include "bpf_internal.h"
#include "dense.h"

namespace bpf {

void *BufPool::take(size_t nbytes, size_t *got) {
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = free.lower_bound(nbytes);
    if (it != free.end() && it->first <= 2 * nbytes) {
      void *p = it->second;
      *got = it->first;
      cached -= it->first;
      free.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    // out of memory with blocks cached: release them and retry once (the
    // failed call's error is cleared, so that a later hipGetLastError()
    // after a kernel launch does not report it)
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, nbytes);
  }
  if (e != hipSuccess) BPF_THROW(BPF_E_NOMEM, "hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
  *got = nbytes;
  return p;
}

void BufPool::give(void *p, size_t nbytes) {
  std::lock_guard<std::mutex> lk(mu);
  if (cap == 0) {
    size_t fr = 0, tot = 0;
    cap = hipMemGetInfo(&fr, &tot) == hipSuccess && tot > 0 ? tot / 4 : ((size_t)8 << 30);
  }
  free.emplace(nbytes, p);
  cached += nbytes;
  while (cached > cap && !free.empty()) {
    auto it = std::prev(free.end());
    (void)hipFree(it->second);
    cached -= it->first;
    free.erase(it);
  }
}

void BufPool::trim() {
  std::lock_guard<std::mutex> lk(mu);
  for (auto &kv : free) (void)hipFree(kv.second);
  free.clear();
  cached = 0;
}

void DevBuf::release() {
  if (ptr) {
    if (pool) pool->give(ptr, bytes);
    else (void)hipFree(ptr);
  }
  ptr = nullptr;
  bytes = 0;
}

void DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && ptr) return;
  release();
  if (nbytes == 0) nbytes = 16;
  if (pool) {
    size_t got = 0;
    ptr = pool->take(nbytes, &got);
    bytes = got;
    return;
  }
  hipError_t e = hipMalloc(&ptr, nbytes);
  if (e != hipSuccess) {
    ptr = nullptr;
    BPF_THROW(BPF_E_NOMEM, "hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
  }
  bytes = nbytes;
}

void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) BPF_THROW(BPF_E_HIP, "%s failed: %s", what, hipGetErrorString(e));
}

void csr_norm_and_cast(hipStream_t s, int n, const int *rowptr, const double *val, float *val32,
                       int nnz, double *d_out);

}  // namespace bpf

using namespace bpf;

void bpf_ctx_s::ensure_small(size_t ndouble) {
  if (ndouble <= h_small_cap) return;
  if (h_small) (void)hipHostFree(h_small);
  h_small = nullptr;
  BPF_CHECK_HIP(hipHostMalloc((void **)&h_small, ndouble * sizeof(double), hipHostMallocDefault));
  h_small_cap = ndouble;
}

void bpf_ctx_s::ensure_stream2(size_t ndouble) {
  if (!stream2) BPF_CHECK_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  if (!ev_a) BPF_CHECK_HIP(hipEventCreateWithFlags(&ev_a, hipEventDisableTiming));
  if (!ev_b) BPF_CHECK_HIP(hipEventCreateWithFlags(&ev_b, hipEventDisableTiming));
  if (ndouble > h_rn_cap) {
    if (h_rn) (void)hipHostFree(h_rn);
    h_rn = nullptr;
    BPF_CHECK_HIP(hipHostMalloc((void **)&h_rn, ndouble * sizeof(double), hipHostMallocDefault));
    h_rn_cap = ndouble;
  }
}

void bpf_csr_s::refresh_derived() {
  ctx->scratch.ensure(64);
  csr_norm_and_cast(ctx->stream, n, rowptr.as<int>(), val.as<double>(), val32.as<float>(), nnz,
                    ctx->scratch.as<double>());
  ctx->ensure_small(1);
  BPF_CHECK_HIP(hipMemcpyAsync(ctx->h_small, ctx->scratch.ptr, sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
  BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  norm_inf = ctx->h_small[0];
  val32_valid = true;
  ++version;
}

extern "C" {

const char *bpf_last_error(void) { return bpf::last_error(); }
const char *bpf_version(void) { return "bpf_amd 0.1 (gfx950)"; }

int bpf_set_option(int option, long long value) {
  return guard([&] {
    if (option == BPF_OPT_SYNCFREE) g_syncfree = value != 0;
    else if (option == BPF_OPT_AUTOTUNE) g_autotune = value != 0;
    else if (option == BPF_OPT_OVERLAP) g_overlap_hess = value != 0;
    else if (option == BPF_OPT_GRAPH) g_graph = value != 0;
    else if (option == BPF_OPT_RELABEL_MIN) {
      BPF_REQUIRE(value >= 0, "relabel_min must be non-negative");
      g_relabel_min = (int)std::min<long long>(value, 1 << 30);
    }
    else if (option == BPF_OPT_GRAPH_REPLAYS) {
      BPF_REQUIRE(value == 0, "graph_replays can only be reset to 0");
      g_graph_replays = 0;
    }
    else if (option == BPF_OPT_WIDE_MIN) {
      BPF_REQUIRE(value >= 1, "wide_min must be at least 1");
      g_wide_min = (int)std::min<long long>(value, 1 << 30);
    }
    else if (option == BPF_OPT_DENSE_INVERSE) {
      BPF_REQUIRE(value >= 0 && value <= 2, "dense_inverse must be 0, 1 or 2");
      g_dense_inverse = (int)value;
    }
    else if (option == BPF_OPT_DENSE_BLOCKS) {
      BPF_REQUIRE(value >= 0 && value <= 256, "dense_blocks must be in 0..256");
      g_dense_blocks = (int)value;
    }
    else if (option == BPF_OPT_TAIL_ROUNDS) {
      BPF_REQUIRE(value >= 0 && value <= 64, "tail_rounds must be in 0..64");
      g_tail_rounds = (int)value;
    }
    else if (option == BPF_OPT_TAIL_SLAB) {
      BPF_REQUIRE(value == 0 || value == 64, "tail_slab must be 0 or 64");
      g_tail_slab = (int)value;
    }
    else if (option == BPF_OPT_BATCH_THREADS) {
      BPF_REQUIRE(value >= 1 && value <= 64, "batch threads must be in 1..64");
      g_batch_threads = (int)value;
    } else BPF_THROW(BPF_E_INVALID, "unknown option %d", option);
  });
}

int bpf_get_option(int option, long long *value) {
  return guard([&] {
    BPF_REQUIRE(value, "null argument");
    if (option == BPF_OPT_SYNCFREE) *value = g_syncfree ? 1 : 0;
    else if (option == BPF_OPT_AUTOTUNE) *value = g_autotune ? 1 : 0;
    else if (option == BPF_OPT_OVERLAP) *value = g_overlap_hess ? 1 : 0;
    else if (option == BPF_OPT_BATCH_THREADS) *value = g_batch_threads;
    else if (option == BPF_OPT_WIDE_MIN) *value = g_wide_min;
    else if (option == BPF_OPT_GRAPH) *value = g_graph ? 1 : 0;
    else if (option == BPF_OPT_RELABEL_MIN) *value = g_relabel_min;
    else if (option == BPF_OPT_TAIL_SLAB) *value = g_tail_slab;
    else if (option == BPF_OPT_DENSE_INVERSE) *value = g_dense_inverse;
    else if (option == BPF_OPT_DENSE_BLOCKS) *value = g_dense_blocks;
    else if (option == BPF_OPT_TAIL_ROUNDS) *value = g_tail_rounds;
    else if (option == BPF_OPT_GRAPH_REPLAYS) *value = g_graph_replays;
    else BPF_THROW(BPF_E_INVALID, "unknown option %d", option);
  });
}

int bpf_device_count(int *count) {
  return guard([&] {
    BPF_REQUIRE(count, "null count");
    int c = 0;
    BPF_CHECK_HIP(hipGetDeviceCount(&c));
    *count = c;
  });
}

int bpf_ctx_create(int device, bpf_ctx *out) {
  return guard([&] {
    BPF_REQUIRE(out, "null output");
    int c = 0;
    BPF_CHECK_HIP(hipGetDeviceCount(&c));
    BPF_REQUIRE(device >= 0 && device < c, "device %d not present (%d devices)", device, c);
    BPF_CHECK_HIP(hipSetDevice(device));
    noalloc_setup(device);   // per-device address of the no-allocation scratch
    auto *ctx = new bpf_ctx_s();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      BPF_THROW(BPF_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    BPF_CHECK_HIP(hipEventCreate(&ctx->ev_start));
    BPF_CHECK_HIP(hipEventCreate(&ctx->ev_stop));
    *out = ctx;
  });
}

int bpf_ctx_destroy(bpf_ctx ctx) {
  return guard([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    blas_release(ctx);
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    if (ctx->h_rn) (void)hipHostFree(ctx->h_rn);
    if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
    if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int bpf_ctx_synchronize(bpf_ctx ctx) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int bpf_malloc(bpf_ctx ctx, void **d_ptr, size_t bytes) {
  return guard([&] {
    BPF_REQUIRE(ctx && d_ptr, "null argument");
    BPF_CHECK_HIP(hipSetDevice(ctx->device));
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) BPF_THROW(BPF_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    *d_ptr = p;
  });
}

int bpf_free(bpf_ctx ctx, void *d_ptr) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    if (d_ptr) {
      BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      BPF_CHECK_HIP(hipFree(d_ptr));
    }
  });
}

int bpf_memcpy_h2d(bpf_ctx ctx, void *d_dst, const void *h_src, size_t bytes) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    if (!bytes) return;
    BPF_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int bpf_memcpy_d2h(bpf_ctx ctx, void *h_dst, const void *d_src, size_t bytes) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    if (!bytes) return;
    BPF_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int bpf_memcpy_d2d(bpf_ctx ctx, void *d_dst, const void *d_src, size_t bytes) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    if (!bytes) return;
    BPF_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}

int bpf_timer_start(bpf_ctx ctx) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    BPF_CHECK_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
  });
}

int bpf_timer_stop(bpf_ctx ctx, double *ms) {
  return guard([&] {
    BPF_REQUIRE(ctx && ms, "null argument");
    BPF_CHECK_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    BPF_CHECK_HIP(hipEventSynchronize(ctx->ev_stop));
    float f = 0.f;
    BPF_CHECK_HIP(hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop));
    *ms = f;
  });
}

// ------------------------------------------------------------------ CSR
int bpf_csr_create(bpf_ctx ctx, int n, int nnz, const int32_t *h_rowptr, const int32_t *h_colidx,
                   const double *h_val, bpf_csr *out) {
  return guard([&] {
    BPF_REQUIRE(ctx && out, "null argument");
    BPF_REQUIRE(n >= 0 && nnz >= 0, "negative size");
    BPF_REQUIRE(h_rowptr && (nnz == 0 || (h_colidx && h_val)), "null CSR array");
    BPF_REQUIRE(h_rowptr[0] == 0 && h_rowptr[n] == nnz, "rowptr must start at 0 and end at nnz");
    for (int i = 0; i < n; ++i)
      BPF_REQUIRE(h_rowptr[i + 1] >= h_rowptr[i], "rowptr not monotone at row %d", i);
    for (int k = 0; k < nnz; ++k)
      BPF_REQUIRE(h_colidx[k] >= 0 && h_colidx[k] < n, "column index %d out of range at %d", h_colidx[k], k);
    BPF_CHECK_HIP(hipSetDevice(ctx->device));
    auto *A = new bpf_csr_s();
    A->ctx = ctx;
    A->n = n;
    A->nnz = nnz;
    A->h_rowptr.assign(h_rowptr, h_rowptr + n + 1);
    // Sort columns within rows (values follow).
    std::vector<int> col(h_colidx, h_colidx + nnz);
    std::vector<double> val(h_val, h_val + nnz);
    bool sorted = true;
    for (int i = 0; i < n && sorted; ++i)
      for (int k = h_rowptr[i] + 1; k < h_rowptr[i + 1]; ++k)
        if (col[k] <= col[k - 1]) { sorted = false; break; }
    if (!sorted) {
      std::vector<int> idx;
      for (int i = 0; i < n; ++i) {
        const int b = h_rowptr[i], e = h_rowptr[i + 1];
        idx.resize(e - b);
        std::iota(idx.begin(), idx.end(), b);
        std::sort(idx.begin(), idx.end(), [&](int x, int y) { return col[x] < col[y]; });
        std::vector<int> c2(e - b);
        std::vector<double> v2(e - b);
        for (int q = 0; q < e - b; ++q) { c2[q] = col[idx[q]]; v2[q] = val[idx[q]]; }
        for (int q = 1; q < e - b; ++q)
          if (c2[q] == c2[q - 1]) { delete A; BPF_THROW(BPF_E_INVALID, "duplicate entry in row %d", i); }
        std::copy(c2.begin(), c2.end(), col.begin() + b);
        std::copy(v2.begin(), v2.end(), val.begin() + b);
      }
    }
    A->h_colidx = col;
    try {
      upload(A->rowptr, A->h_rowptr, ctx->stream);
      upload(A->colidx, A->h_colidx, ctx->stream);
      upload(A->val, val, ctx->stream);
      A->val32.ensure(std::max(1, nnz) * sizeof(float));
      A->refresh_derived();
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}

int bpf_csr_set_values(bpf_csr A, const double *h_val) {
  return guard([&] {
    BPF_REQUIRE(A && h_val, "null argument");
    BPF_CHECK_HIP(hipMemcpyAsync(A->val.ptr, h_val, (size_t)A->nnz * sizeof(double),
                                 hipMemcpyHostToDevice, A->ctx->stream));
    A->refresh_derived();
  });
}

int bpf_csr_set_values_device(bpf_csr A, const double *d_val) {
  return guard([&] {
    BPF_REQUIRE(A && d_val, "null argument");
    BPF_CHECK_HIP(hipMemcpyAsync(A->val.ptr, d_val, (size_t)A->nnz * sizeof(double),
                                 hipMemcpyDeviceToDevice, A->ctx->stream));
    A->refresh_derived();
  });
}

int bpf_csr_info(bpf_csr A, int *n, int *nnz, double *norm_inf) {
  return guard([&] {
    BPF_REQUIRE(A, "null matrix");
    if (n) *n = A->n;
    if (nnz) *nnz = A->nnz;
    if (norm_inf) *norm_inf = A->norm_inf;
  });
}

int bpf_csr_destroy(bpf_csr A) {
  return guard([&] {
    if (!A) return;
    (void)hipStreamSynchronize(A->ctx->stream);
    spmm_forget(A->rowptr.as<int>());
    relabel_free(A);
    delete A;
  });
}

int bpf_spmm(bpf_ctx ctx, bpf_csr A, int dtype, int layout, int p, const void *d_X, int ldx,
             void *d_Y, int ldy, double alpha, double beta) {
  return guard([&] {
    BPF_REQUIRE(ctx && A, "null handle");
    BPF_REQUIRE(dtype == BPF_F64 || dtype == BPF_F32, "bad dtype");
    BPF_REQUIRE(p >= 0, "negative p");
    if (p == 0 || A->n == 0) return;
    BPF_REQUIRE(d_X && d_Y, "null block");
    const bool f32 = dtype == BPF_F32;
    const void *val = f32 ? A->val32.ptr : A->val.ptr;
    if (layout == BPF_COL_MAJOR) {
      BPF_REQUIRE(ldx >= A->n && ldy >= A->n, "leading dimension < n");
      launch_spmm_colmajor(ctx->stream, A->n, A->rowptr.as<int>(), A->colidx.as<int>(), val, f32, p,
                           d_X, ldx, d_Y, ldy, alpha, beta, f32);
    } else if (layout == BPF_ROW_MAJOR) {
      BPF_REQUIRE(ldx >= p && ldy >= p, "leading dimension < p");
      if (f32)
        launch_spmm<float>(ctx->stream, A->n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), val, true, p,
                           static_cast<const float *>(d_X), ldx, static_cast<float *>(d_Y), ldy,
                           alpha, beta);
      else
        launch_spmm<double>(ctx->stream, A->n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), val, false, p,
                            static_cast<const double *>(d_X), ldx, static_cast<double *>(d_Y), ldy,
                            alpha, beta);
    } else {
      BPF_THROW(BPF_E_INVALID, "bad layout");
    }
  });
}

// ------------------------------------------------------- preconditioners
int bpf_prec_factor(bpf_prec P, bpf_csr A) {
  return guard([&] {
    BPF_REQUIRE(P && A, "null handle");
    BPF_REQUIRE(A->n == P->n && A->nnz == P->nnz && A->h_colidx == P->h_colidx &&
                    A->h_rowptr == P->h_rowptr,
                "matrix pattern differs from the analysed pattern");
    if (P->kind == BPF_PREC_BLOCK_JACOBI) {
      bj_factor(P, A);   // LS.update(BJ, A, device): new block inverses
      return;
    }
    hipStream_t s = P->ctx->stream;
    BPF_CHECK_HIP(hipMemcpyAsync(P->fac.ptr, A->val.ptr, (size_t)P->nnz * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
    P->ctx->scratch.ensure(64);
    int *d_status = P->ctx->scratch.as<int>();
    const int init = P->n;
    BPF_CHECK_HIP(hipMemcpyAsync(d_status, &init, sizeof(int), hipMemcpyHostToDevice, s));
    if (P->kind == BPF_PREC_ILU0) {
      launch_ilu_factor(s, P->lo, P->n, P->rowptr.as<int>(), P->colidx.as<int>(), P->diagpos.as<int>(),
                        P->plan, P->fac.as<double>(), P->max_row_nnz, d_status,
                        factor_bytes(P->n, P->nnz));
    } else {
      // IC(0) eliminates the strictly lower part and the diagonal of each row
      double entries = P->n;
      for (int i = 0; i < P->n; ++i) entries += P->h_diagpos[i] - P->h_rowptr[i];
      launch_ic_factor(s, P->lo, P->n, P->rowptr.as<int>(), P->colidx.as<int>(), P->diagpos.as<int>(),
                       P->fac.as<double>(), P->max_row_nnz, d_status, factor_bytes(P->n, entries));
    }
    int status = 0;
    BPF_CHECK_HIP(hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    triop_check(P->lo, s);
    triop_refresh(P->lo, P->fac.as<double>(), s);
    triop_refresh(P->up, P->fac.as<double>(), s);
    ++P->fac_version;
    if (status < P->n)
      BPF_THROW(BPF_E_SINGULAR, "%s: zero or non-positive pivot at row %d",
                P->kind == BPF_PREC_ILU0 ? "ilu0" : "ic0", status);
  });
}

int bpf_prec_create(bpf_ctx ctx, bpf_csr A, int kind, bpf_prec *out) {
  return guard([&] {
    BPF_REQUIRE(ctx && A && out, "null argument");
    BPF_REQUIRE(kind == BPF_PREC_ILU0 || kind == BPF_PREC_IC0, "bad preconditioner kind");
    auto *P = new bpf_prec_s();
    try {
      P->ctx = ctx;
      P->kind = kind;
      P->n = A->n;
      P->nnz = A->nnz;
      P->h_rowptr = A->h_rowptr;
      P->h_colidx = A->h_colidx;
      P->h_diagpos.assign(A->n, -1);
      for (int i = 0; i < A->n; ++i) {
        for (int k = A->h_rowptr[i]; k < A->h_rowptr[i + 1]; ++k)
          if (A->h_colidx[k] == i) P->h_diagpos[i] = k;
        if (P->h_diagpos[i] < 0)
          BPF_THROW(BPF_E_SINGULAR, "structurally zero diagonal in row %d", i);
        P->max_row_nnz = std::max(P->max_row_nnz, A->h_rowptr[i + 1] - A->h_rowptr[i]);
      }
      upload(P->rowptr, P->h_rowptr, ctx->stream);
      upload(P->colidx, P->h_colidx, ctx->stream);
      upload(P->diagpos, P->h_diagpos, ctx->stream);
      P->fac.ensure(std::max(1, P->nnz) * sizeof(double));
      if (kind == BPF_PREC_ILU0) {
        build_elim_plan(P->n, P->h_rowptr, P->h_colidx, P->h_diagpos, ctx->stream, P->plan);
        build_triop(P->lo, P->n, P->h_rowptr, P->h_colidx, 0, ctx->stream);
        build_triop(P->up, P->n, P->h_rowptr, P->h_colidx, 1, ctx->stream);
      } else {
        build_triop(P->lo, P->n, P->h_rowptr, P->h_colidx, 2, ctx->stream);
        build_triop(P->up, P->n, P->h_rowptr, P->h_colidx, 3, ctx->stream);
      }
      int rc = bpf_prec_factor(P, A);
      if (rc != BPF_OK) throw Error{rc};
    } catch (...) {
      delete P;
      throw;
    }
    *out = P;
  });
}

int bpf_prec_apply(bpf_prec P, int dtype, int layout, int p, const void *d_X, int ldx, void *d_Y,
                   int ldy) {
  return guard([&] {
    BPF_REQUIRE(P, "null preconditioner");
    BPF_REQUIRE(dtype == BPF_F64 || dtype == BPF_F32, "bad dtype");
    if (p <= 0 || P->n == 0) return;
    hipStream_t s = P->ctx->stream;
    const bool f32 = dtype == BPF_F32;
    if (layout == BPF_ROW_MAJOR) {
      BPF_REQUIRE(ldx >= p && ldy >= p, "leading dimension < p");
      if (f32)
        prec_apply_rm<float>(s, P, p, static_cast<const float *>(d_X), ldx, static_cast<float *>(d_Y), ldy);
      else
        prec_apply_rm<double>(s, P, p, static_cast<const double *>(d_X), ldx, static_cast<double *>(d_Y), ldy);
    } else if (layout == BPF_COL_MAJOR) {
      BPF_REQUIRE(ldx >= P->n && ldy >= P->n, "leading dimension < n");
      const int n = P->n;
      P->tmp.ensure((size_t)n * p * sizeof(double));
      double *t = P->tmp.as<double>();
      // Float32 blocks are applied in FP64: widened into a column-major copy,
      // the result rounded back.
      DevBuf w(&P->ctx->pool);
      const double *src = static_cast<const double *>(d_X);
      if (f32) {
        w.ensure((size_t)n * p * sizeof(double));
        copy_cast<float, double>(s, p, n, static_cast<const float *>(d_X), ldx, w.as<double>(), n, 1.0);
        src = w.as<double>();
      }
      launch_cm_to_rm(s, n, p, src, f32 ? n : ldx, t, p, false);
      prec_apply_rm<double>(s, P, p, t, p, t, p);
      if (f32) {
        launch_rm_to_cm(s, n, p, t, p, false, w.as<double>(), n);
        copy_cast<double, float>(s, p, n, w.as<double>(), n, static_cast<float *>(d_Y), ldy, 1.0);
      } else {
        launch_rm_to_cm(s, n, p, t, p, false, static_cast<double *>(d_Y), ldy);
      }
    } else {
      BPF_THROW(BPF_E_INVALID, "bad layout");
    }
    prec_check(P, s);
  });
}

int bpf_prec_get_factor(bpf_prec P, double *h_val) {
  return guard([&] {
    BPF_REQUIRE(P && h_val, "null argument");
    BPF_REQUIRE(P->kind != BPF_PREC_BLOCK_JACOBI, "a block-Jacobi preconditioner has no combined factor");
    BPF_CHECK_HIP(hipMemcpyAsync(h_val, P->fac.ptr, (size_t)P->nnz * sizeof(double),
                                 hipMemcpyDeviceToHost, P->ctx->stream));
    BPF_CHECK_HIP(hipStreamSynchronize(P->ctx->stream));
  });
}

int bpf_prec_levels(bpf_prec P, int *nl, int *nu) {
  return guard([&] {
    BPF_REQUIRE(P, "null preconditioner");
    if (nl) *nl = P->lo.nlevels;
    if (nu) *nu = P->up.nlevels;
  });
}

int bpf_prec_destroy(bpf_prec P) {
  return guard([&] {
    if (!P) return;
    (void)hipStreamSynchronize(P->ctx->stream);
    relabel_free(P);
    delete P;
  });
}

}  // extern "C"
