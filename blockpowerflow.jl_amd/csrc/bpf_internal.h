// Internal declarations of libbpf_amd: handle structures, error plumbing,
// device buffers, triangular-solve schedules and kernel launchers.
#pragma once

#include <atomic>
#include <mutex>
#include <map>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bpf.h"
#include "host_common.h"

namespace bpf {

#define BPF_CHECK_HIP(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      BPF_THROW(BPF_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                __FILE__, __LINE__);                                           \
  } while (0)

// Runs f() and converts exceptions into status codes.
template <typename F>
int guard(F &&f) {
  try {
    f();
    return BPF_OK;
  } catch (const Error &e) {
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return BPF_E_NOMEM;
  } catch (...) {
    set_error("unknown exception");
    return BPF_E_INVALID;
  }
}

// ---------------------------------------------------------------- buffers
// A context's cache of device allocations: blocks given back are handed out
// again for requests of up to their size (and at least half of it), so a
// solver that runs again with the same shapes makes no hipMalloc / hipFree
// (which cost ~0.1-0.3 ms each and synchronize).  Blocks are only reused by
// work of the same context, whose kernels run in stream order, so a block
// given back while kernels still use it is safe to hand out again.
struct BufPool {
  std::multimap<size_t, void *> free;
  size_t cached = 0;
  // Cache limit: a quarter of the device's memory (set on first use); a
  // block given back beyond it is freed, largest cached blocks first.
  size_t cap = 0;
  std::mutex mu;
  void *take(size_t nbytes, size_t *got);   // cached block or hipMalloc
  void give(void *p, size_t nbytes);
  void trim();                              // hipFree every cached block
  ~BufPool() { trim(); }
};

// Owning device buffer (hipMalloc, or a context's BufPool when `pool` is
// set); grows, never shrinks.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  BufPool *pool = nullptr;
  DevBuf() = default;
  explicit DevBuf(BufPool *p) : pool(p) {}
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes), pool(o.pool) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr; bytes = o.bytes; pool = o.pool;
      o.ptr = nullptr; o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release();
  void ensure(size_t nbytes);              // contents not preserved
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

void check_hip(hipError_t e, const char *what);

// Synchronous upload of a host vector into a (grown) device buffer.
template <typename T>
void upload(DevBuf &b, const std::vector<T> &h, hipStream_t s) {
  b.ensure((h.empty() ? 1 : h.size()) * sizeof(T));
  if (!h.empty())
    check_hip(hipMemcpyAsync(b.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(upload)");
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(upload)");
}

}  // namespace bpf

// ---------------------------------------------------------------- handles
namespace bpf {
struct CsrRelabel;
struct PrecRelabel;
}  // namespace bpf

struct bpf_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bpf::DevBuf scratch;        // generic reduction scratch
  bpf::DevBuf pinned_dummy;
  double *h_small = nullptr;  // pinned host staging for small D2H reads
  size_t h_small_cap = 0;
  void ensure_small(size_t ndouble);
  bpf::BufPool pool;          // solver work buffers, reused across calls
  // Second stream, its two events and pinned host buffer for the block
  // GMRES Hessenberg updates (created on first use, kept for the context).
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double *h_rn = nullptr;
  size_t h_rn_cap = 0;
  void ensure_stream2(size_t ndouble);
  void *blas = nullptr;       // rocblas_handle for the single-block wide solvers
};

struct bpf_csr_s {
  bpf_ctx ctx = nullptr;
  int n = 0, nnz = 0;
  std::vector<int> h_rowptr, h_colidx;
  bpf::DevBuf rowptr, colidx, val, val32;
  bpf::DevBuf rowlen_order;    // rows sorted for SpMM (unused: natural order)
  double norm_inf = 0.0;       // ||A||_inf of the current values
  bool val32_valid = false;
  unsigned long long version = 0;   // bumped by every value change
  bpf::CsrRelabel *rl = nullptr;    // locality-relabelled copy for the solvers (lazy)
  void refresh_derived();      // after a value change: norm, FP32 copy
};

namespace bpf {

// A triangular operator in level-scheduled form: the host level sets and
// their device copies.
struct TriOp : TriLevels {
  DevBuf order, ptr, col, gather, dgather, lvl;
  DevBuf val, val32, diag, diag32;   // refreshed from the factor values
  // Sync-free execution: per-row completion flags (original row numbering)
  // tagged with the epoch of the launch that produced them, and a status
  // word set when a bounded dependency wait ran out.
  DevBuf flags, status;
  int epoch = 0;
  int checked = 0;             // epoch of the last status check
};

// Device copy of an elimination plan (elim_plan_host).
struct ElimPlan {
  DevBuf pinfo, map;
  DevBuf rdiag;   // 1 / U[k][k] of every row k eliminated so far (written by the factorization)
  bool ready() const { return pinfo.ptr && map.ptr; }
};
// Process-wide switch between the sync-free single-launch triangular
// solves / factorizations and the level-by-level launches (default).
extern std::atomic<bool> g_syncfree;
// Process-wide switch: block-Hessenberg updates on a second stream, beside
// the next iteration (default on).
extern std::atomic<bool> g_overlap_hess;
// Process-wide switch: the inner runs of the mixed form's preconditioner step
// are captured into a HIP graph and replayed (default on), and the number of
// replays so far.
extern std::atomic<bool> g_graph;
extern std::atomic<long long> g_graph_replays;
// Process-wide column width (64, or 0 = one workgroup) of the slabs the
// narrow-level part of a triangular solve is split into, one workgroup each.
extern std::atomic<int> g_tail_slab;
// Process-wide: solves of a refactorization handle's dense tail by explicit
// inverse factors (0 never, 1 FP32 solves, 2 all solves).
extern std::atomic<int> g_dense_inverse;
// Process-wide: most 128-row blocks of a refactorization handle's dense tail.
extern std::atomic<int> g_dense_blocks;
// Process-wide: a triangular-solve level joins the narrow-level launch when
// its rows fit in this many rounds of that launch's wavefronts.
extern std::atomic<int> g_tail_rounds;
// Process-wide number of host threads a batched solve runs its bins on.
extern std::atomic<int> g_batch_threads;
// Process-wide: block GMRES (FP64) with at least this many right-hand sides
// in one Krylov block runs the single-block form on rocBLAS / rocSOLVER
// (default 129: narrower blocks run on the library's own kernels).
extern std::atomic<int> g_wide_min;
// Releases the context's rocBLAS handle (created on first use).
void blas_release(bpf_ctx ctx);
// Process-wide switch for the SpMM launch-shape autotuner (default on), and
// removal of a matrix's cached shapes (its row-pointer array is released).
extern std::atomic<bool> g_autotune;
void spmm_forget(const int *rowptr);
// Reads the op's status word (blocking); throws BPF_E_HIP if a sync-free
// dependency wait ran out, after switching to level-by-level launches.
void triop_check(TriOp &op, hipStream_t s);

// Launch plan: consecutive levels run either one launch per level ("wide")
// or many levels in one single-workgroup launch ("narrow").  A level is wide
// when it has more than `tail_rows` rows.
struct Seg { int l0, l1; bool wide; };
std::vector<Seg> make_plan(const TriOp &op, int tail_rows);

// Builds a TriOp from a CSR pattern (analyze_triop) and uploads it.
void build_triop(TriOp &op, int n, const std::vector<int> &rowptr,
                 const std::vector<int> &colidx, int kind, hipStream_t s,
                 const TriWindow &win = TriWindow());
// Copies values (FP64 and FP32) from a combined factor array on the device.
void triop_refresh(TriOp &op, const double *d_fac, hipStream_t s);

}  // namespace bpf

struct bpf_prec_s {
  bpf_ctx ctx = nullptr;
  int kind = BPF_PREC_ILU0;
  int n = 0, nnz = 0;
  bpf::DevBuf rowptr, colidx, diagpos, fac;   // combined factor on A's pattern
  bpf::ElimPlan plan;                         // ILU(0): elimination plan of the pattern
  unsigned long long fac_version = 0;         // bumped by every factorization
  bpf::PrecRelabel *rl = nullptr;             // relabelled triangular operators (lazy)
  std::vector<int> h_rowptr, h_colidx, h_diagpos;
  bpf::TriOp lo, up;                          // solve operators
  bpf::DevBuf tmp;                            // n x p scratch for apply
  int max_row_nnz = 0;
  // Block Jacobi (kind BPF_PREC_BLOCK_JACOBI): rows grouped into nblocks
  // parts; part b holds rows bj_rows[bj_ptr[b] .. bj_ptr[b+1]) and its dense
  // inverse (row-major, leading dimension bj_mmax) at b * bj_mmax^2 of
  // bj_inv (FP64) and bj_inv32 (FP32 copy).  bj_part[i] = part of row i,
  // bj_loc[i] = position of row i within its part.
  int bj_nblocks = 0, bj_mmax = 0;
  std::vector<int> h_bj_part;
  bpf::DevBuf bj_rows, bj_ptr, bj_part, bj_loc, bj_inv, bj_inv32;
};

namespace bpf {
// y = P^{-1} x for a row-major n x p block (y may alias x), whatever the
// kind of the handle; prec_check reports a failed sync-free wait.
template <typename TV>
void prec_apply_rm(hipStream_t s, bpf_prec P, int p, const TV *x, int ldx, TV *y, int ldy);
void prec_check(bpf_prec P, hipStream_t s);
// Device address of the no-allocation entry point's scratch array on
// `device` (the current device), resolved once per device.
void noalloc_setup(int device);
// Block Jacobi: dense inverses of the diagonal blocks from A's current
// values (throws BPF_E_SINGULAR on a singular block), and their application.
void bj_factor(bpf_prec P, bpf_csr A);
template <typename TV>
void bj_apply(hipStream_t s, bpf_prec P, int p, const TV *x, int ldx, TV *y, int ldy);
}  // namespace bpf

struct bpf_rf_s;
struct bpf_batch_s;

namespace bpf {

// ---------------------------------------------------------------- launchers
// Row-major blocks: element (i, c) at X[i * ld + c].
template <typename TV>
void launch_spmm(hipStream_t s, int n, int nnz, const int *rowptr, const int *colidx,
                 const void *val, bool val32, int p, const TV *X, int ldx, TV *Y, int ldy,
                 double alpha, double beta);
void launch_spmm_colmajor(hipStream_t s, int n, const int *rowptr, const int *colidx,
                          const void *val, bool val32, int p, const void *X, int ldx,
                          void *Y, int ldy, double alpha, double beta, bool f32);

// y = op^{-1} x for p columns (row-major).  x may alias y.
template <typename TV>
void trisolve(hipStream_t s, TriOp &op, int p, const TV *x, int ldx, TV *y, int ldy);

// Numeric IKJ factorization on a combined pattern (ILU0 when the pattern is
// A's, exact LU when it is the full fill pattern).  The values must already
// hold the (permuted) matrix entries.  Uses the lower TriOp's levels.
// `bytes` is the launch's compulsory HBM traffic for the kernel timers: each
// eliminated row's values read and written once, its column indices, and its
// row pointer / diagonal position / order entries (factor_bytes()).
// Rows [r0, r1), independent of each other, eliminated against the (final)
// pivot rows k < kmax only.
void build_elim_plan(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
                     const std::vector<int> &diagpos, hipStream_t s, ElimPlan &out);
void launch_partial_factor(hipStream_t s, int r0, int r1, int kmax, const int *rowptr, const int *colidx,
                           const int *diagpos, const ElimPlan &plan, double *fac, int max_row_nnz,
                           int *d_status, double bytes);
void launch_ilu_factor(hipStream_t s, TriOp &lower, int n, const int *rowptr,
                       const int *colidx, const int *diagpos, const ElimPlan &plan, double *fac,
                       int max_row_nnz, int *d_status, double bytes);
void launch_ic_factor(hipStream_t s, TriOp &lower, int n, const int *rowptr,
                      const int *colidx, const int *diagpos, double *fac, int max_row_nnz,
                      int *d_status, double bytes);
// Compulsory bytes of eliminating `rows` rows holding `entries` stored values.
inline double factor_bytes(double rows, double entries) {
  return entries * (2 * sizeof(double) + sizeof(int)) + rows * 3 * sizeof(int);
}

// Row-permuted layout conversions: dst (row-major) row k = src (column-major)
// row rowmap[k]; dst (column-major) row rowmap[k] = src (row-major) row k.
template <typename TD>
void launch_perm_cm_to_rm(hipStream_t s, int n, int p, const double *src, int lds, const int *rowmap, TD *dst,
                          int ldd);
template <typename TS>
void launch_perm_rm_to_cm(hipStream_t s, int n, int p, const TS *src, int lds, const int *rowmap, double *dst,
                          int ldd);

// Locality relabelling for the Krylov solvers (csrc/relabel.cpp).  Systems
// with at least g_relabel_min unknowns are solved in a reverse Cuthill-McKee
// relabelling of A's (symmetrized) pattern, so that the rows an SpMM or a
// triangular-solve row gathers lie near it in memory: A's relabelled copy
// (each row's entries in the original order, so each product is summed as
// before) and relabelled views of an ILU(0)/IC(0) handle's triangular
// operators (the same levels and rows).  Both are built once per pattern and
// refreshed when A's values or the factor change.
extern std::atomic<int> g_relabel_min;
struct CsrRelabel {
  std::vector<int> perm;       // original row i -> storage row perm[i]
  DevBuf order;                // storage row k holds original row order[k] (device)
  DevBuf emap;                 // original entry e -> entry of the relabelled copy
  bpf_csr_s *Ar = nullptr;
  unsigned long long version = ~0ull;
};
struct PrecRelabel {
  const CsrRelabel *key = nullptr;
  bpf_prec_s *Nr = nullptr;
  unsigned long long fac_version = ~0ull;
};
// The relabelled copy of A with current values (built on first use).
CsrRelabel *relabel_csr(bpf_csr A, hipStream_t s);
// The relabelled view of P for A's relabelling (null for kinds it does not
// cover), its values current.
bpf_prec relabel_prec(bpf_prec P, const CsrRelabel *rl, hipStream_t s);
void relabel_free(bpf_csr A);
void relabel_free(bpf_prec P);

// Layout conversions (column-major <-> row-major) with type conversion.
void launch_cm_to_rm(hipStream_t s, int n, int p, const double *src, int lds, void *dst,
                     int ldd, bool dst32);
void launch_rm_to_cm(hipStream_t s, int n, int p, const void *src, int lds, bool src32,
                     double *dst, int ldd);

}  // namespace bpf
