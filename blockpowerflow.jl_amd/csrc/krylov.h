// Block-Krylov building blocks shared by the single-system and batched
// solvers: a growable block arena, the block-Arnoldi engine (grouped block
// GMRES with CholQR2 block orthonormalization and an incremental Householder
// QR of the block Hessenberg matrix), and small device utilities.
#pragma once

#include <functional>
#include <memory>

#include "dense.h"
#include "poly.h"

namespace bpf {

// n x p blocks of element type T, allocated in chunks; a device table holds
// the block pointers for kernels that walk several blocks.
struct BlockArena {
  size_t block_elems = 0;
  size_t esize = 0;
  int chunk = 8;
  BufPool *pool = nullptr;     // chunks come from / go back to this pool
  std::vector<DevBuf> chunks;
  std::vector<void *> ptrs;
  void init(size_t elems, size_t es, int chunk_blocks = 8) {
    block_elems = elems;
    esize = es;
    chunk = chunk_blocks;
  }
  void *block(int i);   // grows as needed
  int capacity() const { return (int)ptrs.size(); }
};

// Device table of pointers with a pinned host mirror; entries are appended
// with small asynchronous copies (an entry never changes once written).
struct PtrTable {
  void **h = nullptr;
  DevBuf d;
  int cap = 0;
  PtrTable() = default;
  PtrTable(const PtrTable &) = delete;
  PtrTable &operator=(const PtrTable &) = delete;
  ~PtrTable();
  void set(int i, void *ptr, hipStream_t s);
  // Capacity for entries 0..n-1, so that later set() calls never replace the
  // device table (work on another stream may be reading it).
  void reserve(int n, hipStream_t s);
  template <typename T> T *const *dev() const { return d.as<T *>(); }
};

// Buffers of a grouped CholQR2 (per group S x S: Gram matrix, up to three R
// factors and their inverses; the flag words; projection partials).
struct CholQrWork {
  DevBuf G, R1, R1i, R2, R2i, R3, R3i, flags, part;
  DevBuf gflags, gsets;        // per-group bits; the groups split by them (2 ng)
  void init(int ng, int S);
};

// Grouped CholQR2 of W (row-major, ld p): Vout = W R1^{-1} R2^{-1} with
// C = R2 R1 (upper, per group at Cout + g*cstride).  W is overwritten by the
// first pass.  Columns whose first-pass Cholesky pivot is not positive are
// deflated (zero rows of R1) and replaced by pseudo-random directions, and a
// positive pivot at rounding level is flagged; in the groups where either
// happened `reorth` (if set) projects the block off the earlier basis and two
// more passes follow, C = R3 R2 R1.  The host reads one flag word to see
// whether any group broke down; which groups take the extra passes is
// decided on the device, group by group, so no group's arithmetic depends on
// another's.
// n_shift is the row count used in the shift of a failed pivot; w.flags[0]
// gathers bit 0 (shift), bit 1 (deflation) and bit 2 (rounding-level pivot)
// and is cleared by the caller.  reorth(W, groups) acts on the groups given.
// With host_check false the host does not wait for the flag word: every
// group takes the plain second pass (deflated columns keep their
// pseudo-random completion without re-projection).
using ReorthFn = std::function<void(void *, const GroupDesc *)>;
template <typename TV>
void cholqr2_groups(hipStream_t s, const GroupDesc *dg, int ng, int S, int p, int gmax_rows, int rows_sum,
                    int n_shift, void *W, void *Vout, double *Cout, size_t cstride, CholQrWork &w,
                    const ReorthFn &reorth = {}, bool host_check = true, int passes = 2, int vec4 = 0);

// Host-visible description of one Krylov run.
struct ArnoldiConfig {
  int n = 0, p = 0;            // rows, columns of the RHS block
  int s = 0;                   // group width
  bool f32 = false;            // basis / operator precision
  bool flexible = false;       // store Z_k = N^{-1} V_k
  int orth = 0;                // ORTH_BMGS / ORTH_BCGS / ORTH_BCGS2
  int bmgs_blocks = 1;         // BMGS: basis blocks projected together per step
  bool prec32 = false;         // apply the preconditioner in FP32 (FP64 basis)
  // Left preconditioner (null = identity): W = M A N^{-1} V_k, and the run
  // minimises ||M (B - A X)||_F (src/block_gmres.jl:226,300).
  bpf_prec M = nullptr;
  // Explicit groups (row ranges, widths, slots) for many systems stacked in
  // one block-diagonal operator; empty = column slices of width s.
  std::vector<GroupDesc> groups;
  // Flexible runs with an FP64 basis: the preconditioner step is inner_its
  // iterations of an FP32 block GMRES (ILU(0)-right-preconditioned, groups
  // of inner_s columns, 0 = the run's own groups) on A z = V_k, from z = 0
  // (0 = off).  With explicit groups each group's columns are split.
  int inner_its = 0;
  int inner_s = 0;
  // A run nested inside another one's preconditioner step: one stream, its
  // own event and pinned buffer, no second-stream overlap.
  bool nested = false;
  // CholQR passes per orthonormalization: 2 (CholQR2) or 1 (nested runs,
  // where the block only has to be near-orthonormal).
  int cholqr_passes = 2;
};
enum { ORTH_BMGS = 0, ORTH_BCGS = 1, ORTH_BCGS2 = 2 };

// Stopping decision callback: given per-column residual 2-norm estimates of
// the (scaled) problem and the iteration count, return true to stop.
using StopFn = std::function<bool(const double *rnorm, int iter)>;

class BlockArnoldi {
 public:
  BlockArnoldi(bpf_ctx ctx, bpf_csr A, bpf_prec N, const ArnoldiConfig &cfg);
  ~BlockArnoldi();
  // Runs up to maxit block iterations for A D = R0 (R0: FP64 row-major n x p,
  // columns pre-scaled by the caller), then adds alpha * D to X (FP64
  // row-major).  Returns the number of iterations; rnorm_out (host, p) gets
  // the final residual estimates.
  int run(const double *R0, double *X, int maxit, const StopFn &stop, double *rnorm_out);
  // Preconditioner step of an outer flexible run, for a nested run: Z = the
  // correction of maxit iterations on A z = V from z = 0 (V, Z: FP64
  // row-major n x p).  With the graph option on, the fixed-length iteration
  // chain is captured into a HIP graph on the first call (after one eager
  // run that sizes every buffer and table) and replayed on later calls with
  // the same maxit: the input is cast into the run's own block before the
  // replay and the correction copied out after it.  Results are identical to
  // the eager run's.
  void precondition(const double *V, double *Z, int maxit);
  // Forms X + D for the current state without ending the run (used to
  // estimate ||x|| mid-run).
  void peek(const double *Xbase, double *Xout);
  // Column scaling of the right-hand side handed to run(): the correction
  // added to X is multiplied column-wise by d_scale (device, p entries).
  void set_scale(const double *d_scale) { scale_ = d_scale; }
  // Polynomial preconditioner on top of N (flexible runs only); null = off.
  void set_poly(Poly *poly) { poly_ = poly; }
  int iterations() const { return k_; }
  int nslots() const { return nslots_; }
  int ngroups() const { return ng_; }
  int width() const { return S_; }
  const std::vector<GroupDesc> &groups() const { return h_groups_; }
  // Stops (or resumes) the Krylov recurrence of group g; a stopped group keeps
  // the first K block columns for its solution.
  void set_active(int g, bool active, int K);
  void upload_groups();
  bool shifted() const { return shifted_; }

 private:
  template <typename TV> int run_t(const double *R0, double *X, int maxit, const StopFn &stop,
                                   double *rnorm_out);
  template <typename TV> void form(double *X, double alpha);
  // kprev: the last earlier basis block (-1 for V_0)
  template <typename TV> void cholqr2(void *W, void *Vout, double *Cout, size_t cstride, int kprev);
  template <typename TV> void orthogonalize(int k, double *col);
  double *col_region(int k);

  const double *scale_ = nullptr;
  Poly *poly_ = nullptr;
  // Graph replay of a nested run: staged_ = the input is already in W_ and
  // the correction stays in Xtmp_ (no R0 cast, no update of X).
  bool staged_ = false;
  hipGraphExec_t gexec_ = nullptr;
  int gits_ = 0;               // iterations the captured chain runs
  bool gfail_ = false;         // capture or instantiation failed: eager from now on
  std::unique_ptr<BlockArnoldi> inner_;   // nested FP32 block GMRES (inner_its > 0)
  hipEvent_t own_ev_ = nullptr;           // nested runs: their own event ...
  double *own_rn_ = nullptr;              // ... and pinned residual buffer
  bpf_ctx ctx_;
  bpf_csr A_;
  bpf_prec N_;
  ArnoldiConfig cfg_;
  int S_ = 0, ng_ = 0, nslots_ = 0, gmax_rows_ = 0;
  int vec4_ = 0;               // GemmArgs::vec4 holds for this run's groups and blocks
  int rows_sum_ = 0;           // > 0: groups stacked by rows (sum of heights)
  int k_ = 0;                  // completed block columns
  bool shifted_ = false;
  int *h_fl_ = nullptr;        // pinned snapshots of the CholQR flag word
  int h_fl_cap_ = 0;
  hipStream_t st_;
  // Second stream (the context's): the Hessenberg update of iteration k runs
  // there while iteration k+1's preconditioner, SpMM and Gram-Schmidt are
  // queued on st_.
  hipStream_t hs_ = nullptr;
  hipEvent_t ev_main_ = nullptr, ev_hess_ = nullptr;
  std::vector<GroupDesc> h_groups_;
  DevBuf d_groups_;
  BlockArena V_, Z_;           // basis and (flexible) preconditioned basis
  DevBuf W_, P_;               // work blocks (TV), P: preconditioned block
  // Hessenberg arena: one region per block column k (pointer table).
  std::vector<DevBuf> cols_;
  PtrTable coltab_, vtab_, ztab_;
  DevBuf tmpcol_;              // second-pass coefficients
  CholQrWork cq_;
  DevBuf E_;                   // current [Z_k; Lambda] per group
  DevBuf rnorm_;
  DevBuf Y_;                   // back substitution output
  DevBuf Zb_, Rinvb_;          // back substitution scratch
  DevBuf &part_ = cq_.part;    // projection partial sums
  DevBuf &flags_ = cq_.flags;
  DevBuf Xtmp_;                // FP64 n x p scratch (non-flexible solution assembly)
  DevBuf Vf_, Zf_;             // FP32 copies for an FP32 preconditioner
};

// Block GMRES over one n x p Krylov block of any width (FP64; the dense
// p x p algebra on rocBLAS / rocSOLVER), src/block_gmres.jl:170-411.
void gmres_wide(bpf_ctx ctx, bpf_csr A, bpf_prec M, bpf_prec N, int p, const double *dB, int ldb,
                const double *dX0, int ldx0, double *dX, int ldx, const bpf_solver_params &prm,
                bpf_solver_stats &st, double *h_hist, int hist_cap);

// p x p algebra of block BiCGSTAB for blocks wider than 128 on rocBLAS /
// rocSOLVER: n x p blocks row-major, p x p matrices column-major.
struct WideDense {
  void *h = nullptr;           // rocblas_handle of the context
  hipStream_t s = nullptr;
  int p = 0;
  DevBuf ipiv, info;
  void init(bpf_ctx ctx, int p);
  void proj(int n, const double *Vm, const double *Wm, double *out);   // out = Vm^T Wm
  void update(int n, const double *Vm, const double *psi, double *Y, double alpha, double beta);  // Y = a Vm psi + b Y
  void lu(double *K);                                                  // LU with partial pivoting
  void solve(const double *K, double *rhs);                            // rhs = K^{-1} rhs
};

// Row-major FP64 residual R = B - A X.
void residual_rm(hipStream_t s, bpf_csr A, int p, const double *B, const double *X, double *R);
// Per-column backward error from row-major FP64 R, X, B; host output.  With
// b_norms_known, B's column norms are the ones an earlier call on the same B
// left in stat.
void backward_error_rm(bpf_ctx ctx, bpf_csr A, int p, const double *R, const double *X,
                       const double *B, double *h_berr, DevBuf &stat, bool b_norms_known = false);
void scale_columns(hipStream_t s, int n, int p, double *X, int ld, const double *d_scale);

}  // namespace bpf
