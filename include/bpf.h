/*
 * bpf.h -- C ABI of the MI355X-native multi-right-hand-side solve library
 * (libbpf_amd.so).  Hand-written gfx950 HIP kernels behind plain C entry
 * points: opaque handles, plain pointers and sizes, int status (0 = OK,
 * otherwise a BPF_E* code; bpf_last_error() gives the message of the last
 * failure on the calling thread).
 *
 * Every entry point replaces a call site of exanauts/BlockPowerFlow.jl; the
 * citation on each declaration names the reference interface it stands in
 * for.  Dense blocks at this boundary are column-major n x p with a leading
 * dimension, exactly as the reference's Julia matrices (CuMatrix / ROCMatrix);
 * indices are 0-based int32 (the reference converts its 1-based Cint arrays
 * with `.- 1` before handing them to the vendor library, src/cusolverRF.jl:251).
 *
 * Pointers named d_* are device pointers (allocated with bpf_malloc or any
 * HIP allocator on the context's device); pointers named h_* are host memory.
 * All work of a context runs in order on that context's stream.  A handle may
 * be used by one host thread at a time.
 */
#ifndef BPF_H
#define BPF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
  BPF_OK = 0,
  BPF_E_INVALID = 1,     /* bad argument (shape, size, null handle)        */
  BPF_E_HIP = 2,         /* HIP runtime error                              */
  BPF_E_SINGULAR = 3,    /* zero / non-positive pivot                      */
  BPF_E_NOMEM = 4,       /* device or host allocation failed               */
  BPF_E_UNSUPPORTED = 5, /* a size or option this build does not support   */
  BPF_E_BREAKDOWN = 6    /* Krylov breakdown that could not be repaired    */
};

/* ---- element types and layouts ---------------------------------------- */
enum { BPF_F64 = 0, BPF_F32 = 1 };
enum { BPF_COL_MAJOR = 0, BPF_ROW_MAJOR = 1 };

typedef struct bpf_ctx_s *bpf_ctx;
typedef struct bpf_csr_s *bpf_csr;
typedef struct bpf_prec_s *bpf_prec;
typedef struct bpf_rf_s *bpf_rf;
typedef struct bpf_batch_s *bpf_batch;

const char *bpf_last_error(void);
const char *bpf_version(void);

/* Process-wide execution options (no reference counterpart).
 * BPF_OPT_SYNCFREE: 1 = triangular solves and ILU(0)/IC(0)/LU
 * factorizations run as one sync-free launch (one wavefront per row waiting
 * on its dependencies' completion flags); 0 (default) = one launch per wide
 * dependency level and one single-workgroup launch for runs of narrow levels.
 * On MI355X the device-scope flag handshake costs more than the launches it
 * saves (each release/acquire writes back / invalidates the per-XCD L2), so
 * the sync-free form is slower there; it is kept as an option.
 * BPF_OPT_AUTOTUNE: 1 (default) = the first SpMM of a matrix at a given
 * block width times each launch shape (lanes per row x columns per lane) on
 * the actual operands and keeps the fastest; 0 = a fixed shape per width.
 * BPF_OPT_OVERLAP: 1 (default) = block GMRES runs the block-Hessenberg QR
 * update of iteration k (src/block_gmres.jl:318-342) on a second stream
 * while iteration k+1's preconditioner, SpMM and Gram-Schmidt run, testing
 * convergence of k meanwhile (one iteration's work is dropped at the end);
 * 0 = one iteration after another.  Results are identical either way.
 * BPF_OPT_BATCH_THREADS: host threads that solve the bins of a batched solve
 * (systems grouped by RHS count) at the same time, each bin on its own pair
 * of streams (default 4; 1 = one bin after another).  Bins share nothing, so
 * results are identical for every value.
 * BPF_OPT_WIDE_MIN: block GMRES in FP64 with at least this many right-hand
 * sides (and block_width 0 or >= p) runs them as one n x p Krylov block, as
 * the reference does for every p (src/block_gmres.jl:46-56): the sparse
 * operator and preconditioners on the library's kernels, the p x p and
 * 2p x p algebra (Gram-Schmidt products, Householder QR of the basis and
 * Hessenberg blocks, reflector applications, back-substitution) on rocBLAS
 * and rocSOLVER.  Default 129: up to 128 right-hand sides run on the
 * library's own matrix-core kernels; block_width < p splits wider blocks
 * into independent Krylov blocks of at most 128 columns there instead.
 * BPF_OPT_GRAPH: 1 (default) = in the mixed form with an inner FP32 block
 * GMRES as the preconditioner step (inner_its > 0), the inner run's fixed
 * chain of inner_its iterations is captured into a HIP graph on its first
 * call and replayed on every later outer iteration (the input cast into the
 * inner run's own block before the replay, the correction copied out after
 * it); 0 = every launch issued from the host.  Results are identical.
 * BPF_OPT_GRAPH_REPLAYS: read-only count of those replays in this process
 * (set to 0 to reset).
 * BPF_OPT_RELABEL_MIN: block GMRES in the MIXED precision on a system of at
 * least this many unknowns (default 32768; 0 = never), with no left
 * preconditioner and an ILU(0)/IC(0) or no right one, runs in a reverse
 * Cuthill-McKee
 * relabelling of A's pattern (built once per pattern and kept with the
 * handles): A's relabelled copy for the products and relabelled views of the
 * preconditioner's triangular operators (the same factor, levels and
 * arithmetic per row), the blocks permuted at the layout conversion, so that
 * gathered rows lie near the rows gathering them.
 * BPF_OPT_TAIL_SLAB: column width (0 or 64; default 64) of the slabs that
 * the narrow-level part of a level-scheduled triangular solve over more than
 * that many right-hand sides is split into, one workgroup per slab walking
 * the narrow levels on its own (columns are independent); 0 = one workgroup
 * for all columns.  Results are identical.
 * BPF_OPT_DENSE_INVERSE: the dense tail of a refactorization handle (the last
 * rows of a large factor, solved as a dense block) is solved by explicit
 * inverses of its two triangular factors, built by recursive doubling on
 * matrix cores at each refactorization (two block-triangular products per
 * solve instead of a chain of diagonal-block steps; built by the first solve
 * that uses them after a refactorization): 0 = never (blocked substitution),
 * 1 (default) = in the FP32 solves that refinement follows, 2 = in every
 * solve.
 * BPF_OPT_DENSE_BLOCKS: the dense tail of a refactorization handle created
 * from now on spans at most this many 128-row blocks (default 12; 0 = no
 * dense tail), and at most about n / 16 rows.
 * BPF_OPT_TAIL_ROUNDS: a dependency level of a triangular solve joins the
 * single narrow-level launch when its rows fit in this many rounds of that
 * launch's wavefronts (default 2).  Results are identical. */
enum { BPF_OPT_SYNCFREE = 1, BPF_OPT_AUTOTUNE = 2, BPF_OPT_OVERLAP = 3, BPF_OPT_BATCH_THREADS = 4,
       BPF_OPT_WIDE_MIN = 5, BPF_OPT_GRAPH = 6, BPF_OPT_GRAPH_REPLAYS = 7, BPF_OPT_RELABEL_MIN = 8,
       BPF_OPT_TAIL_SLAB = 9, BPF_OPT_DENSE_INVERSE = 10, BPF_OPT_DENSE_BLOCKS = 11, BPF_OPT_TAIL_ROUNDS = 12 };
int bpf_set_option(int option, long long value);
int bpf_get_option(int option, long long *value);

/* ---- context, device memory ------------------------------------------ */
/* Replaces the implicit CUDA.jl / AMDGPU.jl task-local device and default
 * stream that every reference call runs on (src/BlockPowerFlow.jl:9-15). */
int bpf_device_count(int *count);
int bpf_ctx_create(int device, bpf_ctx *ctx);
int bpf_ctx_destroy(bpf_ctx ctx);
int bpf_ctx_synchronize(bpf_ctx ctx);
/* Device allocation and copies; stand in for CuMatrix{T}(undef, n, p),
 * copyto!, Array(x) (e.g. src/block_gmres.jl:45-56, test/cusolverRF.jl:23). */
int bpf_malloc(bpf_ctx ctx, void **d_ptr, size_t bytes);
int bpf_free(bpf_ctx ctx, void *d_ptr);
int bpf_memcpy_h2d(bpf_ctx ctx, void *d_dst, const void *h_src, size_t bytes);
int bpf_memcpy_d2h(bpf_ctx ctx, void *h_dst, const void *d_src, size_t bytes);
int bpf_memcpy_d2d(bpf_ctx ctx, void *d_dst, const void *d_src, size_t bytes);
/* Event timer over the context's stream (milliseconds between start/stop). */
int bpf_timer_start(bpf_ctx ctx);
int bpf_timer_stop(bpf_ctx ctx, double *ms);

/* In-process kernel timing (no reference counterpart; the reference only
 * reports wall time, src/block_gmres.jl:408, src/block_utils.jl:57).  When enabled, instrumented
 * launches on any context are bracketed by HIP events on their stream and
 * summed per family with their algorithmic bytes and flops:
 *   0 SpMM, 1 triangular-solve level launch, 2 block projection (V^T W),
 *   3 block update (V Psi), 4 sparse factorization, 5 dense-tail LU of the
 *   refactorization (its blocked factorization and its solves), 6 and 7 the
 *   projections and updates of blocks at most 32 columns wide (2 and 3 count
 *   widths of 64 and more).
 * Enabling (or disabling) resets the counters. */
enum { BPF_K_SPMM = 0, BPF_K_TRSV = 1, BPF_K_PROJ = 2, BPF_K_GEMM = 3, BPF_K_FACTOR = 4, BPF_K_DENSE_LU = 5,
       BPF_K_PROJ_NARROW = 6, BPF_K_GEMM_NARROW = 7 };
int bpf_profile_enable(bpf_ctx ctx, int on);
int bpf_profile_read(bpf_ctx ctx, int family, long long *launches, double *ms, double *bytes,
                     double *flops);

/* ---- CSR matrices ------------------------------------------------------ */
/* CuSparseMatrixCSR{T,Cint}(J) / ROCSparseMatrixCSR(A): upload of a square
 * CSR matrix with int32 indices and float64 values (benchmarks/benchmarks_amd.jl:43,
 * test/cusolverRF.jl:8).  Column indices are sorted on upload if needed. */
int bpf_csr_create(bpf_ctx ctx, int n, int nnz, const int32_t *h_rowptr,
                   const int32_t *h_colidx, const double *h_val, bpf_csr *A);
/* New values on the same pattern (gJ.nzVal .*= 2.0, test/cusolverRF.jl:24):
 * host or device source array of nnz float64 in the pattern's order. */
int bpf_csr_set_values(bpf_csr A, const double *h_val);
int bpf_csr_set_values_device(bpf_csr A, const double *d_val);
int bpf_csr_info(bpf_csr A, int *n, int *nnz, double *norm_inf);
int bpf_csr_destroy(bpf_csr A);

/* ---- SpMM ---------------------------------------------------------------- */
/* Y = alpha*A*X + beta*Y with X, Y n x p blocks in the given layout and type.
 * Replaces mul!(W, A, P) (src/block_gmres.jl:299, src/block_bicgstab.jl:252,264)
 * which dispatches to cuSPARSE / rocSPARSE SpMM. */
int bpf_spmm(bpf_ctx ctx, bpf_csr A, int dtype, int layout, int p,
             const void *d_X, int ldx, void *d_Y, int ldy, double alpha, double beta);

/* ---- ILU(0) / IC(0) preconditioners -------------------------------------- */
enum { BPF_PREC_ILU0 = 0, BPF_PREC_IC0 = 1, BPF_PREC_BLOCK_JACOBI = 2 };
/* ilu0(A, nrhs) / ilu0(A) (src/preconditioners_nvidia.jl:42-47,
 * src/preconditioners_amd.jl:43-46) and ic0(A, nrhs) / ic0(A)
 * (src/preconditioners_nvidia.jl:6-11, src/preconditioners_amd.jl:5-9):
 * level analysis of A's pattern (host, once) plus the numeric factorization
 * on the device.  The factor keeps A's pattern (combined CSR: strict lower =
 * L with unit diagonal, upper with diagonal = U; for IC0 the lower part with
 * diagonal = L). */
int bpf_prec_create(bpf_ctx ctx, bpf_csr A, int kind, bpf_prec *P);
/* Numeric refactorization with A's current values (same pattern). */
int bpf_prec_factor(bpf_prec P, bpf_csr A);
/* y = P^{-1} x : ldiv!(y, ilu, x) (src/preconditioners_nvidia.jl:49-59,
 * src/preconditioners_amd.jl:55-60); ldiv!(y, ic, x) (nvidia :13-23,
 * amd :18-23).  Also what mul!(Y, N, P) means for block_bicgstab's N
 * (src/block_bicgstab.jl:251,263).  d_Y may alias d_X. */
int bpf_prec_apply(bpf_prec P, int dtype, int layout, int p,
                   const void *d_X, int ldx, void *d_Y, int ldy);
/* LS.BlockJacobiPreconditioner(J, nblocks, device) + LS.update (ExaPF, as
 * passed to block_gmres / block_bicgstab by benchmarks/benchmarks_nvidia.jl:46-51,71
 * and test/block_gmres.jl:5,18-25): the unknowns are split into nblocks parts of
 * the symmetrized graph of A (ceil(n / nblocks) <= 128 unknowns each,
 * BPF_E_UNSUPPORTED beyond), the dense diagonal blocks A[part, part] are
 * inverted on the device (BPF_E_SINGULAR for a singular block).
 * bpf_prec_apply (mul!(y, P, x)) computes y[part] = inv(A[part, part]) x[part];
 * bpf_prec_factor recomputes the inverses from A's current values
 * (LS.update).  bpf_block_jacobi_partition writes the part of each row. */
int bpf_block_jacobi_create(bpf_ctx ctx, bpf_csr A, int nblocks, bpf_prec *P);
int bpf_block_jacobi_partition(bpf_prec P, int32_t *h_part);
/* Copy the combined factor values (nnz float64, A's pattern) to the host. */
int bpf_prec_get_factor(bpf_prec P, double *h_val);
/* Number of level sets of the lower / upper solve schedules. */
int bpf_prec_levels(bpf_prec P, int *nlevels_lower, int *nlevels_upper);
int bpf_prec_destroy(bpf_prec P);

/* ---- block Krylov solvers ------------------------------------------------ */
enum {
  BPF_STATUS_UNKNOWN = 0,
  BPF_STATUS_SOLVED = 1,    /* "solution good enough given atol and rtol" */
  BPF_STATUS_TIRED = 2,     /* "maximum number of iterations exceeded"    */
  BPF_STATUS_OVERTIMED = 3  /* "time limit exceeded"                      */
};
/* FP64: everything in FP64 (the drop-in).
 * MIXED: FP32 SpMM / ILU(0) solves / polynomial inside the preconditioner,
 *   FP64 flexible Krylov basis, Gram matrices, Hessenberg and residual.
 * MIXED32: FP32 basis as well, FP64 iterative refinement on the outside. */
enum { BPF_PRECISION_FP64 = 0, BPF_PRECISION_MIXED = 1, BPF_PRECISION_MIXED32 = 2 };

typedef struct {
  int memory;          /* block-basis memory hint (reference default 20)       */
  int restart;         /* 1: restarted block-GMRES(memory)                      */
  int itmax;           /* 0: 2*div(n,p) for GMRES, 2n for BiCGSTAB              */
  double atol, rtol;   /* block Frobenius criterion ||R|| <= atol + rtol||R0||  */
  double timemax;      /* seconds; <= 0 means no limit                          */
  int verbose;         /* print every `verbose` iterations (0 = silent)         */
  int history;         /* record residual norms                                 */
  /* Additions (no counterpart in the reference; defaults reproduce it): */
  double berr_tol;     /* > 0: also stop when every RHS has backward error
                          ||Ax-b||inf/(||A||inf||x||inf+||b||inf) <= berr_tol   */
  int precision;       /* BPF_PRECISION_FP64 or BPF_PRECISION_MIXED             */
  int block_width;     /* s: RHS per independent Krylov block (0 = p; blocks
                          wider than 128 outside the one-block FP64 form are
                          split into near-equal blocks of at most 128)         */
  int max_refine;      /* mixed: maximum FP64 refinement steps                  */
  double inner_rtol;   /* mixed: inner reduction factor per refinement step     */
  int reorthogonalization; /* second block Gram-Schmidt pass (reference kwarg) */
  int poly_degree;     /* mixed: degree of the GMRES polynomial applied on top of
                          the ILU(0) preconditioner (0 = ILU(0) alone)          */
  int bmgs_blocks;     /* block MGS (src/block_gmres.jl:301-304) projects against
                          this many consecutive basis blocks per step (one
                          projection and one update launch for them; <= 1 =
                          one block per step, the reference's loop)             */
  int inner_its;       /* mixed: > 0 = the flexible preconditioner step is this
                          many iterations of an FP32 ILU(0)-preconditioned block
                          GMRES on A z = V_k (in place of the polynomial)       */
  int inner_width;     /* mixed, inner_its > 0: columns per independent Krylov
                          block of the inner solve (0 = all p; batched solves:
                          < 0 = chosen per bin by timing, see
                          bpf_batch_inner_widths)                               */
} bpf_solver_params;

typedef struct {
  int niter;           /* Krylov iterations (all refinement steps)              */
  int solved;
  int status;          /* BPF_STATUS_*                                          */
  int refine_steps;    /* FP64 refinement steps taken (mixed)                   */
  double timer;        /* seconds                                               */
  int nresiduals;      /* entries written to the history buffer                 */
  double max_berr;     /* final max backward error over RHS (if berr_tol > 0)   */
} bpf_solver_stats;

void bpf_solver_params_default(bpf_solver_params *prm);

/* block_gmres(A, B; memory, N, ldiv=true, restart, atol, rtol, itmax, timemax,
 * verbose, history) -> (X, stats)  (src/block_gmres.jl:127-154,170-411).
 * N may be NULL (identity).  B, X: device, column-major float64 n x p.
 * h_history (optional, capacity hist_cap) receives ||R_k||_F per iteration. */
int bpf_block_gmres(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p,
                    const double *d_B, int ldb, double *d_X, int ldx,
                    const bpf_solver_params *prm, bpf_solver_stats *stats,
                    double *h_history, int hist_cap);
/* Warm start: block_gmres(A, B, X0; ...) (src/block_gmres.jl:127-140); d_X0
 * holds the initial guess (may alias d_X). */
int bpf_block_gmres_x0(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p,
                       const double *d_B, int ldb, const double *d_X0, int ldx0,
                       double *d_X, int ldx, const bpf_solver_params *prm,
                       bpf_solver_stats *stats, double *h_history, int hist_cap);
/* block_bicgstab(A, B; C=B, N, atol, rtol, itmax, timemax, verbose, history)
 * (src/block_bicgstab.jl:182-299).  d_C may be NULL (C = B).  Up to 128
 * right-hand sides the p x p products and LU solves run on the library's
 * kernels, beyond that on rocBLAS / rocSOLVER. */
int bpf_block_bicgstab(bpf_ctx ctx, bpf_csr A, bpf_prec N, int p,
                       const double *d_B, int ldb, const double *d_C, int ldc,
                       double *d_X, int ldx, const bpf_solver_params *prm,
                       bpf_solver_stats *stats, double *h_history, int hist_cap);
/* block_gmres(A, B[, X0]; M, N, ldiv, ...) with both preconditioners
 * (src/block_gmres.jl:127-154,194-203): M (may be NULL) is the left
 * preconditioner, N (may be NULL) the right one; a preconditioner handle is
 * applied as its inverse approximation, which is what mulorldiv! does with
 * either value of ldiv (src/block_utils.jl:113).  With M the run stops on
 * ||M (B - A X)||_F <= atol + rtol ||M B||_F (:226-230,:341-349).  dtype
 * (BPF_F64 or BPF_F32) is the element type of B, X0 and X: the reference's
 * solvers are generic over it (:170-173).  Float32 blocks are widened to
 * FP64 on the device, the solve runs in FP64 and X is rounded back to
 * Float32.  bpf_block_gmres_x0 is this call with M = NULL, dtype BPF_F64.
 * bpf_block_bicgstab_t is block_bicgstab with the element type given the
 * same way (src/block_bicgstab.jl:182-190). */
int bpf_block_gmres_ex(bpf_ctx ctx, bpf_csr A, bpf_prec M, bpf_prec N, int dtype, int p, const void *d_B,
                       int ldb, const void *d_X0, int ldx0, void *d_X, int ldx, const bpf_solver_params *prm,
                       bpf_solver_stats *stats, double *h_history, int hist_cap);
int bpf_block_bicgstab_t(bpf_ctx ctx, bpf_csr A, bpf_prec N, int dtype, int p, const void *d_B, int ldb,
                         const void *d_C, int ldc, void *d_X, int ldx, const bpf_solver_params *prm,
                         bpf_solver_stats *stats, double *h_history, int hist_cap);
/* Thin QR of a device n x p block (n >= p, p <= 128), in place:
 * householder!(Q, R, tau) (src/block_utils.jl:38-53), the orthonormalization
 * of block_gmres (src/block_gmres.jl:276,316).  Computed by CholQR2 (two
 * Cholesky-QR passes, FP64 Gram matrices on matrix cores); R is upper
 * triangular with a non-negative diagonal (geqrf's R up to row signs).  A
 * column in the span of the columns before it (a Cholesky pivot that is not
 * positive) gets a zero row of R and a pseudo-random orthonormal completion
 * of Q, as Householder QR gives an arbitrary one.  *shifted (optional)
 * reports that the factorization broke down: a dependent column was
 * completed, or a failed pivot was repaired by a shift. */
int bpf_householder(bpf_ctx ctx, int n, int p, double *d_Q, int ldq, double *d_R, int ldr,
                    int *shifted);
/* Block-Hessenberg least squares of block GMRES, column by column on the
 * device exactly as inside the solver: the QR update of each new block
 * column with the stored block reflectors (src/block_gmres.jl:318-337), the
 * residual estimate per RHS after each column (:341-342) and the final block
 * back substitution (:367-375).  h_H: (K+1)p x Kp block upper Hessenberg,
 * h_G: p x p right-hand side [G; 0; ...], column-major host arrays.  Output
 * h_Y (Kp x p) minimises ||[G; 0] - H Y||_F column-wise; h_rnorm (K x p,
 * row k = after column k, optional) holds the residual norms. */
int bpf_block_hessenberg_lsq(bpf_ctx ctx, int p, int K, const double *h_H, int ldh,
                             const double *h_G, int ldg, double *h_Y, int ldy, double *h_rnorm);
/* Eigenvalues of a d x d upper Hessenberg matrix (row-major), host-only
 * utility behind the polynomial preconditioner's harmonic Ritz values. */
int bpf_hessenberg_eigenvalues(const double *h_H, int d, double *h_re, double *h_im);
/* Per-RHS FP64 backward error ||Ax-b||inf/(||A||inf||x||inf+||b||inf) into
 * h_berr[p] (no reference counterpart: the reference prints norm(B - A*X),
 * benchmarks/benchmarks_amd.jl:64). */
int bpf_backward_error(bpf_ctx ctx, bpf_csr A, int p, const double *d_X, int ldx,
                       const double *d_B, int ldb, double *h_berr);

/* bpf_backward_error without any device allocation (no reference
 * counterpart: it is for callers that cannot allocate, such as code under
 * HIP graph capture or a tight serving loop).  Writes the p per-RHS backward
 * errors to device memory d_berr, asynchronously on the context's stream.
 * Its scratch is a fixed device array of 3 x 4096 doubles per GPU, so p is
 * at most 4096 (BPF_E_UNSUPPORTED beyond).  Calls that overlap on one GPU
 * (several contexts or host threads at once) share that array and race. */
int bpf_backward_error_noalloc(bpf_ctx ctx, bpf_csr A, int p, const double *d_X, int ldx,
                               const double *d_B, int ldb, double *d_berr);

/* ---- batched systems ------------------------------------------------------- */
/* Many independent systems of uneven size solved together (no reference
 * counterpart; the reference solves one system per call).  Host CSR arrays
 * are given per system; B/X per system are device column-major blocks. */
int bpf_batch_create(bpf_ctx ctx, int nsys, const int32_t *n, const int32_t *nnz,
                     const int32_t *const *h_rowptr, const int32_t *const *h_colidx,
                     const double *const *h_val, const int32_t *p, bpf_batch *batch);
int bpf_batch_solve(bpf_batch batch, const double *const *d_B, double *const *d_X,
                    const bpf_solver_params *prm, bpf_solver_stats *stats,
                    int32_t *h_solved_per_system);
/* Columns per inner Krylov block that the last bpf_batch_solve used for each
 * system's bin (0 = no inner solve).  With inner_width < 0 (mixed,
 * inner_its > 0) a bin's first solve times each width of 8, 16, 32, 64 (up
 * to its RHS count) on full solves of the bin and keeps the fastest for
 * later solves with the same inner_its. */
int bpf_batch_inner_widths(bpf_batch batch, int32_t *h_width);
int bpf_batch_destroy(bpf_batch batch);

/* ---- fixed-pattern LU refactorization (CUSOLVERRF replacement) -------------- */
enum { BPF_ORDER_AMD = 0, BPF_ORDER_MDQ = 1, BPF_ORDER_METIS = 2, BPF_ORDER_RCM = 3, BPF_ORDER_NATURAL = 4 };
/* CusolverRfLU(A; nrhs, ordering, tol, fast_mode) (src/cusolverRF.jl:222-301):
 * host symbolic analysis = fill-reducing ordering + threshold-pivoted sparse
 * LU (glu_analysis_host, :56-220), then device setup and refactorization.
 * Raises BPF_E_SINGULAR when a pivot is below tol (the reference constructs
 * but never throws SingularException, :164-166). nrhs is not limited. */
int bpf_rf_create(bpf_ctx ctx, bpf_csr A, int ordering, double tol, bpf_rf *rf);
/* update!(rflu, A) / lu!(rflu, A) (src/cusolverRF.jl:304-314,334): new values
 * of A on the analysed pattern, same pivots, device refactorization. */
int bpf_rf_refactor(bpf_rf rf, bpf_csr A);
/* ldiv!(x, rflu, b) (src/cusolverRF.jl:317-324) for nrhs columns: direct solve
 * with the current factors in FP64 (dtype BPF_F64) or with FP32 factors and
 * FP32 triangular solves (BPF_F32, followed by nothing else). d_X may alias d_B. */
int bpf_rf_solve(bpf_rf rf, int dtype, int nrhs, const double *d_B, int ldb,
                 double *d_X, int ldx);
/* FP32 triangular solves plus at most `refine` FP64 iterative-refinement
 * steps with the true residual B - A X (no reference counterpart).  With
 * berr_tol > 0 the per-RHS backward error is checked on every residual and
 * refinement stops once all nrhs values are <= berr_tol; berr_tol = 0 runs
 * all `refine` steps.  h_berr (nrhs, optional) receives the final backward
 * errors, h_steps (optional) the refinement steps taken. */
int bpf_rf_solve_refined(bpf_rf rf, bpf_csr A, int nrhs, const double *d_B, int ldb,
                         double *d_X, int ldx, int refine, double berr_tol, double *h_berr,
                         int *h_steps);
/* Host copies of the analysis: permutations P, Q (n each, B = A[P,Q]) and the
 * factor pattern / values (combined L\U CSR of B). */
int bpf_rf_info(bpf_rf rf, int *n, int *nnz_factor, int *nlevels_lower, int *nlevels_upper);
int bpf_rf_get_factor(bpf_rf rf, int32_t *h_P, int32_t *h_Q, int32_t *h_rowptr,
                      int32_t *h_colidx, double *h_val);
int bpf_rf_destroy(bpf_rf rf);   /* cudestroy!(rflu), src/cusolverRF.jl:54 */

/* CusolverRfLUBatch(A, batchsize; ordering, tol, fast_mode)
 * (src/cusolverRF.jl:336-419): one pattern, batchsize value sets. */
int bpf_rf_batch_create(bpf_ctx ctx, bpf_csr A, int batchsize, int ordering,
                        double tol, bpf_rf *rf);
/* rf_batch_refactor!(rflu, Anzval) (src/cusolverRF.jl:472-481): d_nzval is
 * nnz x batchsize column-major float64 (one value set per column). */
int bpf_rf_batch_refactor(bpf_rf rf, const double *d_nzval, int ld);
/* ldiv!(X, rflu, B) for the batch (src/cusolverRF.jl:451-460): column i of B
 * is solved with factor i. */
int bpf_rf_batch_solve(bpf_rf rf, const double *d_B, int ldb, double *d_X, int ldx);

#ifdef __cplusplus
}
#endif
#endif /* BPF_H */
