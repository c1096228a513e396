// Dense block kernels of the block-Krylov solvers: argument structures and
// launchers (see dense_kernels.hip for the kernels).
#pragma once

#include "bpf_internal.h"

namespace bpf {

// One independent Krylov group: columns [c0, c0 + w) and rows [r0, r0 + nr)
// of the n x p blocks.  `slot` indexes the group's columns in per-column
// arrays (residual estimates, scales); `K` caps the block columns used when
// forming the solution; inactive groups are skipped by the Arnoldi kernels.
struct GroupDesc {
  int c0;
  int w;
  int r0;
  int nr;
  int slot;
  int K;
  int active;
  int pad_;
};

// Psi_{i,g} = V_{i0+i}[:, g]^T W[:, g] for i < nb.
struct ProjArgs {
  const void *const *Vp;  // device table of block pointers (block j = Vp[j]),
  const void *V1;         // or a single block when Vp is null (nb = 1)
  int i0, nb, ldv;
  const void *W;
  int ldw;
  int n;
  const GroupDesc *groups;
  int ng;
  int chunk_rows;      // set by the launcher
  double *part;        // set by the launcher
  // Groups stacked by rows (disjoint row ranges): the sum of their row
  // counts, which sizes the chunks; empty chunks are skipped.  0 = column
  // groups over all n rows.
  int rows_sum = 0;
  // Rows x width of the groups that take part (traffic accounting only;
  // negative = all groups).
  double cells = -1.0;
  // 1: width-16 FP32 blocks accumulate each row chunk's sums in FP32 on
  // FP32 matrix cores (the chunks are still added in FP64).
  int f32acc = 0;
};

// Y[:, g] = beta*Y[:, g] + alpha * sum_{i<nb} V_{i0+i}[:, g] * Psi_{i,g}
struct GemmArgs {
  const void *const *Vp;
  const void *V1;
  int i0, nb, ldv;
  const double *psi;   // Psi_{i,g} at psi + i*psi_istride + g*psi_gstride (S x S, ld ldpsi)
  size_t psi_istride, psi_gstride;
  int ldpsi;
  void *Y;
  int ldy;
  int n;
  double alpha, beta;
  const GroupDesc *groups;
  // 0: inactive groups are skipped.  1 (solution assembly): a group takes
  // part iff K > 0, over its first min(nb, K) blocks.
  int use_k = 0;
  // 1: groups that do not take part leave Y untouched even when beta == 0
  // (passes over a subset of the groups); 0: their fresh Y is zeroed.
  int keep_skipped = 0;
  int rows_sum = 0;    // stacked groups: sum of heights (traffic accounting)
  double cells = -1.0; // rows x width of the groups that take part (accounting; negative = all)
  // 1: every group's c0 and width, ldv and ldy are multiples of 4 and the
  // blocks and Y are 16-byte aligned (the FP32 width-16 update then loads a
  // row's four columns per lane as one vector).
  int vec4 = 0;
};

struct CholArgs {
  const double *G;     // per group at G + g*gstride, S x S
  double *R, *Rinv;
  size_t gstride;
  int S;
  int n;               // rows of the block (for the shift)
  const GroupDesc *groups;
  int *flags;          // [0]: bits kept over a solve, [1]: bits of this factorization;
                       // bit 0: a shift was needed, bit 1: a column was deflated,
                       // bit 2: a pivot at rounding level (kept, re-projected)
  // 1: a pivot that is not positive (column k lies in the span of the columns
  // before it to rounding) gives R a zero row k instead of a shift, with 1 on
  // the diagonal of the factor that Rinv inverts; launch_fill_deflated then
  // replaces column k of W Rinv.
  int deflate = 0;
  // Per-group bits of this factorization (deflation 2, rounding-level pivot
  // 4) at gflags[g], written for active groups when deflate is set (may be null).
  int *gflags = nullptr;
};

// Hessenberg arena: block column k is a separate region cols[k] holding, per
// group g, k+2 blocks of S x S (Psi_0..Psi_k, C; after the QR update the
// rotated R blocks and the compact reflector of column k), then the tau
// vectors (ng x S), the final right-hand-side blocks Z_k (ng x S x S) and the
// triangular factors T of the compact WY form Omega_k = I - Y T Y^T (ng x S x S).
//   block (k, g, i) = cols[k] + (g*(k+2) + i)*S*S
//   tau   (k, g)    = cols[k] + ng*(k+2)*S*S + g*S
//   zfin  (k, g)    = cols[k] + ng*(k+2)*S*S + ng*S + g*S*S
//   tmat  (k, g)    = cols[k] + ng*(k+2)*S*S + ng*S + ng*S*S + g*S*S
__host__ __device__ inline size_t hcol_elems(int k, int ng, int S) {
  return (size_t)ng * (k + 2) * S * S + (size_t)ng * S + 2 * (size_t)ng * S * S;
}

struct HessArgs {
  double *const *cols;
  int k;               // current column
  int S;
  int ng;
  const GroupDesc *groups;
};

// QR update of column k; E (per group 2S x S at E + g*2*S*S) holds [Z_k; 0]
// on entry.  On exit zfin(k) = rotated Z_k, E = [Lambda; 0] (the next Z),
// rnorm[c0 + c] = ||Lambda[:, c]||_2.
struct HessQrArgs {
  double *const *cols;
  int k;
  int S;
  int ng;
  double *E;
  const GroupDesc *groups;
  double *rnorm;
};

struct BacksubArgs {
  double *const *cols;
  int K;               // number of block columns
  int S;
  int ng;
  double *Y;           // Y_i of group g at Y + (i*ng + g)*S*S
  const GroupDesc *groups;
  double *Z;           // scratch, same layout as Y: the running right-hand side
  double *Rinv;        // scratch, same layout as Y: inverses of the diagonal blocks
};

int pad_width(int w);
void fill_int(hipStream_t s, int *p, int v, int n);
void fill_double(hipStream_t s, double *p, double v, size_t n);
void fill_float(hipStream_t s, float *p, float v, size_t n);

// out block for (i, g) at out + i*o_istride + g*o_gstride (S x S).
template <typename TVV, typename TW>
void launch_proj(hipStream_t s, ProjArgs a, int S, double *out, size_t o_istride, size_t o_gstride,
                 bool accumulate, DevBuf &part);
template <typename TVV, typename TY>
void launch_block_gemm(hipStream_t s, GemmArgs a, int S, int ng);
void launch_chol(hipStream_t s, CholArgs a, int ng);
// Splits the groups by their per-group CholQR bits: out[g] (breakdown, bits
// 2 or 4) and out[ng + g] (none) copy groups[g], each active only where
// groups[g] is active and the bits say so.
void launch_split_groups(hipStream_t s, const GroupDesc *groups, const int *gflags, int ng, GroupDesc *out);
// After a deflating Cholesky (bit 1 of *flags): column k of group g's rows of
// Q (row-major, ld) becomes pseudo-random values of unit expected norm
// wherever R[g] (S x S at g*stride) has a zero diagonal entry k.  Returns at
// once when no column was deflated.
template <typename TV>
void launch_fill_deflated(hipStream_t s, TV *Q, int ld, const double *R, int S, size_t stride,
                          const GroupDesc *groups, int ng, int gmax_rows, const int *flags);
// C_g = A_g B_g for upper-triangular S x S blocks (A_g, B_g at g*stride,
// C_g at g*cstride).
void launch_trmm_upper(hipStream_t s, const double *A, const double *B, double *C, int S,
                       size_t stride, size_t cstride, const GroupDesc *groups, int ng);
void launch_hess_apply(hipStream_t s, HessArgs a, int ng);
void launch_hess_qr(hipStream_t s, HessQrArgs a, int ng);
void launch_backsub(hipStream_t s, BacksubArgs a, int ng);

// Y_i[:, b] *= scale[c0 + b] for the K x ng blocks of a back-substitution result.
void launch_scale_ycols(hipStream_t s, double *Y, int K, int ng, int S, const GroupDesc *groups,
                        const double *scale);
// X[:, c] *= scale[c] (row-major n x p), or /= when invert.
void launch_scale_cols(hipStream_t s, int n, int p, double *X, int ld, const double *scale, bool invert);
// Per-group forms for systems stacked by rows (block-diagonal operator):
// out[slot + c] = max_r |X[r, c0 + c]| (mode 0) or sum_r X[r, c0 + c]^2
// (mode 1) over the group's rows; out holds nslots entries.  Mode 1 needs
// part with colstat_groups_blocks(gmax_rows) * nslots doubles (sums are
// formed from per-block partials in a fixed order: reproducible).
int colstat_groups_blocks(int gmax_rows);
template <typename T>
void colstat_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                    const T *X, int ld, int mode, double *out, int nslots, double *part);
// X[r, c0 + c] /= scale[slot + c] (invert) or *= over each group's rows.
void launch_scale_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                         double *X, int ld, const double *scale, bool invert);
// Between per-group column-major blocks cm[g] (nr x w, leading dimension
// ldc[g]) and the stacked row-major block rm (leading dimension ld).
void launch_pack_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                        double *const *cm, const int *ldc, double *rm, int ld, bool to_rm);
// Small dense LU with partial pivoting (one workgroup per group): A (w x w,
// ld S) overwritten by L\U, piv (S ints); solve applies it to B (w x w).
void launch_lu_small(hipStream_t s, double *A, int *piv, int S, const GroupDesc *groups, int ng);
void launch_lu_small_solve(hipStream_t s, const double *A, const int *piv, double *B, int S,
                           const GroupDesc *groups, int ng);
// out = sum over the n x p row-major block of x .* y (FP64), device scalar,
// from kDotBlocks-or-fewer per-block partials in part, added in a fixed order.
constexpr int kDotBlocks = 2048;
template <typename T>
void launch_dot(hipStream_t s, size_t m, const T *x, const T *y, double *out, double *part);

void colmax_colmajor(hipStream_t s, int n, int p, const double *X, int ld, double *out);
// Per-column inf-norms (mode 0) or sums of squares (mode 1) of a row-major
// n x p block into out[p]; mode 1 needs part with colstat_blocks(n) * p doubles.
int colstat_blocks(int n);
template <typename T>
void colstat_rowmajor(hipStream_t s, int n, int p, const T *X, int ld, int mode, double *out,
                      double *part = nullptr);
void axpby(hipStream_t s, size_t m, double alpha, const double *x, double beta, double *y);
// y[g*stride + e] += x[g*stride + e] for e < count, g < ng (one launch).
void axpby_groups(hipStream_t s, int ng, size_t count, size_t stride, const double *x, double *y);
template <typename TS, typename TD>
void copy_cast(hipStream_t s, int n, int p, const TS *src, int lds, TD *dst, int ldd, double scale);

}  // namespace bpf
