// Dense tail of a fixed-pivot sparse LU refactorization.
//
// With a fill-reducing ordering the last rows of the factor form a deep,
// narrow elimination tail (one long row per dependency level).  Eliminated
// row by row it is a long sequential chain; held as a dense m x m Schur
// complement S it is a blocked right-looking LU (no pivoting: the pivots are
// the ones fixed by the host analysis) whose updates are GEMMs on matrix
// cores, and the matching triangular solves are blocked the same way:
//
//   for each nb-column panel j0 of S:
//     dlu_diag_kernel     LU of the nb x nb diagonal block in LDS (16-column
//                         sub-panels, MFMA updates) + inverses of its L and U
//     dgemm_batch_kernel  P = S21 U11^{-1},  Q = L11^{-1} S12   (scratch)
//     copy                S21 <- P, S12 <- Q
//     dgemm_batch_kernel  S22 -= P Q
//
//   solves (m x p right-hand sides, FP64), blocked substitution:
//     Y_b = L_bb^{-1} X_b, X_{>b} -= L_{>b,b} Y_b, then Z_b = U_bb^{-1} Y_b,
//     Y_{<b} -= U_{<b,b} Z_b;
//   or by explicit inverse factors L^{-1}, U^{-1} (built from the diagonal-
//   block inverses by recursive doubling): two block-triangular products.
#include "dense_lu.h"
#include "profile.h"

#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

namespace bpf {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 acc) {
  // D(16x16) += A(16x4) B(4x16): lane l supplies A[l%16][l/16] and
  // B[l/16][l%16] and holds D[l/16 + 4q][l%16] in acc[q].
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

constexpr int kLd = kDluNb + 1;

__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Lane L of each row of 16 lanes, broadcast to that row (DPP row_newbcast).
template <int L>
__device__ __forceinline__ double row_bcast(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + L, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + L, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double pick4(const double (&x)[4], int q) {
  return q == 0 ? x[0] : q == 1 ? x[1] : q == 2 ? x[2] : x[3];
}

// Register layout of a 16 x 16 tile in one wavefront: lane l holds column
// c = l % 16, rows 4 (l / 16) + q in x[q].  The steps below are templates on
// the pivot index so that every lane index and register index is a constant.
struct TileLane {
  int c, g, lane, pw;
};

// Elimination step KK: pivot by one lane read, its column spread along the
// rows of 16 lanes by DPP, its row to the other lane groups by a permute;
// multipliers are products with the pivot's reciprocal.  Every lane runs the
// same instructions: rows at or above the pivot get a zero multiplier,
// columns at or left of it a zero pivot-row entry, and the pivot column
// takes the multipliers.  Outside the pw x pw block the tile is the
// identity, which the steps leave unchanged.
template <int KK>
__device__ __forceinline__ void lu_step(double (&x)[4], const TileLane &t, int *status, int zbase) {
  constexpr int SRC = (KK >> 2) * 16 + KK;   // lane holding G[KK][KK] in x[KK & 3]
  double d = readlane_d(x[KK & 3], SRC);
  if (d == 0.0) {
    if (t.lane == 0) atomicMin(status, zbase + KK);
    d = 1.0;
    if (t.lane == SRC) x[KK & 3] = 1.0;
  }
  const double dinv = 1.0 / d;
  const double u = __shfl(x[KK & 3], (KK >> 2) * 16 + t.c);   // G[KK][c]
  const double up = t.c > KK ? u : 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool below = 4 * t.g + q > KK;
    const double m = below ? row_bcast<KK>(x[q]) * dinv : 0.0;   // G[r][KK] / d
    const double xn = fma(-m, up, x[q]);
    x[q] = (below && t.c == KK) ? m : xn;
  }
}

// Step I of the two substitutions on the identity: U X = I at row 15 - I
// (rdiag: 1 / U[c][c] on the lanes of group c / 4) and L X = I at row I.
template <int I>
__device__ __forceinline__ void inv_step(const double (&x)[4], double (&bu)[4], double (&bl)[4], double rdiag,
                                         const TileLane &t) {
  constexpr int KU = 15 - I, KL = I;
  const double rk = readlane_d(rdiag, (KU >> 2) * 16 + KU);
  if (t.g == (KU >> 2)) bu[KU & 3] *= rk;
  const double xu = __shfl(bu[KU & 3], (KU >> 2) * 16 + t.c);   // X[KU][c]
  const double xl = __shfl(bl[KL & 3], (KL >> 2) * 16 + t.c);   // X[KL][c]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 4 * t.g + q;
    const double urk = r < KU ? row_bcast<KU>(x[q]) : 0.0;   // U[r][KU]
    const double lrk = r > KL ? row_bcast<KL>(x[q]) : 0.0;   // L[r][KL]
    bu[q] = fma(-urk, xu, bu[q]);
    bl[q] = fma(-lrk, xl, bl[q]);
  }
}

template <int... K>
__device__ __forceinline__ void lu_steps(std::integer_sequence<int, K...>, double (&x)[4], const TileLane &t,
                                         int *status, int zbase) {
  (lu_step<K>(x, t, status, zbase), ...);
}
template <int... K>
__device__ __forceinline__ void inv_steps(std::integer_sequence<int, K...>, const double (&x)[4], double (&bu)[4],
                                          double (&bl)[4], double rdiag, const TileLane &t) {
  (inv_step<K>(x, bu, bl, rdiag, t), ...);
}

// One wavefront: LU without pivoting of the pw x pw diagonal tile at
// G[p0][p0] (pw <= 16) and the inverses of both its triangles.  Lane l holds
// column c = l % 16, rows 4 (l / 16) + q in x[q] (the identity outside the
// pw x pw block).  Each pivot is read by one lane read, its column spread
// along the rows of 16 lanes by DPP, its row to the other lane groups by a
// permute; multipliers are products with the pivot's reciprocal.  U_PP^{-1}
// and L_PP^{-1} then follow by column-oriented substitution on the identity
// (interleaved), with the reciprocals of U's diagonal formed once.  The
// tile goes back to G, the inverses to sU / sL and to the diagonal blocks of
// Uinv / Linv.  A zero pivot at row r sets *status = min(*status, base + p0 + r).
__device__ __forceinline__ void tile_lu(double *G, int p0, int pw, double (*sU)[17], double (*sL)[17],
                                        double *Linv, double *Uinv, int *status, int base, int lane) {
  const TileLane t{lane & 15, lane >> 4, lane, pw};
  const int c = t.c, g = t.g;
  double x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 4 * g + q;
    x[q] = (r < pw && c < pw) ? G[(p0 + r) * kLd + p0 + c] : (r == c ? 1.0 : 0.0);
  }
  lu_steps(std::make_integer_sequence<int, 16>{}, x, t, status, base + p0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 4 * g + q;
    if (r < pw && c < pw) G[(p0 + r) * kLd + p0 + c] = x[q];
  }
  // 1 / U[c][c], held by the lanes of group c / 4 (register c % 4)
  const double rdiag = g == (c >> 2) ? 1.0 / pick4(x, c & 3) : 1.0;
  double bu[4], bl[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) bu[q] = bl[q] = (4 * g + q == c) ? 1.0 : 0.0;
  inv_steps(std::make_integer_sequence<int, 16>{}, x, bu, bl, rdiag, t);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 4 * g + q;
    sU[r][c] = bu[q];
    sL[r][c] = bl[q];
    if (r < pw && c < pw) {
      Uinv[(size_t)(p0 + r) * kDluNb + p0 + c] = bu[q];
      Linv[(size_t)(p0 + r) * kDluNb + p0 + c] = bl[q];
    }
  }
}

// Diagonal tile B of a kDluNb x kDluNb inverse in the MFMA output layout
// (lane l: rows 16 B + l / 16 + 4 q, column 16 B + l % 16); the identity
// outside the w x w block.
__device__ __forceinline__ dbl4 diag_tile(const double *Inv, int B, int w, int lane) {
  const int lr = lane >> 4, lc = lane & 15;
  dbl4 t;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 16 * B + lr + 4 * q, c = 16 * B + lc;
    t[q] = (r < w && c < w) ? Inv[(size_t)r * kDluNb + c] : (r == c ? 1.0 : 0.0);
  }
  return t;
}

// D = A_II B for the diagonal inverse tile A_II (Inv, block I) and B in the
// MFMA output layout (which is the B-operand layout of step q = 0..3).
__device__ __forceinline__ dbl4 diag_times(const double *Inv, int I, int w, dbl4 b, int lane) {
  const int lr = lane >> 4, lc = lane & 15;
  dbl4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int r = 16 * I + lc, c = 16 * I + 4 * s + lr;
    const double a = (r < w && c < w) ? Inv[(size_t)r * kDluNb + c] : (r == c ? 1.0 : 0.0);
    d = mfma4(a, b[s], d);
  }
  return d;
}

// Block column J of X = U^{-1} (U: upper part of G): X_JJ from the diagonal
// inverse, then for I = J-1 .. 0: X_IJ = -X_II sum_{I<M<=J} U_IM X_MJ.
// Writes the column's tiles (zeros below the diagonal block) to Uinv.
template <int J>
__device__ __forceinline__ void inv_ucol(const double *G, double *Uinv, int w, int nbk, int lane) {
  if (J >= nbk) return;
  const int lr = lane >> 4, lc = lane & 15;
  dbl4 X[J + 1];
  X[J] = diag_tile(Uinv, J, w, lane);
#pragma unroll
  for (int I = J - 1; I >= 0; --I) {
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int M = I + 1; M <= J; ++M)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mfma4(G[(16 * I + lc) * kLd + 16 * M + 4 * s + lr], X[M][s], acc);
    X[I] = -diag_times(Uinv, I, w, acc, lane);
  }
#pragma unroll
  for (int I = 0; I < 8; ++I) {
    if (I >= nbk) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * I + lr + 4 * q, c = 16 * J + lc;
      if (I != J) Uinv[(size_t)r * kDluNb + c] = I < J ? X[I < J ? I : 0][q] : 0.0;
    }
  }
}

// Block column J of Z = L^{-1} (L: strictly lower part of G, unit
// diagonal): Z_JJ from the diagonal inverse, then for I = J+1 .. nbk-1:
// Z_IJ = -Z_II sum_{J<=M<I} L_IM Z_MJ.  Writes the column's tiles (zeros
// above the diagonal block) to Linv.
template <int J>
__device__ __forceinline__ void inv_lcol(const double *G, double *Linv, int w, int nbk, int lane) {
  if (J >= nbk) return;
  const int lr = lane >> 4, lc = lane & 15;
  dbl4 Z[8];
  Z[J] = diag_tile(Linv, J, w, lane);
#pragma unroll
  for (int I = J + 1; I < 8; ++I) {
    if (I >= nbk) break;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int M = J; M < I; ++M)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mfma4(G[(16 * I + lc) * kLd + 16 * M + 4 * s + lr], Z[M][s], acc);
    Z[I] = -diag_times(Linv, I, w, acc, lane);
  }
#pragma unroll
  for (int I = 0; I < 8; ++I) {
    if (I >= nbk) break;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 16 * I + lr + 4 * q, c = 16 * J + lc;
      if (I != J) Linv[(size_t)r * kDluNb + c] = I > J ? Z[I][q] : 0.0;
    }
  }
}

// LU without pivoting of the w x w block at S (ld lds), w <= kDluNb, in LDS:
// 16-column sub-panels.  One wavefront factors the sub-panel's 16 x 16
// diagonal block and inverts both of its triangles (tile_lu); the L column
// block (G_rP U_PP^{-1}), the U row block (L_PP^{-1} G_Pc) and the trailing
// update then run on matrix cores, the next diagonal tile updated and
// factored by wave 0 while the others update the rest.  The 16 x 16 inverses
// are also the diagonal blocks of U^{-1} and L^{-1}, which are completed
// blockwise at the end.
// A zero pivot at block row r sets *status = min(*status, base + r).
__global__ __launch_bounds__(1024) void dlu_diag_kernel(double *S, int lds, int w, double *Linv, double *Uinv,
                                                        int *status, int base) {
  __shared__ double G[kDluNb * kLd];
  __shared__ double sU[16][17], sL[16][17];   // current sub-panel's U_PP^{-1}, L_PP^{-1}
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  {
    // 128 x 128 as 8,192 pairs of columns, eight per thread, loads first
    double2 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + i * 1024, r = e >> 6, c = (e & 63) * 2;
      const double *src = S + (size_t)r * lds + c;
      if (r < w && c + 1 < w) v[i] = *reinterpret_cast<const double2 *>(src);
      else v[i] = double2{r < w && c < w ? src[0] : 0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + i * 1024, r = e >> 6, c = (e & 63) * 2;
      G[r * kLd + c] = v[i].x;
      G[r * kLd + c + 1] = v[i].y;
    }
  }
  __syncthreads();
  // Sub-panel 0's diagonal tile, then per sub-panel: its L column and U row
  // blocks (all wavefronts), and the trailing update with look-ahead: wave 0
  // updates the next diagonal tile first and factors it while the other
  // wavefronts update the rest of the trailing block.
  if (wave == 0) tile_lu(G, 0, min(16, w), sU, sL, Linv, Uinv, status, base, lane);
  __syncthreads();
  for (int p0 = 0; p0 < w; p0 += 16) {
    const int pw = min(16, w - p0);
    const int t0 = p0 + 16;
    const int nt = (w - t0 + 15) / 16;
    if (nt <= 0) break;
    // L column block G[r][P] <- G[r][P] U_PP^{-1} and U row block
    // G[P][c] <- L_PP^{-1} G[P][c], one 16 x 16 tile per wavefront turn; each
    // wavefront reads its whole tile before overwriting it.
    for (int tt = wave; tt < 2 * nt; tt += 16) {
      const int t = tt >> 1;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      if ((tt & 1) == 0) {
        const int rb = t0 + t * 16;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const int k = k0 + lr;
          acc = mfma4(rb + lc < w ? G[(rb + lc) * kLd + p0 + k] : 0.0, sU[k][lc], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = rb + lr + 4 * q;
          if (r < w && lc < pw) G[r * kLd + p0 + lc] = acc[q];
        }
      } else {
        const int cb = t0 + t * 16;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const int k = k0 + lr;
          acc = mfma4(sL[lc][k], cb + lc < w ? G[(p0 + k) * kLd + cb + lc] : 0.0, acc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = lr + 4 * q, c = cb + lc;
          if (r < pw && c < w) G[(p0 + r) * kLd + c] = acc[q];
        }
      }
    }
    __syncthreads();
    // trailing update G[r][c] -= sum_k L[r][p0+k] U[p0+k][c], r, c >= t0:
    // tile 0 (the next diagonal tile) by wave 0, which then factors it; the
    // other tiles by wavefronts 1-15
    auto trail = [&](int tt) {
      const int ti = tt / nt, tj = tt % nt;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4) {
        const int k = p0 + k0 + lr;
        acc = mfma4(G[(t0 + ti * 16 + lc) * kLd + k], G[k * kLd + t0 + tj * 16 + lc], acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = t0 + ti * 16 + lr + 4 * q, j = t0 + tj * 16 + lc;
        if (i < w && j < w) G[i * kLd + j] -= acc[q];
      }
    };
    if (wave == 0) {
      trail(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      tile_lu(G, t0, min(16, w - t0), sU, sL, Linv, Uinv, status, base, lane);
    } else {
      for (int tt = wave; tt < nt * nt; tt += 15) trail(tt);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int e = tid + i * 1024, r = e >> 6, c = (e & 63) * 2;
    double *dst = S + (size_t)r * lds + c;
    if (r < w && c + 1 < w) *reinterpret_cast<double2 *>(dst) = double2{G[r * kLd + c], G[r * kLd + c + 1]};
    else if (r < w && c < w) dst[0] = G[r * kLd + c];
  }
  // The inverses from the 16 x 16 diagonal-block inverses (written to Uinv /
  // Linv by tile_lu): block column J of U^{-1} by wavefront J, of L^{-1} by
  // wavefront 8 + J, each column's tiles kept in registers (an MFMA output
  // tile is laid out as the B operand of the next product).
  const int nbk = (w + 15) / 16;
  if (wave < 8) {
    switch (wave) {
      case 0: inv_ucol<0>(G, Uinv, w, nbk, lane); break;
      case 1: inv_ucol<1>(G, Uinv, w, nbk, lane); break;
      case 2: inv_ucol<2>(G, Uinv, w, nbk, lane); break;
      case 3: inv_ucol<3>(G, Uinv, w, nbk, lane); break;
      case 4: inv_ucol<4>(G, Uinv, w, nbk, lane); break;
      case 5: inv_ucol<5>(G, Uinv, w, nbk, lane); break;
      case 6: inv_ucol<6>(G, Uinv, w, nbk, lane); break;
      default: inv_ucol<7>(G, Uinv, w, nbk, lane); break;
    }
  } else {
    switch (wave - 8) {
      case 0: inv_lcol<0>(G, Linv, w, nbk, lane); break;
      case 1: inv_lcol<1>(G, Linv, w, nbk, lane); break;
      case 2: inv_lcol<2>(G, Linv, w, nbk, lane); break;
      case 3: inv_lcol<3>(G, Linv, w, nbk, lane); break;
      case 4: inv_lcol<4>(G, Linv, w, nbk, lane); break;
      case 5: inv_lcol<5>(G, Linv, w, nbk, lane); break;
      case 6: inv_lcol<6>(G, Linv, w, nbk, lane); break;
      default: inv_lcol<7>(G, Linv, w, nbk, lane); break;
    }
  }
}

__global__ __launch_bounds__(256) void copy_block_kernel(int rows, int cols, const double *src, int lds,
                                                         double *dst, int ldd) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)rows * cols) return;
  const int r = (int)(e / cols), c = (int)(e % cols);
  dst[(size_t)r * ldd + c] = src[(size_t)r * lds + c];
}

__global__ __launch_bounds__(256) void tail_move_kernel(int cnt, const int *pos, const int *didx, double *fac,
                                                        double *S, int to_dense) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= cnt) return;
  if (to_dense) S[didx[e]] = fac[pos[e]];
  else fac[pos[e]] = S[didx[e]];
}

void copy_block(hipStream_t s, int rows, int cols, const double *src, int lds, double *dst, int ldd) {
  const size_t m = (size_t)rows * cols;
  if (m == 0) return;
  hipLaunchKernelGGL(copy_block_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, rows, cols, src,
                     lds, dst, ldd);
}

// ------------------------------------------------ inverse triangular factors
// One GEMM of a batch: C = alpha A B + beta C, row-major FP64.  M and K are
// multiples of 16, N any; tri = 1: A is block lower triangular (a row tile in
// kDluNb block b takes K < (b + 1) kDluNb), tri = 2: block upper
// (K >= b kDluNb).  C must not overlap A or B.
struct GemmDesc {
  const double *A, *B;
  double *C;
  int M, N, K, lda, ldb, ldc;
  double alpha, beta;
  int tri;
  int tiles_n;   // output tiles across N
  int tile0;     // first output tile of this GEMM in the batch
};
constexpr int kMaxGemm = 8;
struct GemmBatch {
  GemmDesc d[kMaxGemm];
  int nd, ntiles;
};

// One workgroup per (16 TM) x (16 TN) output tile of one GEMM of the batch;
// its four wavefronts split the k range (wavefront w takes the 16-wide
// chunks w, w + 4, ...) and their partial tiles are summed in LDS in a fixed
// order (bitwise reproducible).  Workgroups go to the 8 XCDs in turn, so
// when the tile count is a multiple of 8 workgroup b takes tile
// (b % 8) * (ntiles / 8) + b / 8: each XCD gets a contiguous run of tiles
// (the column tiles of the same row tiles) and fetches its A rows into its
// own L2 once.  Operands come straight from global memory:
// per chunk, lane group g = lane / 16 owns k = k0 + 4g .. k0 + 4g + 3 (a
// 32-byte run of each of its A rows and four B rows), so MFMA step s pairs A
// and B on k = k0 + 4g + s.  The next chunk's operands are loaded before the
// current chunk's products.
template <int TM, int TN>
__global__ __launch_bounds__(256) void dgemm_batch_kernel(GemmBatch b) {
  __shared__ dbl4 red[3][TM * TN][64];
  const int t = b.ntiles % 8 == 0 ? (int)(blockIdx.x % 8) * (b.ntiles / 8) + (int)(blockIdx.x / 8) : blockIdx.x;
  int id = 0;
  while (id + 1 < b.nd && t >= b.d[id + 1].tile0) ++id;
  const GemmDesc g = b.d[id];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g4 = lane >> 4, l16 = lane & 15;
  const int tt = t - g.tile0;
  const int m0 = (tt / g.tiles_n) * 16 * TM, n0 = (tt % g.tiles_n) * 16 * TN;
  int kb = 0, ke = g.K;
  if (g.tri == 1) ke = min(g.K, (m0 / kDluNb + 1) * kDluNb);
  else if (g.tri == 2) kb = (m0 / kDluNb) * kDluNb;
  const double *ar[TM];
#pragma unroll
  for (int i = 0; i < TM; ++i) ar[i] = g.A + (size_t)(m0 + 16 * i + l16) * g.lda + 4 * g4;
  const double *bc[TN];
  bool ok[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int c = n0 + 16 * j + l16;
    ok[j] = c < g.N;
    bc[j] = g.B + (size_t)(4 * g4) * g.ldb + (ok[j] ? c : 0);
  }
  dbl4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  double2 a[TM][2], an[TM][2];
  double bv[TN][4], bn[TN][4];
  auto load = [&](int k, double2 (&aa)[TM][2], double (&bb)[TN][4]) {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      aa[i][0] = *reinterpret_cast<const double2 *>(ar[i] + k);
      aa[i][1] = *reinterpret_cast<const double2 *>(ar[i] + k + 2);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) bb[j][q] = ok[j] ? bc[j][(size_t)(k + q) * g.ldb] : 0.0;
  };
  const int k1 = kb + 16 * wave;
  if (k1 < ke) load(k1, a, bv);
  for (int k = k1; k < ke; k += 64) {
    if (k + 64 < ke) load(k + 64, an, bn);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const double av = q == 0 ? a[i][0].x : q == 1 ? a[i][0].y : q == 2 ? a[i][1].x : a[i][1].y;
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = mfma4(av, bv[j][q], acc[i][j]);
      }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      a[i][0] = an[i][0];
      a[i][1] = an[i][1];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) bv[j][q] = bn[j][q];
  }
  if (wave > 0)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) red[wave - 1][i * TN + j][lane] = acc[i][j];
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] += red[w][i * TN + j][lane];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    if (!ok[j]) continue;
    const int c = n0 + 16 * j + l16;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double *cp = g.C + (size_t)(m0 + 16 * i + g4 + 4 * q) * g.ldc + c;
        *cp = g.beta == 0.0 ? g.alpha * acc[i][j][q] : fma(g.beta, *cp, g.alpha * acc[i][j][q]);
      }
  }
}

// The inverses of the diagonal blocks into the diagonal blocks of Zl / Zu.
__global__ __launch_bounds__(256) void diag_inv_copy_kernel(int nblk, int m, const double *Linv,
                                                            const double *Uinv, double *Zl, double *Zu) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nblk * kDluNb * kDluNb) return;
  const int b = (int)(e / (kDluNb * kDluNb)), r = (int)(e / kDluNb % kDluNb), c = (int)(e % kDluNb);
  const size_t d = (size_t)(b * kDluNb + r) * m + b * kDluNb + c;
  Zl[d] = Linv[e];
  Zu[d] = Uinv[e];
}

void add_gemm(std::vector<GemmDesc> &v, int M, int N, int K, double alpha, const double *A, int lda,
              const double *B, int ldb, double beta, double *C, int ldc, int tri) {
  BPF_REQUIRE(M % 16 == 0 && K % 16 == 0 && lda % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0,
              "dense-tail GEMM operand shape");
  if (M == 0 || N <= 0) return;
  GemmDesc d{A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, tri, 0, 0};
  v.push_back(d);
}

// Runs the GEMMs of v (independent of each other), up to kMaxGemm per launch.
template <int TM, int TN>
void run_gemms(hipStream_t s, const std::vector<GemmDesc> &v) {
  for (size_t i0 = 0; i0 < v.size(); i0 += kMaxGemm) {
    GemmBatch b{};
    b.nd = (int)std::min<size_t>(kMaxGemm, v.size() - i0);
    int tiles = 0;
    for (int i = 0; i < b.nd; ++i) {
      GemmDesc d = v[i0 + i];
      d.tiles_n = (d.N + 16 * TN - 1) / (16 * TN);
      d.tile0 = tiles;
      tiles += d.M / (16 * TM) * d.tiles_n;
      b.d[i] = d;
    }
    b.ntiles = tiles;
    if (tiles == 0) continue;
    hipLaunchKernelGGL((dgemm_batch_kernel<TM, TN>), dim3(tiles), dim3(256), 0, s, b);
  }
}

struct Comb {
  int lo, mid, hi;   // block ranges [lo, mid) and [mid, hi)
};

void split_blocks(int lo, int hi, int depth, std::vector<std::vector<Comb>> &lv) {
  if (hi - lo <= 1) return;
  const int mid = lo + (hi - lo + 1) / 2;
  split_blocks(lo, mid, depth + 1, lv);
  split_blocks(mid, hi, depth + 1, lv);
  if ((int)lv.size() <= depth) lv.resize(depth + 1);
  lv[depth].push_back({lo, mid, hi});
}

}  // namespace

void dense_lu_invert(hipStream_t s, int m, const double *S, const double *Linv, const double *Uinv,
                     double *Zl, double *Zu, double *T) {
  BPF_REQUIRE(m % kDluNb == 0, "dense tail of %d rows is not a multiple of %d", m, kDluNb);
  const int nb = kDluNb, nblk = m / nb;
  const size_t cnt = (size_t)nblk * nb * nb;
  hipLaunchKernelGGL(diag_inv_copy_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, nblk, m, Linv,
                     Uinv, Zl, Zu);
  // Recursive doubling, deepest combines first.  For ranges A = [lo, mid),
  // C = [mid, hi) of a unit lower L and an upper U:
  //   (L^{-1})_CA = -(L^{-1})_CC (L_CA (L^{-1})_AA)
  //   (U^{-1})_AC = -(U^{-1})_AA (U_AC (U^{-1})_CC)
  // with the inner products staged in T at the block's own position (the
  // combines of one depth cover disjoint blocks).
  std::vector<std::vector<Comb>> lv;
  split_blocks(0, nblk, 0, lv);
  for (int d = (int)lv.size() - 1; d >= 0; --d) {
    std::vector<GemmDesc> first, second;
    for (const Comb &c : lv[d]) {
      const int a0 = c.lo * nb, c0 = c.mid * nb, a = (c.mid - c.lo) * nb, cc = (c.hi - c.mid) * nb;
      add_gemm(first, cc, a, a, 1.0, S + (size_t)c0 * m + a0, m, Zl + (size_t)a0 * m + a0, m, 0.0,
               T + (size_t)c0 * m + a0, m, 0);
      add_gemm(first, a, cc, cc, 1.0, S + (size_t)a0 * m + c0, m, Zu + (size_t)c0 * m + c0, m, 0.0,
               T + (size_t)a0 * m + c0, m, 0);
      add_gemm(second, cc, a, cc, -1.0, Zl + (size_t)c0 * m + c0, m, T + (size_t)c0 * m + a0, m, 0.0,
               Zl + (size_t)c0 * m + a0, m, 1);
      add_gemm(second, a, cc, a, -1.0, Zu + (size_t)a0 * m + a0, m, T + (size_t)a0 * m + c0, m, 0.0,
               Zu + (size_t)a0 * m + c0, m, 2);
    }
    run_gemms<2, 2>(s, first);
    run_gemms<2, 2>(s, second);
  }
  BPF_CHECK_HIP(hipGetLastError());
}

void dense_lu_solve_inv(hipStream_t s, int m, int p, const double *S, const double *Linv, const double *Uinv,
                        bool build, double *Zl, double *Zu, double *T, double *X, double *Y) {
  const int tok = prof::begin(s);
  double bytes = 0.0, flops = 0.0;
  if (build) {
    dense_lu_invert(s, m, S, Linv, Uinv, Zl, Zu, T);
    // two triangular inverses from their diagonal blocks (m^3 / 3 each): S
    // read, both written once
    flops += 2.0 / 3.0 * m * (double)m * m;
    bytes += 2.0 * m * (double)m * sizeof(double);
  }
  std::vector<GemmDesc> v;
  add_gemm(v, m, p, m, 1.0, Zl, m, X, p, 0.0, Y, p, 1);   // Y = L^{-1} X
  if (p >= 32) run_gemms<1, 2>(s, v); else run_gemms<1, 1>(s, v);
  v.clear();
  add_gemm(v, m, p, m, 1.0, Zu, m, Y, p, 0.0, X, p, 2);   // X = U^{-1} Y
  if (p >= 32) run_gemms<1, 2>(s, v); else run_gemms<1, 1>(s, v);
  BPF_CHECK_HIP(hipGetLastError());
  // two triangular m x m products (m^2 p flops each); compulsory traffic:
  // both inverse factors' triangles once, X read and written twice
  const double mp = (double)m * p;
  prof::end(tok, s, prof::DENSE_LU, bytes + (double)m * m * sizeof(double) + 4.0 * mp * sizeof(double),
            flops + 2.0 * m * mp);
}

void dense_tail_move(hipStream_t s, int cnt, const int *pos, const int *didx, double *fac, double *S,
                     bool to_dense) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(tail_move_kernel, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, pos, didx, fac, S,
                     to_dense ? 1 : 0);
  BPF_CHECK_HIP(hipGetLastError());
}

void dense_lu_factor(hipStream_t s, int m, double *S, double *Linv, double *Uinv, double *P, double *Q,
                     int *status, int base) {
  // (16-row GEMM tiles; the diagonal-block kernel moves column pairs of S)
  BPF_REQUIRE(m % 16 == 0 && reinterpret_cast<uintptr_t>(S) % 16 == 0, "dense tail of %d rows", m);
  const int nb = kDluNb;
  const int tok = prof::begin(s);
  double flops = 0.0;
  for (int j0 = 0, b = 0; j0 < m; j0 += nb, ++b) {
    const int w = std::min(nb, m - j0);
    const int rest = m - j0 - w;
    // diagonal-block LU (2/3 w^3) and the inverses of its two triangles (2/3 w^3)
    flops += 4.0 / 3.0 * w * w * w + 4.0 * rest * w * w + 2.0 * rest * rest * w;
    double *S11 = S + (size_t)j0 * m + j0;
    double *Li = Linv + (size_t)b * nb * nb, *Ui = Uinv + (size_t)b * nb * nb;
    hipLaunchKernelGGL(dlu_diag_kernel, dim3(1), dim3(1024), 0, s, S11, m, w, Li, Ui, status, base + j0);
    if (rest > 0) {
      double *S21 = S + (size_t)(j0 + w) * m + j0, *S12 = S11 + w, *S22 = S21 + w;
      std::vector<GemmDesc> v;
      add_gemm(v, rest, w, w, 1.0, S21, m, Ui, nb, 0.0, P, nb, 0);    // P = S21 U11^{-1}
      add_gemm(v, w, rest, w, 1.0, Li, nb, S12, m, 0.0, Q, m, 0);     // Q = L11^{-1} S12
      run_gemms<2, 2>(s, v);
      copy_block(s, rest, w, P, nb, S21, m);
      copy_block(s, w, rest, Q, m, S12, m);
      v.clear();
      add_gemm(v, rest, rest, w, -1.0, P, nb, Q, m, 1.0, S22, m, 0);  // S22 -= P Q
      run_gemms<2, 2>(s, v);
    }
  }
  BPF_CHECK_HIP(hipGetLastError());
  // compulsory traffic: S read and written once
  prof::end(tok, s, prof::DENSE_LU, 2.0 * m * (double)m * sizeof(double), flops);
}

void dense_lu_solve(hipStream_t s, int m, int p, const double *S, const double *Linv, const double *Uinv,
                    double *X, double *Y) {
  // Forward: Y_b = L_bb^{-1} X_b, X_{>b} -= L_{>b,b} Y_b.
  const int nb = kDluNb;
  const int nblk = (m + nb - 1) / nb;
  const int tok = prof::begin(s);
  std::vector<GemmDesc> v;
  auto run = [&]() {
    if (p >= 32) run_gemms<1, 2>(s, v); else run_gemms<1, 1>(s, v);
    v.clear();
  };
  for (int b = 0; b < nblk; ++b) {
    const int j0 = b * nb, w = std::min(nb, m - j0), rest = m - j0 - w;
    add_gemm(v, w, p, w, 1.0, Linv + (size_t)b * nb * nb, nb, X + (size_t)j0 * p, p, 0.0, Y + (size_t)j0 * p, p, 0);
    run();
    if (rest > 0) {
      add_gemm(v, rest, p, w, -1.0, S + (size_t)(j0 + w) * m + j0, m, Y + (size_t)j0 * p, p, 1.0,
               X + (size_t)(j0 + w) * p, p, 0);
      run();
    }
  }
  // Backward: X_b = U_bb^{-1} Y_b, Y_{<b} -= U_{<b,b} X_b.
  for (int b = nblk - 1; b >= 0; --b) {
    const int j0 = b * nb, w = std::min(nb, m - j0);
    add_gemm(v, w, p, w, 1.0, Uinv + (size_t)b * nb * nb, nb, Y + (size_t)j0 * p, p, 0.0, X + (size_t)j0 * p, p, 0);
    run();
    if (j0 > 0) {
      add_gemm(v, j0, p, w, -1.0, S + j0, m, X + (size_t)j0 * p, p, 1.0, Y, p, 0);
      run();
    }
  }
  BPF_CHECK_HIP(hipGetLastError());
  // two sweeps of m x m triangles plus the full diagonal-block inverses:
  // 2 m^2 p + 2 nb m p flops; compulsory traffic S once, X read and written
  const double mp = (double)m * p;
  prof::end(tok, s, prof::DENSE_LU, (double)m * m * sizeof(double) + 2.0 * mp * sizeof(double),
            2.0 * m * mp + 2.0 * nb * mp);
}

}  // namespace bpf
