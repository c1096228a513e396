// Dense block kernels of the block-Krylov solvers (gfx950).
//
// Every n x p block is row-major (element (r, c) at X[r*ld + c]).  The p
// right-hand sides are split into groups of width s; each group is an
// independent block-Krylov instance, so its blocks are the column slices
// [c0, c0+s) of the n x p blocks.  Small per-group matrices (s x s) are FP64
// row-major with leading dimension S (S = s rounded up to 16/32/64/128).
//
//   proj_partial / proj_reduce  Psi_i = V_i^T W over row chunks  (TN GEMM, FP64 accumulate)
//   block_gemm                  Y = beta*Y + alpha*sum_i V_i Psi_i  (NN GEMM, FP64 accumulate)
//   chol_kernel                 R = chol(G) (upper), R^{-1}, shift on breakdown
//   hess_apply_kernel           apply the stored block reflectors to a new block column
//   hess_qr_kernel              Householder QR of [Psi_kk; C] and rotation of [Z_k; 0]
//   backsub_kernel              block upper-triangular solve R Y = Z
//   colstats kernels            per-column inf-norms / 2-norms, backward error
#include "dense.h"
#include "profile.h"

#include <algorithm>
#include <cmath>

namespace bpf {

namespace {

constexpr int kB = 256;
// Dynamic LDS per workgroup used before falling back to global memory.
constexpr size_t kLdsLimit = 64 * 1024;

template <typename T> __device__ __forceinline__ double to_d(T v) { return static_cast<double>(v); }

// ---------------------------------------------------------------- helpers
__global__ void fill_int_kernel(int *p, int v, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void fill_kernel(T *p, T v, size_t n) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
  // For non-negative doubles the IEEE bit pattern orders like the value.
  atomicMax(reinterpret_cast<unsigned long long *>(addr), __double_as_longlong(v));
}

__global__ void csr_norm_cast_kernel(int n, const int *rowptr, const double *val, float *val32,
                                     double *out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      s += fabs(val[k]);
      val32[k] = static_cast<float>(val[k]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, s);
}

// --------------------------------------------------------- TN projection
// part[((chunk * nb + i) * ng + g) * S*S + a*S + b] = sum_{r in chunk} V_i[r, c0+a] W[r, c0+b]
template <typename TVV, typename TW, int S>
__global__ __launch_bounds__(kB) void proj_partial_kernel(ProjArgs a) {
  constexpr int TM = S / 16;
  constexpr int RB = 16;
  __shared__ double sv[RB][S + 1];
  __shared__ double sw[RB][S + 1];
  const int i = blockIdx.x;          // basis block (fastest: blocks sharing W run together)
  const int chunk = blockIdx.y;
  const int g = blockIdx.z;
  const GroupDesc gd = a.groups[g];
  const TVV *V = static_cast<const TVV *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0;
  const TW *W = static_cast<const TW *>(a.W) + gd.c0;
  const int r0 = chunk * a.chunk_rows;
  const int r1 = min(a.n, r0 + a.chunk_rows);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[TM][TM];
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int q = 0; q < TM; ++q) acc[m][q] = 0.0;
  for (int rb = r0; rb < r1; rb += RB) {
    for (int e = threadIdx.x; e < RB * S; e += kB) {
      const int rr = e / S, cc = e % S;
      const int r = rb + rr;
      const bool ok = r < r1 && cc < gd.w;
      sv[rr][cc] = ok ? to_d(V[(size_t)r * a.ldv + cc]) : 0.0;
      sw[rr][cc] = ok ? to_d(W[(size_t)r * a.ldw + cc]) : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < RB; ++rr) {
      double va[TM], wb[TM];
#pragma unroll
      for (int m = 0; m < TM; ++m) va[m] = sv[rr][ty * TM + m];
#pragma unroll
      for (int q = 0; q < TM; ++q) wb[q] = sw[rr][tx * TM + q];
#pragma unroll
      for (int m = 0; m < TM; ++m)
#pragma unroll
        for (int q = 0; q < TM; ++q) acc[m][q] = fma(va[m], wb[q], acc[m][q]);
    }
    __syncthreads();
  }
  double *out = a.part + ((size_t)(chunk * a.nb + i) * a.ng + g) * S * S;
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int q = 0; q < TM; ++q) out[(ty * TM + m) * S + tx * TM + q] = acc[m][q];
}

// out[i*o_istride + g*o_gstride + e] (+)= sum_chunk part[...]
__global__ __launch_bounds__(kB) void proj_reduce_kernel(const double *part, int nchunk, int nb,
                                                         int ng, int SS, double *out,
                                                         size_t o_istride, size_t o_gstride,
                                                         int accumulate) {
  const size_t total = (size_t)nb * ng * SS;
  const size_t e = (size_t)blockIdx.x * kB + threadIdx.x;
  if (e >= total) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * total + e];
  const size_t ig = e / SS, el = e % SS;
  const size_t i = ig / ng, g = ig % ng;
  double *o = out + i * o_istride + g * o_gstride + el;
  *o = accumulate ? *o + s : s;
}

// ------------------------------------------------------------- NN GEMM
// Y[r, c0+b] = beta*Y[r, c0+b] + alpha * sum_i sum_a V_i[r, c0+a] * Psi_i[a][b]
// Psi_i for group g at psi + (i*psi_istride) + g*psi_gstride, row-major S x S.
template <typename TVV, typename TY, int S>
__global__ __launch_bounds__(kB) void block_gemm_kernel(GemmArgs a) {
  constexpr int RT = 64;            // rows per workgroup
  constexpr int TR = RT / 16;       // rows per thread
  constexpr int TC = S / 16;        // cols per thread
  constexpr int KA = 16;            // inner chunk
  __shared__ double sv[RT][KA + 1];
  __shared__ double sp[KA][S + 1];
  const int g = blockIdx.y;
  const GroupDesc gd = a.groups[g];
  const int r0 = blockIdx.x * RT;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[TR][TC];
#pragma unroll
  for (int m = 0; m < TR; ++m)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[m][q] = 0.0;
  for (int i = 0; i < a.nb; ++i) {
    const TVV *V = static_cast<const TVV *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0;
    const double *P = a.psi + (size_t)i * a.psi_istride + (size_t)g * a.psi_gstride;
    for (int k0 = 0; k0 < gd.w; k0 += KA) {
      for (int e = threadIdx.x; e < RT * KA; e += kB) {
        const int rr = e / KA, kk = e % KA;
        const int r = r0 + rr, k = k0 + kk;
        sv[rr][kk] = (r < a.n && k < gd.w) ? to_d(V[(size_t)r * a.ldv + k]) : 0.0;
      }
      for (int e = threadIdx.x; e < KA * S; e += kB) {
        const int kk = e / S, cc = e % S;
        const int k = k0 + kk;
        sp[kk][cc] = (k < gd.w && cc < gd.w) ? P[(size_t)k * a.ldpsi + cc] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < KA; ++kk) {
        double vr[TR], pc[TC];
#pragma unroll
        for (int m = 0; m < TR; ++m) vr[m] = sv[ty * TR + m][kk];
#pragma unroll
        for (int q = 0; q < TC; ++q) pc[q] = sp[kk][tx * TC + q];
#pragma unroll
        for (int m = 0; m < TR; ++m)
#pragma unroll
          for (int q = 0; q < TC; ++q) acc[m][q] = fma(vr[m], pc[q], acc[m][q]);
      }
      __syncthreads();
    }
  }
  TY *Y = static_cast<TY *>(a.Y) + gd.c0;
#pragma unroll
  for (int m = 0; m < TR; ++m) {
    const int r = r0 + ty * TR + m;
    if (r >= a.n) continue;
#pragma unroll
    for (int q = 0; q < TC; ++q) {
      const int c = tx * TC + q;
      if (c >= gd.w) continue;
      TY *y = Y + (size_t)r * a.ldy + c;
      const double v = a.alpha * acc[m][q];
      *y = static_cast<TY>(a.beta == 0.0 ? v : a.beta * to_d(*y) + v);
    }
  }
}

// -------------------------------------------------------- Cholesky (FP64)
// One workgroup per group: G (w x w, ld S) -> R upper (R^T R = G + shift I),
// Rinv = R^{-1}.  If a pivot is not positive, restart with a shift
// sigma = 11 (n w + w (w+1)) u ||G||_F (Fukaya et al. 2020) and flag it.
__global__ __launch_bounds__(1024) void chol_kernel(CholArgs a, int use_lds) {
  extern __shared__ double sm[];
  const int g = blockIdx.x;
  const int w = a.groups[g].w;
  const int S = a.S;
  // w x w working copy (ld S): in LDS when it fits, else in the R output.
  double *G = use_lds ? sm : a.R + (size_t)g * a.gstride;
  __shared__ int s_fail;
  __shared__ double s_norm;
  const double *Gin = a.G + (size_t)g * a.gstride;
  double *R = a.R + (size_t)g * a.gstride;
  double *Ri = a.Rinv + (size_t)g * a.gstride;
  if (threadIdx.x == 0) s_norm = 0.0;
  __syncthreads();
  double loc = 0.0;
  for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
    const double v = Gin[(e / w) * S + e % w];
    loc += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) loc += __shfl_xor(loc, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&s_norm, loc);
  __syncthreads();
  const double gnorm = sqrt(s_norm);
  double shift = 0.0;
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
      const int r = e / w, c = e % w;
      G[r * S + c] = Gin[r * S + c] + (r == c ? shift : 0.0);
    }
    if (threadIdx.x == 0) s_fail = 0;
    __syncthreads();
    // Right-looking Cholesky, upper factor stored in G's upper triangle.
    for (int k = 0; k < w; ++k) {
      const double d = G[k * S + k];
      if (!(d > 0.0) || !isfinite(d)) {
        if (threadIdx.x == 0) s_fail = 1;
      }
      __syncthreads();
      if (s_fail) break;
      const double rkk = sqrt(d);
      // row k: R[k][j] = G[k][j] / rkk for j > k
      for (int j = k + 1 + threadIdx.x; j < w; j += blockDim.x) G[k * S + j] /= rkk;
      __syncthreads();
      if (threadIdx.x == 0) G[k * S + k] = rkk;
      // trailing update G[i][j] -= R[k][i] R[k][j], i <= j
      const int m = w - k - 1;
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int i = k + 1 + e / m, j = k + 1 + e % m;
        if (j >= i) G[i * S + j] -= G[k * S + i] * G[k * S + j];
      }
      __syncthreads();
    }
    __syncthreads();
    if (!s_fail) break;
    shift = shift == 0.0 ? 11.0 * ((double)a.n * w + (double)w * (w + 1)) * 1.1102230246251565e-16 * gnorm
                         : shift * 100.0;
    if (shift == 0.0) shift = 1e-300;
    if (threadIdx.x == 0) atomicOr(a.flags, 1);
  }
  // R = upper(G); Rinv by column-oriented back substitution, one column per thread.
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) {
    const int r = e / S, c = e % S;
    const double v = (r < w && c < w && c >= r) ? G[r * S + c] : (r == c && r >= w ? 1.0 : 0.0);
    R[r * S + c] = v;
  }
  __syncthreads();
  G = R;
  for (int c = threadIdx.x; c < S; c += blockDim.x) {
    // solve R x = e_c (upper) for x = column c of Rinv
    for (int r = S - 1; r >= 0; --r) {
      double v = (r == c) ? 1.0 : 0.0;
      if (r < w && c < w) {
        for (int j = r + 1; j <= c; ++j) v -= G[r * S + j] * Ri[j * S + c];
        Ri[r * S + c] = (r <= c) ? v / G[r * S + r] : 0.0;
      } else {
        Ri[r * S + c] = (r == c) ? 1.0 : 0.0;
      }
    }
  }
}

// C = A * B for upper-triangular S x S blocks per group (C may not alias).
__global__ __launch_bounds__(kB) void trmm_upper_kernel(const double *A, const double *B, double *C,
                                                        int S, size_t stride, size_t cstride,
                                                        const GroupDesc *groups) {
  const int g = blockIdx.y;
  const int w = groups[g].w;
  const int e = blockIdx.x * kB + threadIdx.x;
  if (e >= S * S) return;
  const int r = e / S, c = e % S;
  const double *a = A + (size_t)g * stride, *b = B + (size_t)g * stride;
  double v = 0.0;
  if (r < w && c < w && c >= r)
    for (int k = r; k <= c; ++k) v += a[r * S + k] * b[k * S + c];
  C[(size_t)g * cstride + e] = v;
}

// ------------------------------------------------------ block Hessenberg
__device__ __forceinline__ double *hblock(double *const *cols, int k, int g, int i, int S) {
  return cols[k] + ((size_t)g * (k + 2) + i) * S * S;
}
__device__ __forceinline__ double *htau(double *const *cols, int k, int g, int ng, int S) {
  return cols[k] + (size_t)ng * (k + 2) * S * S + (size_t)g * S;
}
__device__ __forceinline__ double *hzfin(double *const *cols, int k, int g, int ng, int S) {
  return cols[k] + (size_t)ng * (k + 2) * S * S + (size_t)ng * S + (size_t)g * S * S;
}

// Column b of group g: apply reflector blocks Omega_0..Omega_{k-1} (block i
// is the compact 2S x S Householder factor stored in blocks i, i+1 of block
// column i) to the new block column [Psi_0; ...; Psi_k], in place.
// Column b of the block column is staged in LDS when (k+1)*S doubles fit,
// otherwise updated in place in global memory (stride S).
struct ColView {
  double *base;
  int S;
  bool lds;
  __device__ __forceinline__ double &operator[](int e) const {
    return lds ? base[e] : base[(size_t)(e / S) * S * S + (size_t)(e % S) * S];
  }
};

__global__ __launch_bounds__(kB) void hess_apply_kernel(HessArgs a, int use_lds) {
  extern __shared__ double ys[];      // (k+1)*S entries of column b
  const int b = blockIdx.x, g = blockIdx.y;
  const int S = a.S, k = a.k;
  const int w = a.groups[g].w;
  if (b >= w) return;
  double *hcol = hblock(a.cols, k, g, 0, S);
  if (use_lds) {
    for (int e = threadIdx.x; e < (k + 1) * S; e += kB) ys[e] = hcol[(size_t)(e / S) * S * S + (e % S) * S + b];
  }
  __syncthreads();
  const ColView yv{use_lds ? ys : hcol + b, S, use_lds != 0};
  __shared__ double red[kB / 64];
  for (int i = 0; i < k; ++i) {
    const double *H = hblock(a.cols, i, g, i, S);   // 2S x S row-major
    const double *tau = htau(a.cols, i, g, a.ng, S);
    const int off = i * S;
    for (int j = 0; j < w; ++j) {
      // v = [0..0, 1, H[j+1..2S-1][j]]; w = tau_j * v^T seg
      double part = 0.0;
      for (int r = j + threadIdx.x; r < 2 * S; r += kB) {
        const double v = (r == j) ? 1.0 : H[(size_t)r * S + j];
        part += v * yv[off + r];
      }
      for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
      __syncthreads();
      const double dot = red[0] + red[1] + red[2] + red[3];
      const double t = tau[j] * dot;
      for (int r = j + threadIdx.x; r < 2 * S; r += kB) {
        const double v = (r == j) ? 1.0 : H[(size_t)r * S + j];
        yv[off + r] -= t * v;
      }
      __syncthreads();
    }
  }
  if (use_lds)
    for (int e = threadIdx.x; e < (k + 1) * S; e += kB) hcol[(size_t)(e / S) * S * S + (e % S) * S + b] = ys[e];
}

// One workgroup per group: D = [Psi_kk; C] (2S x S, row-major, in a.D) is
// factored in place by Householder QR (LAPACK geqrf convention: reflector j
// is v = e_j + D[j+1.., j], H = I - tau v v^T); the reflectors are applied
// to E = [Z_k; 0] (2S x S) as they are formed.  Padding columns (>= w) are
// left as identity.  R_kk = upper(D[0:S]), Z_k = E[0:S], Lambda = E[S:2S].
__global__ __launch_bounds__(1024) void hess_qr_kernel(HessQrArgs a) {
  const int g = blockIdx.x;
  const int S = a.S;
  const int w = a.groups[g].w;
  double *D = hblock(a.cols, a.k, g, a.k, S);
  double *E = a.E + (size_t)g * 2 * S * S;
  double *tau = htau(a.cols, a.k, g, a.ng, S);
  __shared__ double red[16];
  __shared__ double s_beta, s_tau, s_scale;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = 0; j < w; ++j) {
    // norm of D[j+1.., j]
    double part = 0.0;
    for (int r = j + 1 + threadIdx.x; r < 2 * S; r += blockDim.x) {
      const double v = D[(size_t)r * S + j];
      part += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == 0) red[wid] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sig = 0.0;
      for (int q = 0; q < nw; ++q) sig += red[q];
      const double alpha = D[(size_t)j * S + j];
      if (sig == 0.0) {
        s_tau = 0.0; s_beta = alpha; s_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + sig);
        const double beta = alpha <= 0.0 ? nrm : -nrm;
        s_tau = (beta - alpha) / beta;
        s_scale = 1.0 / (alpha - beta);
        s_beta = beta;
      }
      tau[j] = s_tau;
    }
    __syncthreads();
    for (int r = j + 1 + threadIdx.x; r < 2 * S; r += blockDim.x) D[(size_t)r * S + j] *= s_scale;
    __syncthreads();
    if (threadIdx.x == 0) D[(size_t)j * S + j] = s_beta;
    const double t = s_tau;
    // apply to remaining columns of D (c > j, c < w) and all w columns of E:
    // one wave per column.
    const int ncol = (w - j - 1) + w;
    for (int cc = wid; cc < ncol; cc += nw) {
      double *M = cc < (w - j - 1) ? D : E;
      const int c = cc < (w - j - 1) ? j + 1 + cc : cc - (w - j - 1);
      double dot = 0.0;
      for (int r = j + lane; r < 2 * S; r += 64) {
        const double v = (r == j) ? 1.0 : D[(size_t)r * S + j];
        dot += v * M[(size_t)r * S + c];
      }
      for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
      const double f = t * dot;
      for (int r = j + lane; r < 2 * S; r += 64) {
        const double v = (r == j) ? 1.0 : D[(size_t)r * S + j];
        M[(size_t)r * S + c] -= f * v;
      }
    }
    __syncthreads();
  }
  for (int j = w + threadIdx.x; j < S; j += blockDim.x) tau[j] = 0.0;
  // residual norms of the new Lambda = E[S:2S, :w] per column
  for (int c = wid; c < w; c += nw) {
    double s = 0.0;
    for (int r = S + lane; r < 2 * S; r += 64) {
      const double v = E[(size_t)r * S + c];
      s += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) a.rnorm[a.groups[g].c0 + c] = sqrt(s);
  }
  __syncthreads();
  // zfin(k) = E[0:S]; E[0:S] = E[S:2S] (next right-hand side); E[S:2S] = 0.
  double *zf = hzfin(a.cols, a.k, g, a.ng, S);
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) {
    zf[e] = E[e];
    E[e] = E[S * S + e];
    E[S * S + e] = 0.0;
  }
}

// ---------------------------------------------------- back substitution
// Column b of group g: solve the block upper-triangular system with blocks
// Rcol(j)[i] (block i of block column j, S x S row-major) for y = Zrhs(:, b).
struct YView {
  double *base;   // LDS array, or &Y[g-block 0][0][b]
  int S, ng;
  bool lds;
  __device__ __forceinline__ double &operator[](int e) const {
    return lds ? base[e] : base[((size_t)(e / S) * ng) * S * S + (size_t)(e % S) * S];
  }
};

__global__ __launch_bounds__(kB) void backsub_kernel(BacksubArgs a, int use_lds) {
  extern __shared__ double ys[];     // K*S
  const int b = blockIdx.x, g = blockIdx.y;
  const int S = a.S, K = a.K;
  const int w = a.groups[g].w;
  if (b >= w) return;
  const YView y{use_lds ? ys : a.Y + (size_t)g * S * S + b, S, a.ng, use_lds != 0};
  for (int e = threadIdx.x; e < K * S; e += kB) {
    const int i = e / S, r = e % S;
    y[e] = hzfin(a.cols, i, g, a.ng, S)[(size_t)r * S + b];
  }
  __syncthreads();
  for (int j = K - 1; j >= 0; --j) {
    const double *Rjj = hblock(a.cols, j, g, j, S);
    // dense upper-triangular solve of block j (sequential over rows, w entries)
    for (int r = w - 1; r >= 0; --r) {
      __syncthreads();
      if (threadIdx.x == 0) y[j * S + r] /= Rjj[(size_t)r * S + r];
      __syncthreads();
      const double yr = y[j * S + r];
      for (int q = threadIdx.x; q < r; q += kB) y[j * S + q] -= Rjj[(size_t)q * S + r] * yr;
    }
    __syncthreads();
    // y_i -= R_ij y_j for i < j
    for (int e = threadIdx.x; e < j * S; e += kB) {
      const int i = e / S, r = e % S;
      if (r >= w) continue;
      const double *Rij = hblock(a.cols, j, g, i, S);
      double v = 0.0;
      for (int q = 0; q < w; ++q) v += Rij[(size_t)r * S + q] * y[j * S + q];
      y[e] -= v;
    }
    __syncthreads();
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * S; e += kB) {
    const int i = e / S, r = e % S;
    const double v = (r < w) ? y[e] : 0.0;
    if (use_lds || r >= w) a.Y[((size_t)i * a.ng + g) * S * S + (size_t)r * S + b] = v;
  }
}

// ------------------------------------------------------------ statistics
// Column-major n x p: out[c] = max_i |X[i + c*ld]|  (one workgroup per column)
template <typename T>
__global__ __launch_bounds__(kB) void colmax_cm_kernel(int n, const T *X, int ld, double *out) {
  const int c = blockIdx.x;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += kB) m = fmax(m, fabs(to_d(X[(size_t)c * ld + i])));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[kB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[c] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Row-major n x p: per-column inf-norm (mode 0) or sum of squares (mode 1),
// accumulated with atomics into out[p] (must be zeroed by the caller).
template <typename T>
__global__ __launch_bounds__(kB) void colstat_rm_kernel(int n, int p, const T *X, int ld, int mode,
                                                        double *out) {
  const int c = blockIdx.y * 64 + (threadIdx.x & 63);
  const int rstride = (kB / 64) * gridDim.x;
  double m = 0.0;
  if (c < p) {
    for (int r = blockIdx.x * (kB / 64) + (threadIdx.x >> 6); r < n; r += rstride) {
      const double v = to_d(X[(size_t)r * ld + c]);
      m = mode == 0 ? fmax(m, fabs(v)) : m + v * v;
    }
    if (mode == 0) atomic_max_nonneg(out + c, m);
    else atomicAdd(out + c, m);
  }
}

// R = B - A*X for row-major FP64 blocks (residual in working precision).
__global__ __launch_bounds__(kB) void axpby_kernel(size_t m, double alpha, const double *x,
                                                   double beta, double *y) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i < m) y[i] = alpha * x[i] + beta * y[i];
}

template <typename TS, typename TD>
__global__ __launch_bounds__(kB) void copy_cast_kernel(int n, int p, const TS *src, int lds, TD *dst,
                                                      int ldd, double scale) {
  const size_t e = (size_t)blockIdx.x * kB + threadIdx.x;
  if (e >= (size_t)n * p) return;
  const int r = (int)(e / p), c = (int)(e % p);
  dst[(size_t)r * ldd + c] = static_cast<TD>(scale * to_d(src[(size_t)r * lds + c]));
}

}  // namespace

// ----------------------------------------------------------------- hosts
void csr_norm_and_cast(hipStream_t s, int n, const int *rowptr, const double *val, float *val32,
                       int, double *d_out) {
  hipLaunchKernelGGL((fill_kernel<double>), dim3(1), dim3(kB), 0, s, d_out, 0.0, (size_t)1);
  if (n > 0)
    hipLaunchKernelGGL(csr_norm_cast_kernel, dim3((n + kB - 1) / kB), dim3(kB), 0, s, n, rowptr, val,
                       val32, d_out);
  BPF_CHECK_HIP(hipGetLastError());
}

void fill_int(hipStream_t s, int *p, int v, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_int_kernel, dim3((n + kB - 1) / kB), dim3(kB), 0, s, p, v, n);
  BPF_CHECK_HIP(hipGetLastError());
}

void fill_double(hipStream_t s, double *p, double v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL((fill_kernel<double>), dim3((unsigned)((n + kB - 1) / kB)), dim3(kB), 0, s, p, v, n);
  BPF_CHECK_HIP(hipGetLastError());
}

void fill_float(hipStream_t s, float *p, float v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL((fill_kernel<float>), dim3((unsigned)((n + kB - 1) / kB)), dim3(kB), 0, s, p, v, n);
  BPF_CHECK_HIP(hipGetLastError());
}

int pad_width(int w) {
  if (w <= 16) return 16;
  if (w <= 32) return 32;
  if (w <= 64) return 64;
  if (w <= 128) return 128;
  return -1;
}

template <typename TVV, typename TW>
void launch_proj(hipStream_t s, ProjArgs a, int S, double *out, size_t o_istride, size_t o_gstride,
                 bool accumulate, DevBuf &part) {
  if (a.nb <= 0 || a.ng <= 0) return;
  // Chunk rows so that the grid has >= ~2048 workgroups.
  int target = std::max(1, 2048 / std::max(1, a.nb * a.ng));
  int nchunk = std::max(1, std::min(target, (a.n + 255) / 256));
  a.chunk_rows = (a.n + nchunk - 1) / nchunk;
  a.chunk_rows = ((a.chunk_rows + 15) / 16) * 16;
  nchunk = std::max(1, (a.n + a.chunk_rows - 1) / a.chunk_rows);
  part.ensure((size_t)nchunk * a.nb * a.ng * S * S * sizeof(double));
  a.part = part.as<double>();
  dim3 grid(a.nb, nchunk, a.ng);
  const int tok = prof::begin(s);
  switch (S) {
    case 16: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 16>), grid, dim3(kB), 0, s, a); break;
    case 32: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 32>), grid, dim3(kB), 0, s, a); break;
    case 64: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 64>), grid, dim3(kB), 0, s, a); break;
    case 128: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 128>), grid, dim3(kB), 0, s, a); break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", S);
  }
  const size_t total = (size_t)a.nb * a.ng * S * S;
  hipLaunchKernelGGL(proj_reduce_kernel, dim3((unsigned)((total + kB - 1) / kB)), dim3(kB), 0, s,
                     a.part, nchunk, a.nb, a.ng, S * S, out, o_istride, o_gstride,
                     accumulate ? 1 : 0);
  BPF_CHECK_HIP(hipGetLastError());
  // Algorithmic traffic: each basis block and W read once; the s x s
  // products are 2 n w^2 flops per (block, group).
  const double cols = (double)a.ng * S;   // upper bound on sum of group widths
  prof::end(tok, s, prof::PROJ, (double)a.n * cols * ((a.nb) * sizeof(TVV) + sizeof(TW)),
            2.0 * a.n * (double)S * S * a.nb * a.ng);
}
template void launch_proj<double, double>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);
template void launch_proj<float, float>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);
template void launch_proj<float, double>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);

template <typename TVV, typename TY>
void launch_block_gemm(hipStream_t s, GemmArgs a, int S, int ng) {
  if (a.n <= 0 || ng <= 0) return;
  dim3 grid((a.n + 63) / 64, ng);
  if (a.nb <= 0) {
    if (a.beta == 1.0) return;
  }
  const int tok = prof::begin(s);
  switch (S) {
    case 16: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 16>), grid, dim3(kB), 0, s, a); break;
    case 32: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 32>), grid, dim3(kB), 0, s, a); break;
    case 64: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 64>), grid, dim3(kB), 0, s, a); break;
    case 128: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 128>), grid, dim3(kB), 0, s, a); break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", S);
  }
  BPF_CHECK_HIP(hipGetLastError());
  const double cols = (double)ng * S;
  prof::end(tok, s, prof::GEMM,
            (double)a.n * cols * (a.nb * sizeof(TVV) + (a.beta == 0.0 ? 1.0 : 2.0) * sizeof(TY)),
            2.0 * a.n * (double)S * S * a.nb * ng);
}
template void launch_block_gemm<double, double>(hipStream_t, GemmArgs, int, int);
template void launch_block_gemm<float, float>(hipStream_t, GemmArgs, int, int);
template void launch_block_gemm<float, double>(hipStream_t, GemmArgs, int, int);


void launch_chol(hipStream_t s, CholArgs a, int ng) {
  const size_t shm = (size_t)a.S * a.S * sizeof(double);
  const int lds = shm <= kLdsLimit ? 1 : 0;
  hipLaunchKernelGGL(chol_kernel, dim3(ng), dim3(1024), lds ? shm : 0, s, a, lds);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_trmm_upper(hipStream_t s, const double *A, const double *B, double *C, int S,
                       size_t stride, size_t cstride, const GroupDesc *groups, int ng) {
  dim3 grid((S * S + kB - 1) / kB, ng);
  hipLaunchKernelGGL(trmm_upper_kernel, grid, dim3(kB), 0, s, A, B, C, S, stride, cstride, groups);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_hess_apply(hipStream_t s, HessArgs a, int ng) {
  if (a.k <= 0) return;
  const size_t shm = (size_t)(a.k + 1) * a.S * sizeof(double);
  const int lds = shm <= kLdsLimit ? 1 : 0;
  hipLaunchKernelGGL(hess_apply_kernel, dim3(a.S, ng), dim3(kB), lds ? shm : 0, s, a, lds);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_hess_qr(hipStream_t s, HessQrArgs a, int ng) {
  hipLaunchKernelGGL(hess_qr_kernel, dim3(ng), dim3(1024), 0, s, a);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_backsub(hipStream_t s, BacksubArgs a, int ng) {
  if (a.K <= 0) return;
  const size_t shm = (size_t)a.K * a.S * sizeof(double);
  const int lds = shm <= kLdsLimit ? 1 : 0;
  hipLaunchKernelGGL(backsub_kernel, dim3(a.S, ng), dim3(kB), lds ? shm : 0, s, a, lds);
  BPF_CHECK_HIP(hipGetLastError());
}

void colmax_colmajor(hipStream_t s, int n, int p, const double *X, int ld, double *out) {
  if (p <= 0) return;
  hipLaunchKernelGGL((colmax_cm_kernel<double>), dim3(p), dim3(kB), 0, s, n, X, ld, out);
  BPF_CHECK_HIP(hipGetLastError());
}

template <typename T>
void colstat_rowmajor(hipStream_t s, int n, int p, const T *X, int ld, int mode, double *out) {
  if (p <= 0) return;
  fill_double(s, out, 0.0, p);
  if (n <= 0) return;
  const int gx = std::max(1, std::min(1024, (n + 255) / 256));
  dim3 grid(gx, (p + 63) / 64);
  hipLaunchKernelGGL((colstat_rm_kernel<T>), grid, dim3(kB), 0, s, n, p, X, ld, mode, out);
  BPF_CHECK_HIP(hipGetLastError());
}
template void colstat_rowmajor<double>(hipStream_t, int, int, const double *, int, int, double *);
template void colstat_rowmajor<float>(hipStream_t, int, int, const float *, int, int, double *);

void axpby(hipStream_t s, size_t m, double alpha, const double *x, double beta, double *y) {
  if (m == 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((m + kB - 1) / kB)), dim3(kB), 0, s, m, alpha, x, beta, y);
  BPF_CHECK_HIP(hipGetLastError());
}

template <typename TS, typename TD>
void copy_cast(hipStream_t s, int n, int p, const TS *src, int lds, TD *dst, int ldd, double scale) {
  const size_t m = (size_t)n * p;
  if (m == 0) return;
  hipLaunchKernelGGL((copy_cast_kernel<TS, TD>), dim3((unsigned)((m + kB - 1) / kB)), dim3(kB), 0, s, n,
                     p, src, lds, dst, ldd, scale);
  BPF_CHECK_HIP(hipGetLastError());
}
template void copy_cast<double, double>(hipStream_t, int, int, const double *, int, double *, int, double);
template void copy_cast<double, float>(hipStream_t, int, int, const double *, int, float *, int, double);
template void copy_cast<float, double>(hipStream_t, int, int, const float *, int, double *, int, double);
template void copy_cast<float, float>(hipStream_t, int, int, const float *, int, float *, int, double);

}  // namespace bpf

namespace bpf {
namespace {

__global__ __launch_bounds__(256) void scale_ycols_kernel(double *Y, int K, int ng, int S,
                                                         const GroupDesc *groups, const double *scale) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t SS = (size_t)S * S;
  if (e >= (size_t)K * ng * SS) return;
  const int g = (int)((e / SS) % ng);
  const int b = (int)(e % S);
  if (b < groups[g].w) Y[e] *= scale[groups[g].c0 + b];
}

__global__ __launch_bounds__(256) void scale_cols_kernel(int n, int p, double *X, int ld,
                                                        const double *scale, int invert) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * p) return;
  const int r = (int)(e / p), c = (int)(e % p);
  const double sc = scale[c];
  X[(size_t)r * ld + c] = invert ? X[(size_t)r * ld + c] / sc : X[(size_t)r * ld + c] * sc;
}

__global__ __launch_bounds__(1024) void lu_small_kernel(double *A0, int *piv0, int S,
                                                       const GroupDesc *groups) {
  const int g = blockIdx.x;
  const int w = groups[g].w;
  double *A = A0 + (size_t)g * S * S;
  int *piv = piv0 + (size_t)g * S;
  __shared__ double sv[16];
  __shared__ int si[16];
  __shared__ int s_p;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < w; ++k) {
    double best = -1.0;
    int bi = k;
    for (int i = k + threadIdx.x; i < w; i += blockDim.x) {
      const double v = fabs(A[(size_t)i * S + k]);
      if (v > best) { best = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { sv[wid] = best; si[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = -1.0;
      int ii = k;
      for (int q = 0; q < nw; ++q)
        if (sv[q] > b || (sv[q] == b && si[q] < ii)) { b = sv[q]; ii = si[q]; }
      s_p = ii;
      piv[k] = ii;
    }
    __syncthreads();
    const int pr = s_p;
    if (pr != k)
      for (int j = threadIdx.x; j < w; j += blockDim.x) {
        const double t = A[(size_t)k * S + j];
        A[(size_t)k * S + j] = A[(size_t)pr * S + j];
        A[(size_t)pr * S + j] = t;
      }
    __syncthreads();
    const double d = A[(size_t)k * S + k];
    for (int i = k + 1 + threadIdx.x; i < w; i += blockDim.x) A[(size_t)i * S + k] = d != 0.0 ? A[(size_t)i * S + k] / d : 0.0;
    __syncthreads();
    const int m = w - k - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      A[(size_t)i * S + j] -= A[(size_t)i * S + k] * A[(size_t)k * S + j];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void lu_small_solve_kernel(const double *A0, const int *piv0,
                                                            double *B0, int S, const GroupDesc *groups) {
  const int g = blockIdx.y;
  const int w = groups[g].w;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= w) return;
  const double *A = A0 + (size_t)g * S * S;
  const int *piv = piv0 + (size_t)g * S;
  double *B = B0 + (size_t)g * S * S;
  for (int k = 0; k < w; ++k) {
    const int pr = piv[k];
    if (pr != k) {
      const double t = B[(size_t)k * S + c];
      B[(size_t)k * S + c] = B[(size_t)pr * S + c];
      B[(size_t)pr * S + c] = t;
    }
  }
  for (int i = 0; i < w; ++i) {
    double v = B[(size_t)i * S + c];
    for (int j = 0; j < i; ++j) v -= A[(size_t)i * S + j] * B[(size_t)j * S + c];
    B[(size_t)i * S + c] = v;
  }
  for (int i = w - 1; i >= 0; --i) {
    double v = B[(size_t)i * S + c];
    for (int j = i + 1; j < w; ++j) v -= A[(size_t)i * S + j] * B[(size_t)j * S + c];
    B[(size_t)i * S + c] = v / A[(size_t)i * S + i];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void dot_kernel(size_t m, const T *x, const T *y, double *out) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (size_t)gridDim.x * 256)
    s += static_cast<double>(x[i]) * static_cast<double>(y[i]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

}  // namespace

void launch_scale_ycols(hipStream_t s, double *Y, int K, int ng, int S, const GroupDesc *groups,
                        const double *scale) {
  const size_t m = (size_t)K * ng * S * S;
  if (m == 0) return;
  hipLaunchKernelGGL(scale_ycols_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, Y, K, ng,
                     S, groups, scale);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_scale_cols(hipStream_t s, int n, int p, double *X, int ld, const double *scale, bool invert) {
  const size_t m = (size_t)n * p;
  if (m == 0) return;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, n, p, X, ld,
                     scale, invert ? 1 : 0);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_lu_small(hipStream_t s, double *A, int *piv, int S, const GroupDesc *groups, int ng) {
  hipLaunchKernelGGL(lu_small_kernel, dim3(ng), dim3(1024), 0, s, A, piv, S, groups);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_lu_small_solve(hipStream_t s, const double *A, const int *piv, double *B, int S,
                           const GroupDesc *groups, int ng) {
  hipLaunchKernelGGL(lu_small_solve_kernel, dim3((S + 255) / 256, ng), dim3(256), 0, s, A, piv, B, S,
                     groups);
  BPF_CHECK_HIP(hipGetLastError());
}

template <typename T>
void launch_dot(hipStream_t s, size_t m, const T *x, const T *y, double *out) {
  fill_double(s, out, 0.0, 1);
  if (m == 0) return;
  const unsigned nb = (unsigned)std::min<size_t>(2048, (m + 255) / 256);
  hipLaunchKernelGGL((dot_kernel<T>), dim3(nb), dim3(256), 0, s, m, x, y, out);
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_dot<double>(hipStream_t, size_t, const double *, const double *, double *);
template void launch_dot<float>(hipStream_t, size_t, const float *, const float *, double *);

}  // namespace bpf

namespace bpf {
namespace {
template <typename T>
__global__ __launch_bounds__(256) void lincomb_kernel(size_t m, T alpha, const T *x, T beta, T *y) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < m) y[i] = beta == T(0) ? alpha * x[i] : fma(alpha, x[i], beta * y[i]);
}
}  // namespace

template <typename T>
void lincomb(hipStream_t s, size_t m, double alpha, const T *x, double beta, T *y) {
  if (m == 0) return;
  hipLaunchKernelGGL((lincomb_kernel<T>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                     static_cast<T>(alpha), x, static_cast<T>(beta), y);
  BPF_CHECK_HIP(hipGetLastError());
}
template void lincomb<float>(hipStream_t, size_t, double, const float *, double, float *);
template void lincomb<double>(hipStream_t, size_t, double, const double *, double, double *);

}  // namespace bpf
