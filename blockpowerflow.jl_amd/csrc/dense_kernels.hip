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
//   hess_qr_kernel              Householder QR of [Psi_kk; C], compact-WY factor T
//   hess_rot_kernel             rotation of [Z_k; 0], residual estimates
//   backsub_kernel              block upper-triangular solve R Y = Z
//   colstats kernels            per-column inf-norms / 2-norms, backward error
#include "dense.h"
#include "profile.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

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
//
// Matrix cores: v_mfma_f64_16x16x4f64 computes a 16x16 tile D += A(16x4) B(4x16)
// per wavefront.  Lane l supplies A[l%16][l/16] and B[l/16][l%16]; the result
// element i of lane l is D[l/16 + 4i][l%16].  Here A = V_i^T (a x r) and
// B = W (r x b), so each lane reads V[r0 + l/16][a0 + l%16] and
// W[r0 + l/16][b0 + l%16] from 16-row tiles staged in LDS.  The S/16 x S/16
// output tiles are spread over the 4 wavefronts (a 2x2 grid of tile blocks);
// for S < 64 the wavefronts split the rows instead and are summed in LDS.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Workgroups are handed to the 8 XCDs of an MI355X in turn (workgroup w runs
// on XCD w % 8, each with its own L2).  Narrow groups share the cache lines
// of a row (a 128-byte line holds 32 FP32 or 16 FP64 columns), so all groups
// of one row chunk are placed on one XCD: chunk c on XCD c % 8, its groups
// in consecutive turns.  Grids are padded to a multiple of 8 chunks.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_chunk_group(int ng, int &chunk, int &g) {
  const int xcd = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
  g = j % ng;
  chunk = (j / ng) * kXcds + xcd;
}
inline unsigned xcd_grid(long nchunk, int ng) {
  const long padded = (nchunk + kXcds - 1) / kXcds * kXcds;
  return (unsigned)(padded * ng);
}

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 acc) {
  // D(16x16) += A(16x4) B(4x16): lane l supplies A[l%16][l/16] and
  // B[l/16][l%16] and holds D[l/16 + 4q][l%16] in acc[q].
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

template <int S>
struct TnShape {
  static constexpr int T = S / 16;                 // tiles per dimension
  static constexpr int WA = T >= 2 ? 2 : 1;        // wave grid over a-tiles
  static constexpr int WB = T >= 2 ? 2 : 1;        // ... and b-tiles
  static constexpr int KS = 4 / (WA * WB);         // waves splitting the rows
  static constexpr int TA = T / WA, TB = T / WB;   // tiles per wave
};

template <typename TVV, typename TW, int S>
__global__ __launch_bounds__(kB) void proj_partial_kernel(ProjArgs a) {
  using Sh = TnShape<S>;
  constexpr int RB = 16;                           // rows staged per round
  __shared__ double sv[RB][S + 2];
  __shared__ double sw[RB][S + 2];
  const int i = blockIdx.x;
  const int chunk = blockIdx.y;
  const int g = blockIdx.z;
  const GroupDesc gd = a.groups[g];
  const TVV *V = static_cast<const TVV *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0;
  const TW *W = static_cast<const TW *>(a.W) + gd.c0;
  const int r0 = gd.r0 + chunk * a.chunk_rows;
  const int r1 = gd.active ? min(gd.r0 + gd.nr, r0 + a.chunk_rows) : r0;
  // Stacked groups: the reduction reads only the chunks that hold rows.
  if (a.rows_sum > 0 && r0 >= r1) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ks = wave % Sh::KS;                    // row split index
  const int wg = wave / Sh::KS;                    // tile-block index
  const int wa = wg % Sh::WA, wb = wg / Sh::WA;
  dbl4 acc[Sh::TA][Sh::TB];
#pragma unroll
  for (int x = 0; x < Sh::TA; ++x)
#pragma unroll
    for (int y = 0; y < Sh::TB; ++y) acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int lr = lane >> 4, lc = lane & 15;
  // Register prefetch of the next 16-row tile while the current one is used.
  constexpr int PE = RB * S / kB;   // elements per thread per tile (>= 1)
  double pv[PE], pw[PE];
  auto fetch = [&](int rb) {
#pragma unroll
    for (int q = 0; q < PE; ++q) {
      const int e = threadIdx.x + q * kB;
      const int rr = e / S, cc = e % S;
      const int r = rb + rr;
      const bool ok = r < r1 && cc < gd.w;
      pv[q] = ok ? to_d(V[(size_t)r * a.ldv + cc]) : 0.0;
      pw[q] = ok ? to_d(W[(size_t)r * a.ldw + cc]) : 0.0;
    }
  };
  if (r0 < r1) fetch(r0);
  for (int rb = r0; rb < r1; rb += RB) {
#pragma unroll
    for (int q = 0; q < PE; ++q) {
      const int e = threadIdx.x + q * kB;
      sv[e / S][e % S] = pv[q];
      sw[e / S][e % S] = pw[q];
    }
    __syncthreads();
    if (rb + RB < r1) fetch(rb + RB);
#pragma unroll
    for (int k4 = ks; k4 < RB / 4; k4 += Sh::KS) {
      const int rr = k4 * 4 + lr;
      double fa[Sh::TA], fb[Sh::TB];
#pragma unroll
      for (int x = 0; x < Sh::TA; ++x) fa[x] = sv[rr][(wa * Sh::TA + x) * 16 + lc];
#pragma unroll
      for (int y = 0; y < Sh::TB; ++y) fb[y] = sw[rr][(wb * Sh::TB + y) * 16 + lc];
#pragma unroll
      for (int x = 0; x < Sh::TA; ++x)
#pragma unroll
        for (int y = 0; y < Sh::TB; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  double *out = a.part + ((size_t)(chunk * a.nb + i) * a.ng + g) * S * S;
  if (Sh::KS == 1) {
#pragma unroll
    for (int x = 0; x < Sh::TA; ++x)
#pragma unroll
      for (int y = 0; y < Sh::TB; ++y)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ra = (wa * Sh::TA + x) * 16 + lr + 4 * q;
          const int cb = (wb * Sh::TB + y) * 16 + lc;
          out[ra * S + cb] = acc[x][y][q];
        }
  } else {
    // Sum the row-split partial tiles of the KS wavefronts through LDS.
    double *red = &sv[0][0];   // S*S <= RB*(S+2) for S <= 32
    for (int e = threadIdx.x; e < S * S; e += kB) red[e] = 0.0;
    __syncthreads();
    for (int w = 0; w < Sh::KS; ++w) {
      if (ks == w) {
#pragma unroll
        for (int x = 0; x < Sh::TA; ++x)
#pragma unroll
          for (int y = 0; y < Sh::TB; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int ra = (wa * Sh::TA + x) * 16 + lr + 4 * q;
              const int cb = (wb * Sh::TB + y) * 16 + lc;
              red[ra * S + cb] += acc[x][y][q];
            }
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < S * S; e += kB) out[e] = red[e];
  }
}

// Psi_i = V_i^T W for width-16 groups (S = 16).  The matrix-core fragments
// are read straight from global memory: lane l holds row r + l/16, column
// l%16 of both V_i and W (16 consecutive values per row), so there is no LDS
// staging and no barrier.  A workgroup owns one row chunk of one group and
// all nb blocks: wavefront w takes blocks w, w + 4, w + 8, ... up to
// kNarrowBatch at a time, so each wavefront reads the chunk of W once per
// batch.  Loads are unconditional (rows past the chunk re-read its last row,
// and their W values are zeroed), the block pointers are wavefront-uniform
// scalars, and the number of live blocks is a template parameter, so the
// loop holds no per-load branches.  Partial sums have the layout of
// proj_partial_kernel's.
constexpr int kNarrowBatch = 4;

template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *global_ptr(const void *p) {
  // The table entries are wavefront-uniform: keep them in scalar registers.
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
  return reinterpret_cast<const __attribute__((address_space(1))) T *>(
      (static_cast<unsigned long long>(hi) << 32) | lo);
}

template <int NV, typename TVV, typename TW>
__device__ __forceinline__ void proj16_rows(const __attribute__((address_space(1))) TVV *const (&V)[kNarrowBatch],
                                            const __attribute__((address_space(1))) TW *W, int ldv, int ldw,
                                            int r0, int r1, int lr, int lc, bool colok, dbl4 (&acc)[kNarrowBatch]) {
  constexpr int RU = 4;   // 4-row MFMA steps per round
  const int lcl = colok ? lc : 0;
  for (int rr = r0; rr < r1; rr += 4 * RU) {
    double bw[RU], av[RU][NV];
#pragma unroll
    for (int t = 0; t < RU; ++t) {
      const int r = rr + 4 * t + lr;
      const int rc = min(r, r1 - 1);
      // 32-bit element offsets (blocks hold fewer than 2^32 values): scalar
      // base + vector offset addressing.  Lanes past the group's width (a
      // narrower group padded to 16) read column 0 and are zeroed below.
      const unsigned ow = static_cast<unsigned>(rc) * static_cast<unsigned>(ldw) + lcl;
      const unsigned ov = static_cast<unsigned>(rc) * static_cast<unsigned>(ldv) + lcl;
      const double w = to_d(W[ow]);
      bw[t] = (colok && r < r1) ? w : 0.0;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const double v = to_d(V[u][ov]);
        av[t][u] = colok ? v : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < RU; ++t)
#pragma unroll
      for (int u = 0; u < NV; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][u], bw[t], acc[u], 0, 0, 0);
  }
}

// The same on FP32 matrix cores for an FP32 basis (ProjArgs::f32acc): each
// chunk's sums accumulate in FP32 (v_mfma_f32_16x16x4f32, result in
// acc[q] = D[4 (l/16) + q][l%16]) and the chunks are added in FP64.
template <int NV>
__device__ __forceinline__ void proj16_rows_f32(const __attribute__((address_space(1))) float *const (&V)[kNarrowBatch],
                                                const __attribute__((address_space(1))) float *W, int ldv, int ldw,
                                                int r0, int r1, int lr, int lc, bool colok,
                                                flt4 (&acc)[kNarrowBatch]) {
  constexpr int RU = 4;
  const int lcl = colok ? lc : 0;
  // The next round's loads are in flight during this round's MFMAs (masks
  // applied at use, loads at clamped rows).
  auto load = [&](int rr, float (&b)[RU], float (&v)[RU][NV]) {
#pragma unroll
    for (int t = 0; t < RU; ++t) {
      const int rc = min(rr + 4 * t + lr, r1 - 1);
      const unsigned ow = static_cast<unsigned>(rc) * static_cast<unsigned>(ldw) + lcl;
      const unsigned ov = static_cast<unsigned>(rc) * static_cast<unsigned>(ldv) + lcl;
      b[t] = W[ow];
#pragma unroll
      for (int u = 0; u < NV; ++u) v[t][u] = V[u][ov];
    }
  };
  float bw[RU], av[RU][NV];
  if (r0 < r1) load(r0, bw, av);
  for (int rr = r0; rr < r1; rr += 4 * RU) {
    const bool more = rr + 4 * RU < r1;
    float nbw[RU], nav[RU][NV];
    if (more) load(rr + 4 * RU, nbw, nav);
#pragma unroll
    for (int t = 0; t < RU; ++t) {
      const float b = (colok && rr + 4 * t + lr < r1) ? bw[t] : 0.0f;
#pragma unroll
      for (int u = 0; u < NV; ++u)
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(colok ? av[t][u] : 0.0f, b, acc[u], 0, 0, 0);
    }
    if (more) {
#pragma unroll
      for (int t = 0; t < RU; ++t) {
        bw[t] = nbw[t];
#pragma unroll
        for (int u = 0; u < NV; ++u) av[t][u] = nav[t][u];
      }
    }
  }
}

// The occupancy hint keeps the kernel at 61-67 VGPRs and no AGPRs (7-8
// wavefronts per SIMD); without it the compiler parks the accumulators in
// AGPRs and settles at 4.
template <typename TVV, typename TW, bool F32 = false>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(6))) void proj16_kernel(ProjArgs a) {
  // one wavefront per block slot (up to 4; fewer blocks, fewer wavefronts
  // and shorter chunks, so that every call keeps as many rows in flight)
  constexpr int NB = kNarrowBatch;
  const int NW = blockDim.x >> 6;
  int g, chunk;
  xcd_chunk_group(a.ng, chunk, g);
  if (chunk * a.chunk_rows >= a.n) return;   // grid padding
  const GroupDesc gd = a.groups[g];
  const int r0 = gd.r0 + chunk * a.chunk_rows;
  const int r1 = gd.active ? min(gd.r0 + gd.nr, r0 + a.chunk_rows) : r0;
  if (a.rows_sum > 0 && r0 >= r1) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane >> 4, lc = lane & 15;
  const bool colok = lc < gd.w;
  const auto *W = global_ptr<TW>(static_cast<const TW *>(a.W) + gd.c0);
  for (int ib = wave; ib < a.nb; ib += NW * NB) {
    const int nv = min(NB, (a.nb - ib + NW - 1) / NW);   // live blocks ib, ib + NW, ...
    const __attribute__((address_space(1))) TVV *V[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int i = min(ib + u * NW, a.nb - 1);
      V[u] = global_ptr<TVV>(static_cast<const TVV *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0);
    }
    if constexpr (F32) {
      flt4 acc[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) acc[u] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
      switch (nv) {
        case 4: proj16_rows_f32<4>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        case 3: proj16_rows_f32<3>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        case 2: proj16_rows_f32<2>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        default: proj16_rows_f32<1>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        if (u >= nv) break;
        double *out = a.part + ((size_t)(chunk * a.nb + ib + u * NW) * a.ng + g) * 256;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(4 * lr + q) * 16 + lc] = acc[u][q];
      }
    } else {
      dbl4 acc[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) acc[u] = dbl4{0.0, 0.0, 0.0, 0.0};
      switch (nv) {
        case 4: proj16_rows<4>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        case 3: proj16_rows<3>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        case 2: proj16_rows<2>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
        default: proj16_rows<1>(V, W, a.ldv, a.ldw, r0, r1, lr, lc, colok, acc); break;
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        if (u >= nv) break;
        double *out = a.part + ((size_t)(chunk * a.nb + ib + u * NW) * a.ng + g) * 256;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(lr + 4 * q) * 16 + lc] = acc[u][q];
      }
    }
  }
}

// Stacked groups hold only a few chunks each: one thread per output sums
// its group's chunks in chunk order.
__global__ __launch_bounds__(kB) void proj_reduce_stacked_kernel(const double *part, int nchunk, int nb, int ng,
                                                                 int SS, double *out, size_t o_istride,
                                                                 size_t o_gstride, int accumulate,
                                                                 const GroupDesc *groups, int stacked_rows) {
  const size_t total = (size_t)nb * ng * SS;
  const size_t e = (size_t)blockIdx.x * kB + threadIdx.x;
  if (e >= total) return;
  const GroupDesc gd = groups[(e / SS) % ng];
  const int nc = gd.active ? min(nchunk, (gd.nr + stacked_rows - 1) / stacked_rows) : 0;
  // Eight loads in flight, then the additions in chunk order: the same sum,
  // bit for bit, as one chunk at a time.
  double s = 0.0;
  int c = 0;
  for (; c + 8 <= nc; c += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = part[(size_t)(c + q) * total + e];
#pragma unroll
    for (int q = 0; q < 8; ++q) s += v[q];
  }
  for (; c < nc; ++c) s += part[(size_t)c * total + e];
  const size_t ig = e / SS, el = e % SS;
  const size_t i = ig / ng, g = ig % ng;
  double *o = out + i * o_istride + g * o_gstride + el;
  *o = accumulate ? *o + s : s;
}

// out[i*o_istride + g*o_gstride + e] (+)= sum_chunk part[...]: each
// workgroup owns 64 consecutive outputs; its 4 wavefronts split the chunks
// and combine through LDS.
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(64 * kRedWaves) void proj_reduce_kernel(const double *part, int nchunk, int nb,
                                                                     int ng, int SS, double *out,
                                                                     size_t o_istride, size_t o_gstride,
                                                                     int accumulate, const GroupDesc *groups,
                                                                     int stacked_rows) {
  __shared__ double red[kRedWaves][64];
  const size_t total = (size_t)nb * ng * SS;
  const size_t e = (size_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int w = threadIdx.x >> 6;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (e < total) {
    int nc = nchunk;
    if (stacked_rows > 0) {
      // chunk rows for stacked groups; chunks past the group's rows are unwritten
      const GroupDesc gd = groups[(e / SS) % ng];
      nc = gd.active ? min(nchunk, (gd.nr + stacked_rows - 1) / stacked_rows) : 0;
    }
    // four independent sums keep several loads in flight per lane
    int c = w;
    for (; c + 3 * kRedWaves < nc; c += 4 * kRedWaves) {
      s0 += part[(size_t)c * total + e];
      s1 += part[(size_t)(c + kRedWaves) * total + e];
      s2 += part[(size_t)(c + 2 * kRedWaves) * total + e];
      s3 += part[(size_t)(c + 3 * kRedWaves) * total + e];
    }
    for (; c < nc; c += kRedWaves) s0 += part[(size_t)c * total + e];
  }
  red[w][threadIdx.x & 63] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (w == 0 && e < total) {
    double s = 0.0;
    for (int q = 0; q < kRedWaves; ++q) s += red[q][threadIdx.x];
    const size_t ig = e / SS, el = e % SS;
    const size_t i = ig / ng, g = ig % ng;
    double *o = out + i * o_istride + g * o_gstride + el;
    *o = accumulate ? *o + s : s;
  }
}

// ------------------------------------------------------------- NN GEMM
// Y[r, c0+b] = beta*Y[r, c0+b] + alpha * sum_i sum_a V_i[r, c0+a] * Psi_i[a][b]
// Psi_i for group g at psi + (i*psi_istride) + g*psi_gstride, row-major S x S.
//
// Matrix cores as in the projection: A = V (rows x a), B = Psi (a x b).  A
// workgroup of 8 wavefronts owns 128 rows (one 16-row tile per wavefront)
// and all S columns.  V (128 x 16) and Psi (16 x S) chunks are staged in LDS
// (double buffered: the next chunk's global loads are issued before the
// current chunk's MFMAs); each wavefront runs S/16 independent accumulators.
// Each output row is read and written only by its own workgroup, so Y may
// alias the (single) V block.
constexpr int kGemmThreads = 512;

template <typename TVV, typename TY, int S>
__global__ __launch_bounds__(kGemmThreads, 4) void block_gemm_kernel(GemmArgs a) {
  constexpr int RT = 128;           // rows per workgroup
  constexpr int KA = 16;            // inner chunk
  constexpr int T = S / 16;
  constexpr int NT = kGemmThreads;
  constexpr int VE = RT * KA / NT;  // V elements per thread per chunk (4)
  constexpr int PE = (KA * S + NT - 1) / NT;
  __shared__ double sv[2][RT][KA + 1];
  __shared__ double sp[2][KA][S + 2];
  const int g = blockIdx.y;
  const GroupDesc gd = a.groups[g];
  if (blockIdx.x * RT >= gd.nr) return;
  const int r0 = gd.r0 + blockIdx.x * RT;
  const int rend = gd.r0 + gd.nr;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  const bool takes_part = a.use_k ? gd.K > 0 : gd.active != 0;
  if (!takes_part) {
    // Skipped group: a fresh block (beta == 0) is zeroed, anything else kept.
    if (a.beta == 0.0 && !a.keep_skipped) {
      TY *Y = static_cast<TY *>(a.Y) + gd.c0;
      for (int e = tid; e < RT * gd.w; e += blockDim.x) {
        const int r = r0 + e / gd.w, c = e % gd.w;
        if (r < rend) Y[(size_t)r * a.ldy + c] = TY(0);
      }
    }
    return;
  }
  dbl4 acc[T];
#pragma unroll
  for (int y = 0; y < T; ++y) acc[y] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int kchunks = (gd.w + KA - 1) / KA;
  const int nchunks = (a.use_k ? min(a.nb, gd.K) : a.nb) * kchunks;
  double rv[VE], rp[PE];
  auto fetch = [&](int c) {
    const int i = c / kchunks, k0 = (c % kchunks) * KA;
    const TVV *V = static_cast<const TVV *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0;
    const double *P = a.psi + (size_t)i * a.psi_istride + (size_t)g * a.psi_gstride;
#pragma unroll
    for (int q = 0; q < VE; ++q) {
      const int e = tid + q * NT;
      const int rr = e / KA, kk = e % KA;
      const int r = r0 + rr, k = k0 + kk;
      // Last use of the basis block in this pass: streamed without displacing
      // W from the caches (W is read again by the next projection).
      rv[q] = (r < rend && k < gd.w) ? to_d(__builtin_nontemporal_load(&V[(size_t)r * a.ldv + k])) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PE; ++q) {
      const int e = tid + q * NT;
      const int kk = e / S, cc = e % S;
      const int k = k0 + kk;
      rp[q] = (e < KA * S && k < gd.w && cc < gd.w) ? P[(size_t)k * a.ldpsi + cc] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int q = 0; q < VE; ++q) {
      const int e = tid + q * NT;
      sv[buf][e / KA][e % KA] = rv[q];
    }
#pragma unroll
    for (int q = 0; q < PE; ++q) {
      const int e = tid + q * NT;
      if (e < KA * S) sp[buf][e / S][e % S] = rp[q];
    }
  };
  if (nchunks > 0) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) fetch(c + 1);      // global loads in flight during the MFMAs
#pragma unroll
    for (int k4 = 0; k4 < KA / 4; ++k4) {
      const double fa = sv[buf][wave * 16 + lc][k4 * 4 + lr];   // A[m = lc][k = lr]
      double fb[T];
#pragma unroll
      for (int y = 0; y < T; ++y) fb[y] = sp[buf][k4 * 4 + lr][y * 16 + lc];
#pragma unroll
      for (int y = 0; y < T; ++y) acc[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa, fb[y], acc[y], 0, 0, 0);
    }
    if (c + 1 < nchunks) stash(buf ^ 1);
    __syncthreads();
  }
  TY *Y = static_cast<TY *>(a.Y) + gd.c0;
#pragma unroll
  for (int y = 0; y < T; ++y) {
    const int c = y * 16 + lc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + wave * 16 + lr + 4 * q;
      if (r < rend && c < gd.w) {
        TY *yp = Y + (size_t)r * a.ldy + c;
        const double v = a.alpha * acc[y][q];
        *yp = static_cast<TY>(a.beta == 0.0 ? v : a.beta * to_d(*yp) + v);
      }
    }
  }
}

// Y[r, c0+b] = beta*Y + alpha * sum_i sum_a V_i[r, c0+a] * Psi_i[a][b] for
// FP64 blocks of width 128 (groups with widths and offsets that are
// multiples of 4, GemmArgs::vec4): the outer block-MGS update.  A workgroup
// of 16 wavefronts owns 256 rows, one 16-row tile per wavefront, and keeps
// one whole Psi_i (128 x 128, padded rows) in LDS, loaded once per basis
// block, so its wavefronts meet at a barrier once per block instead of once
// per 16-column chunk.  The basis rows go straight from memory into the
// matrix-core operands: in each 16-column chunk lane l loads columns
// 4(l/16)..4(l/16)+3 of row l%16 of its tile as two 16-byte vectors, and
// MFMA step m takes column 4(l/16)+m as A[i = l%16][k = l/16] with
// Psi[4(l/16)+m][16y + l%16] as B[k][j] for output tile y, which sums over
// every column once.  The next chunk's loads are in flight during the
// current chunk's 32 MFMAs.  Rows past the group re-read its last row and
// column groups past its width read columns 0..3 (their Psi rows are zero);
// each wavefront writes only the rows it read, so Y may alias the (single)
// V block.
constexpr int kG128Waves = 16;
constexpr int kG128Rows = 16 * kG128Waves;
constexpr int kG128Ld = 130;   // Psi row stride in LDS (doubles)
typedef double dbl2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(64 * kG128Waves, 1) void gemm128d_kernel(GemmArgs a) {
  __shared__ double sp[128 * kG128Ld];
  const int g = blockIdx.y;
  const GroupDesc gd = a.groups[g];
  if (blockIdx.x * kG128Rows >= gd.nr) return;
  const int r0 = gd.r0 + blockIdx.x * kG128Rows;
  const int rend = gd.r0 + gd.nr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int w = gd.w;
  const bool takes_part = a.use_k ? gd.K > 0 : gd.active != 0;
  double *Y = static_cast<double *>(a.Y) + gd.c0;
  if (!takes_part) {
    if (a.beta == 0.0 && !a.keep_skipped)
      for (int e = tid; e < kG128Rows * w; e += blockDim.x) {
        const int r = r0 + e / w, c = e % w;
        if (r < rend) Y[(size_t)r * a.ldy + c] = 0.0;
      }
    return;
  }
  const int nblk = a.use_k ? min(a.nb, gd.K) : a.nb;
  const int kch = (w + 15) / 16;
  const int nchunks = nblk * kch;
  const int row = min(r0 + wave * 16 + i16, rend - 1);
  auto vrow = [&](int c) {
    const int i = c / kch, k = (c % kch) * 16 + 4 * kq;
    const double *V = static_cast<const double *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0;
    return reinterpret_cast<const dbl2 *>(V + (size_t)row * a.ldv + (k < w ? k : 0));
  };
  dbl4 acc[8];
#pragma unroll
  for (int y = 0; y < 8; ++y) acc[y] = dbl4{0.0, 0.0, 0.0, 0.0};
  dbl2 n0 = dbl2{0.0, 0.0}, n1 = dbl2{0.0, 0.0};
  if (nchunks > 0) {
    const dbl2 *vp = vrow(0);
    n0 = __builtin_nontemporal_load(vp);
    n1 = __builtin_nontemporal_load(vp + 1);
  }
  for (int c = 0; c < nchunks; ++c) {
    const int ch = c % kch;
    if (ch == 0) {
      // Psi of the next basis block into LDS (zero past the group's width)
      const double *P = a.psi + (size_t)(c / kch) * a.psi_istride + (size_t)g * a.psi_gstride;
      __syncthreads();
      for (int e = tid; e < 128 * 128; e += 64 * kG128Waves) {
        const int k = e >> 7, j = e & 127;
        sp[k * kG128Ld + j] = (k < w && j < w) ? P[(size_t)k * a.ldpsi + j] : 0.0;
      }
      __syncthreads();
    }
    const double av[4] = {n0.x, n0.y, n1.x, n1.y};
    if (c + 1 < nchunks) {
      const dbl2 *vp = vrow(c + 1);
      n0 = __builtin_nontemporal_load(vp);
      n1 = __builtin_nontemporal_load(vp + 1);
    }
    const double *spk = sp + (ch * 16 + 4 * kq) * kG128Ld + i16;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double fb[8];
#pragma unroll
      for (int y = 0; y < 8; ++y) fb[y] = spk[m * kG128Ld + y * 16];
#pragma unroll
      for (int y = 0; y < 8; ++y) acc[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], fb[y], acc[y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int y = 0; y < 8; ++y) {
    const int c = y * 16 + i16;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + wave * 16 + kq + 4 * q;
      if (r < rend && c < w) {
        double *yp = Y + (size_t)r * a.ldy + c;
        const double v = a.alpha * acc[y][q];
        *yp = a.beta == 0.0 ? v : a.beta * *yp + v;
      }
    }
  }
}

// Y[r, c0+b] = beta*Y + alpha * sum_i sum_a V_i[r, c0+a] * Psi_i[a][b] for
// width-16 groups of an FP32 basis (the inner solve's updates): FP32 matrix
// cores (v_mfma_f32_16x16x4f32) accumulating the 16 nb products of each
// output in FP32, Psi rounded to FP32.  Lane l owns row l%16 of a 16-row
// tile and loads its columns 4(l/16)..4(l/16)+3 as one 16-byte vector, so a
// tile is one load per block with no LDS staging: MFMA step m takes column
// 4(l/16)+m of the loaded row as A[i = l%16][k = l/16] and Psi[4(l/16)+m][l%16]
// as B[k][j = l%16], which sums over every column once.  The result lands as
// D[4(l/16)+q][l%16] in acc[q].  A workgroup (4 wavefronts, 4 tiles each)
// owns 256 rows of one group; all groups of a row chunk run on one XCD
// (xcd_chunk_group), so the lines of a row are fetched into one L2 only.  The next block's loads are in flight during the current
// block's MFMAs.
constexpr int kG16Tiles = 4;
constexpr int kG16Rows = 4 * 16 * kG16Tiles;

// VEC: every group's width and column offset and both leading dimensions are
// multiples of 4 and the blocks 16-byte aligned (GemmArgs::vec4); otherwise
// the same lanes load their four columns one by one.  All loads are issued
// at clamped addresses (rows past the group re-read its last row, columns
// past its width read column 0): rows past the group only feed outputs that
// are not stored, and columns past the width meet zero coefficients.
template <bool VEC>
__global__ __launch_bounds__(256) void gemm16f_kernel(GemmArgs a, int ng) {
  int g, chunk;
  xcd_chunk_group(ng, chunk, g);
  const GroupDesc gd = a.groups[g];
  if (chunk * kG16Rows >= gd.nr) return;
  const int r0 = gd.r0 + chunk * kG16Rows, rend = gd.r0 + gd.nr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  float *Y = static_cast<float *>(a.Y) + gd.c0;
  const bool takes_part = a.use_k ? gd.K > 0 : gd.active != 0;
  if (!takes_part) {
    // Skipped group: a fresh block (beta == 0) is zeroed, anything else kept.
    if (a.beta == 0.0 && !a.keep_skipped) {
      for (int e = threadIdx.x; e < kG16Rows * gd.w; e += 256) {
        const int r = r0 + e / gd.w, c = e % gd.w;
        if (r < rend) Y[(size_t)r * a.ldy + c] = 0.0f;
      }
    }
    return;
  }
  const int nb = a.use_k ? min(a.nb, gd.K) : a.nb;
  const int w = gd.w;
  // element offsets of this lane's four columns (clamped into the group)
  const int cq = 4 * kq < w ? 4 * kq : 0;
  int ce[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) ce[m] = min(4 * kq + m, w - 1);
  unsigned rowoff[kG16Tiles];
#pragma unroll
  for (int t = 0; t < kG16Tiles; ++t) {
    const int r = min(r0 + wave * 16 * kG16Tiles + 16 * t + i16, rend - 1);
    rowoff[t] = static_cast<unsigned>(r) * static_cast<unsigned>(a.ldv);
  }
  const int jc = min(i16, w - 1);
  auto load = [&](int i, flt4 (&v)[kG16Tiles], double (&b)[4]) {
    const auto *V = global_ptr<float>(static_cast<const float *>(a.Vp ? a.Vp[a.i0 + i] : a.V1) + gd.c0);
    const auto *P = global_ptr<double>(a.psi + (size_t)i * a.psi_istride + (size_t)g * a.psi_gstride);
#pragma unroll
    for (int m = 0; m < 4; ++m) b[m] = P[ce[m] * a.ldpsi + jc];
#pragma unroll
    for (int t = 0; t < kG16Tiles; ++t) {
      if (VEC) {
        v[t] = *reinterpret_cast<const __attribute__((address_space(1))) flt4 *>(V + rowoff[t] + cq);
      } else {
#pragma unroll
        for (int m = 0; m < 4; ++m) v[t][m] = V[rowoff[t] + ce[m]];
      }
    }
  };
  flt4 acc[kG16Tiles];
#pragma unroll
  for (int t = 0; t < kG16Tiles; ++t) acc[t] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
  flt4 cv[kG16Tiles];
  double cb[4];
  if (nb > 0) load(0, cv, cb);
  for (int i = 0; i < nb; ++i) {
    flt4 nv[kG16Tiles];
    double nbv[4];
    if (i + 1 < nb) load(i + 1, nv, nbv);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      // coefficient rows past the group's width are zero
      const float bm = 4 * kq + m < w ? static_cast<float>(cb[m]) : 0.0f;
#pragma unroll
      for (int t = 0; t < kG16Tiles; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[t][m], bm, acc[t], 0, 0, 0);
    }
    if (i + 1 < nb) {
#pragma unroll
      for (int t = 0; t < kG16Tiles; ++t) cv[t] = nv[t];
#pragma unroll
      for (int m = 0; m < 4; ++m) cb[m] = nbv[m];
    }
  }
  if (i16 >= w) return;
#pragma unroll
  for (int t = 0; t < kG16Tiles; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + wave * 16 * kG16Tiles + 16 * t + 4 * kq + q;
      if (r < rend) {
        float *yp = Y + (size_t)r * a.ldy + i16;
        const double v = a.alpha * static_cast<double>(acc[t][q]);
        *yp = static_cast<float>(a.beta == 0.0 ? v : a.beta * static_cast<double>(*yp) + v);
      }
    }
  }
}

// -------------------------------------------------------- Cholesky (FP64)
// One workgroup (16 wavefronts) per group: G (w x w, ld S) -> R upper with
// R^T R = G + shift I, and Rinv = R^{-1}.  Blocked right-looking Cholesky in
// LDS with 16-column panels: the diagonal block by one wavefront, the panel
// row by forward substitution (one thread per column), the trailing update
// on matrix cores.  If a pivot is not positive the factorization restarts
// with a shift sigma = 11 (n w + w (w+1)) u ||G||_F (Fukaya et al. 2020) and
// flags it (bit 0).  With a.deflate, a pivot that is not positive (column k
// lies in the span of the columns before it to rounding, as with linearly
// dependent right-hand sides) instead gives R a zero row k (bit 1); the
// factor that Rinv inverts keeps 1 there, so column k of W Rinv is the
// rounding-level remainder of w_k, which launch_fill_deflated replaces.  Bits
// go to flags[0] (kept over a solve) and flags[1] (this factorization).
// A positive pivot d_k <= 1e-10 G_kk (w_k's component off the columns
// before it below 1e-5 of its norm: rounding noise of a dependent column, or
// a severely ill-conditioned block) is kept but flagged (bit 2): 1/sqrt(d_k)
// amplifies the noise, whose components along the earlier basis would
// survive the second pass, so the caller re-projects the block and adds a
// third pass as after a deflation.
// R^{-1} is formed blockwise: diagonal 16 x 16 inverses, then
// X_IJ = -X_II sum_{I<M<=J} R_IM X_MJ, one block column per wavefront; X is
// kept transposed in the strict lower triangle with its diagonal in dinv.
constexpr int kCholLd = 129;

__device__ __forceinline__ double chol_xblk(const double *G, const double *dinv, int I, int J, int k, int n) {
  // X_{IJ}[k][n] for I <= J (upper-triangular inverse, stored transposed)
  const int r = I * 16 + k, c = J * 16 + n;
  return r < c ? G[c * kCholLd + r] : (r == c ? dinv[r] : 0.0);
}

__global__ __launch_bounds__(1024) void chol_kernel(CholArgs a) {
  __shared__ double G[128 * kCholLd];
  __shared__ double dinv[128];
  __shared__ double scr[8][16][17];
  __shared__ int s_fail;
  __shared__ double s_norm, s_part[16];
  __shared__ unsigned char s_dep[128];
  __shared__ int s_small;
  const int g = blockIdx.x;
  if (!a.groups[g].active) return;
  const int w = a.groups[g].w;
  const int S = a.S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  const double *Gin = a.G + (size_t)g * a.gstride;
  double *R = a.R + (size_t)g * a.gstride;
  double *Ri = a.Rinv + (size_t)g * a.gstride;
  if (tid == 0) {
    s_norm = 0.0;
    s_small = 0;
  }
  if (tid < 128) s_dep[tid] = 0;
  __syncthreads();
  double loc = 0.0;
  for (int e = tid; e < w * w; e += 1024) {
    const double v = Gin[(e / w) * S + e % w];
    loc += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) loc += __shfl_xor(loc, off, 64);
  if (lane == 0) s_part[wave] = loc;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < 16; ++q) t += s_part[q];   // fixed order
    s_norm = t;
  }
  __syncthreads();
  const double gnorm = sqrt(s_norm);
  const int nb = (w + 15) / 16;
  double shift = 0.0;
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (int e = tid; e < 128 * kCholLd; e += 1024) {
      const int r = e / kCholLd, c = e % kCholLd;
      G[e] = (r < w && c < w) ? Gin[r * S + c] + (r == c ? shift : 0.0) : 0.0;
    }
    if (tid == 0) s_fail = 0;
    __syncthreads();
    for (int j0 = 0; j0 < w; j0 += 16) {
      const int pw = min(16, w - j0);
      // diagonal block by wavefront 0
      if (wave == 0) {
        for (int kk = 0; kk < pw; ++kk) {
          const int k = j0 + kk;
          const double d = G[k * kCholLd + k];
          if (a.deflate && isfinite(d) && d <= 0.0) {
            // column k depends on columns < k: zero row k of R, no update
            if (lane > kk && lane < pw) G[k * kCholLd + j0 + lane] = 0.0;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (lane == 0) {
              G[k * kCholLd + k] = 1.0;
              s_dep[k] = 1;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            continue;
          }
          if (!(d > 0.0) || !isfinite(d)) {
            if (lane == 0) s_fail = 1;
            break;
          }
          if (a.deflate && lane == 0 && d <= 1e-10 * Gin[k * S + k]) s_small = 1;
          const double rkk = sqrt(d);
          if (lane > kk && lane < pw) G[k * kCholLd + j0 + lane] /= rkk;
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          if (lane == 0) G[k * kCholLd + k] = rkk;
          for (int e = lane; e < 256; e += 64) {
            const int i = e >> 4, j = e & 15;
            if (i > kk && j >= i && j < pw)
              G[(j0 + i) * kCholLd + j0 + j] -= G[k * kCholLd + j0 + i] * G[k * kCholLd + j0 + j];
          }
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
      }
      __syncthreads();
      if (s_fail) break;
      // panel row: R[j0:j0+pw, c] = R_PP^{-T} G[j0:j0+pw, c] for trailing columns c
      for (int c = j0 + pw + tid; c < w; c += 1024) {
        for (int kk = 0; kk < pw; ++kk) {
          const int k = j0 + kk;
          if (s_dep[k]) {
            G[k * kCholLd + c] = 0.0;
            continue;
          }
          double v = G[k * kCholLd + c];
          for (int m = 0; m < kk; ++m) v -= G[(j0 + m) * kCholLd + k] * G[(j0 + m) * kCholLd + c];
          G[k * kCholLd + c] = v / G[k * kCholLd + k];
        }
      }
      __syncthreads();
      // trailing update of the upper tiles: G[i][j] -= sum_k R[j0+k][i] R[j0+k][j]
      const int t0 = j0 + 16;
      const int nt = (w - t0 + 15) / 16;
      const int ntiles = nt > 0 ? nt * (nt + 1) / 2 : 0;
      for (int tt = wave; tt < ntiles; tt += 16) {
        int ti = 0, rem = tt;
        while (rem >= nt - ti) { rem -= nt - ti; ++ti; }
        const int tj = ti + rem;
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const int kr = (j0 + k0 + lr) * kCholLd;
          acc = mfma4(G[kr + t0 + ti * 16 + lc], G[kr + t0 + tj * 16 + lc], acc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = t0 + ti * 16 + lr + 4 * q, j = t0 + tj * 16 + lc;
          if (i < w && j < w) G[i * kCholLd + j] -= acc[q];
        }
      }
      __syncthreads();
    }
    __syncthreads();
    if (!s_fail) break;
    if (tid < 128) s_dep[tid] = 0;
    shift = shift == 0.0 ? 11.0 * ((double)a.n * w + (double)w * (w + 1)) * 1.1102230246251565e-16 * gnorm
                         : shift * 100.0;
    if (shift == 0.0) shift = 1e-300;
    if (tid == 0) {
      atomicOr(a.flags, 1);
      atomicOr(a.flags + 1, 1);
    }
  }
  // diagonal 16 x 16 inverses (one thread per column)
  for (int c = tid; c < w; c += 1024) dinv[c] = 1.0 / G[c * kCholLd + c];
  __syncthreads();
  for (int c = tid; c < w; c += 1024) {
    const int rb = (c / 16) * 16;
    for (int r = c - 1; r >= rb; --r) {
      double v = G[r * kCholLd + c] * dinv[c];
      for (int m = r + 1; m < c; ++m) v += G[r * kCholLd + m] * G[c * kCholLd + m];
      G[c * kCholLd + r] = -v * dinv[r];
    }
  }
  __syncthreads();
  // off-diagonal blocks, block column J by wavefront J
  if (wave < nb) {
    const int J = wave;
    double (*sw)[17] = scr[wave];
    for (int I = J - 1; I >= 0; --I) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int M = I + 1; M <= J; ++M) {
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const int k = k0 + lr;
          const int rr = I * 16 + lc, cc = M * 16 + k;
          const double av = (rr < w && cc < w) ? G[rr * kCholLd + cc] : 0.0;
          const double bv = (M * 16 + k < w && J * 16 + lc < w) ? chol_xblk(G, dinv, M, J, k, lc) : 0.0;
          acc = mfma4(av, bv, acc);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) sw[lr + 4 * q][lc] = acc[q];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      dbl4 x4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4) {
        const int k = k0 + lr;
        const double av = (I * 16 + lc < w && I * 16 + k < w) ? chol_xblk(G, dinv, I, I, lc, k) : 0.0;
        x4 = mfma4(av, sw[k][lc], x4);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = I * 16 + lr + 4 * q, c = J * 16 + lc;
        if (r < w && c < w) G[c * kCholLd + r] = -x4[q];
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
  }
  __syncthreads();
  for (int e = tid; e < S * S; e += 1024) {
    const int r = e / S, c = e % S;
    double rv, iv;
    if (r < w && c < w) {
      rv = c >= r && !s_dep[r] ? G[r * kCholLd + c] : 0.0;
      iv = c > r ? G[c * kCholLd + r] : (c == r ? dinv[r] : 0.0);
    } else {
      rv = iv = (r == c) ? 1.0 : 0.0;
    }
    R[r * S + c] = rv;
    Ri[r * S + c] = iv;
  }
  if (tid == 0) {
    int dep = 0;
    for (int k = 0; k < w; ++k) dep |= s_dep[k];
    if (dep) {
      atomicOr(a.flags, 2);
      atomicOr(a.flags + 1, 2);
    }
    if (s_small) {
      atomicOr(a.flags, 4);
      atomicOr(a.flags + 1, 4);
    }
    if (a.deflate && a.gflags) a.gflags[g] = (dep ? 2 : 0) | (s_small ? 4 : 0);
  }
}

// Pseudo-random fill of deflated columns: splitmix64 of (group, column, row)
// mapped to a uniform value in [-1, 1), times sqrt(3 / rows) so a column has
// unit expected norm.  One thread per row of a group.
__device__ __forceinline__ double fill_value(unsigned long long key) {
  key += 0x9E3779B97F4A7C15ull;
  key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
  key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
  key ^= key >> 31;
  return (double)(key >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

template <typename TV>
__global__ __launch_bounds__(kB) void fill_deflated_kernel(TV *Q, int ld, const double *R, int S, size_t stride,
                                                           const GroupDesc *groups, const int *flags) {
  if (!(*flags & 2)) return;
  __shared__ unsigned char dep[128];
  const int g = blockIdx.y;
  const GroupDesc gd = groups[g];
  if (!gd.active) return;
  const double *Rg = R + (size_t)g * stride;
  if (threadIdx.x < gd.w) dep[threadIdx.x] = Rg[(size_t)threadIdx.x * S + threadIdx.x] == 0.0;
  __syncthreads();
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= gd.nr) return;
  const double scale = sqrt(3.0 / gd.nr);
  TV *row = Q + (size_t)(gd.r0 + i) * ld + gd.c0;
  for (int k = 0; k < gd.w; ++k) {
    if (!dep[k]) continue;
    const unsigned long long key =
        ((unsigned long long)g << 48) ^ ((unsigned long long)k << 32) ^ (unsigned long long)i;
    row[k] = static_cast<TV>(scale * fill_value(key));
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Cholesky of a block of width w <= 16 (S = 16): one wavefront per group,
// the same unblocked steps as chol_kernel's 16-column diagonal panel (so the
// same results), the inverse by the same per-column recurrence; shifts,
// deflation and flags as there.
__global__ __launch_bounds__(64) void chol16_kernel(CholArgs a) {
  constexpr int S = 16, LD = 17;
  __shared__ double G[S * LD];
  __shared__ double dinv[S];
  __shared__ unsigned char dep[S];
  const int g = blockIdx.x;
  if (!a.groups[g].active) return;
  const int w = a.groups[g].w;
  const int lane = threadIdx.x;
  const double *Gin = a.G + (size_t)g * a.gstride;
  double *R = a.R + (size_t)g * a.gstride;
  double *Ri = a.Rinv + (size_t)g * a.gstride;
  double loc = 0.0;
  for (int e = lane; e < w * w; e += 64) {
    const double v = Gin[(e / w) * a.S + e % w];
    loc += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) loc += __shfl_xor(loc, off, 64);
  const double gnorm = sqrt(loc);
  double shift = 0.0;
  bool small = false;
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (int e = lane; e < S * LD; e += 64) {
      const int r = e / LD, c = e % LD;
      G[e] = (r < w && c < w) ? Gin[r * a.S + c] + (r == c ? shift : 0.0) : 0.0;
    }
    if (lane < S) dep[lane] = 0;
    small = false;
    bool fail = false;
    wave_sync();
    for (int k = 0; k < w; ++k) {
      const double d = G[k * LD + k];
      if (a.deflate && isfinite(d) && d <= 0.0) {
        if (lane > k && lane < w) G[k * LD + lane] = 0.0;
        wave_sync();
        if (lane == 0) {
          G[k * LD + k] = 1.0;
          dep[k] = 1;
        }
        wave_sync();
        continue;
      }
      if (!(d > 0.0) || !isfinite(d)) {
        fail = true;
        break;
      }
      if (a.deflate && d <= 1e-10 * Gin[k * a.S + k]) small = true;
      const double rkk = sqrt(d);
      if (lane > k && lane < w) G[k * LD + lane] /= rkk;
      wave_sync();
      if (lane == 0) G[k * LD + k] = rkk;
      for (int e = lane; e < 256; e += 64) {
        const int i = e >> 4, j = e & 15;
        if (i > k && j >= i && j < w) G[i * LD + j] -= G[k * LD + i] * G[k * LD + j];
      }
      wave_sync();
    }
    if (!fail) break;
    shift = shift == 0.0 ? 11.0 * ((double)a.n * w + (double)w * (w + 1)) * 1.1102230246251565e-16 * gnorm
                         : shift * 100.0;
    if (shift == 0.0) shift = 1e-300;
    if (lane == 0) {
      atomicOr(a.flags, 1);
      atomicOr(a.flags + 1, 1);
    }
  }
  if (lane < w) dinv[lane] = 1.0 / G[lane * LD + lane];
  wave_sync();
  if (lane < w) {
    const int c = lane;
    for (int r = c - 1; r >= 0; --r) {
      double v = G[r * LD + c] * dinv[c];
      for (int m = r + 1; m < c; ++m) v += G[r * LD + m] * G[c * LD + m];
      G[c * LD + r] = -v * dinv[r];
    }
  }
  wave_sync();
  for (int e = lane; e < S * S; e += 64) {
    const int r = e / S, c = e % S;
    double rv, iv;
    if (r < w && c < w) {
      rv = c >= r && !dep[r] ? G[r * LD + c] : 0.0;
      iv = c > r ? G[c * LD + r] : (c == r ? dinv[r] : 0.0);
    } else {
      rv = iv = (r == c) ? 1.0 : 0.0;
    }
    R[r * a.S + c] = rv;
    Ri[r * a.S + c] = iv;
  }
  if (lane == 0) {
    int dp = 0;
    for (int k = 0; k < w; ++k) dp |= dep[k];
    if (dp) {
      atomicOr(a.flags, 2);
      atomicOr(a.flags + 1, 2);
    }
    if (small) {
      atomicOr(a.flags, 4);
      atomicOr(a.flags + 1, 4);
    }
    if (a.deflate && a.gflags) a.gflags[g] = (dp ? 2 : 0) | (small ? 4 : 0);
  }
}

__global__ __launch_bounds__(kB) void split_groups_kernel(const GroupDesc *groups, const int *gflags, int ng,
                                                          GroupDesc *out) {
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g >= ng) return;
  GroupDesc a = groups[g], b = groups[g];
  const bool brk = (gflags[g] & 6) != 0;
  a.active = groups[g].active && brk;
  b.active = groups[g].active && !brk;
  out[g] = a;
  out[ng + g] = b;
}

// C = A * B for upper-triangular S x S blocks per group (C may not alias).
__global__ __launch_bounds__(kB) void trmm_upper_kernel(const double *A, const double *B, double *C,
                                                        int S, size_t stride, size_t cstride,
                                                        const GroupDesc *groups) {
  const int g = blockIdx.y;
  if (!groups[g].active) return;
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

__device__ __forceinline__ double *htmat(double *const *cols, int k, int g, int ng, int S) {
  return cols[k] + (size_t)ng * (k + 2) * S * S + (size_t)ng * S + (size_t)ng * S * S + (size_t)g * S * S;
}

// Reflector j of a stored block (compact 2S x S factor H, LAPACK geqrf
// convention): Y[r][j] = 1 for r == j, H[r][j] for r > j, 0 above.
__device__ __forceinline__ double yref(const double *H, int S, int r, int j) {
  return r == j ? 1.0 : (r > j ? H[(size_t)r * S + j] : 0.0);
}

// A 2S x 16 column tile `seg` (LDS, rows padded to 17) is multiplied by
// Omega^T = I - Y T^T Y^T, with Y the unit lower-trapezoidal compact factor in
// H (2S x S row-major) and T upper triangular (S x S row-major): m1 = Y^T seg,
// m1 = T^T m1, seg -= Y m1, on matrix cores by the workgroup's 8 wavefronts.
// H and T stream through LDS (`hs`) in chunks of kWyRows(S) rows, the next
// chunk prefetched into registers while the current one is used.  Only rows
// r < rmax (S or 2S) of seg may be non-zero on entry.  Ends on a barrier.
constexpr int kWyThreads = 512;
__host__ __device__ constexpr int kWyRows(int S) { return S < 64 ? S : 64; }

template <int S>
__device__ __forceinline__ void wy_apply_tile(const double *__restrict__ H, const double *__restrict__ T,
                                              double (*seg)[17], double (*m1)[17], double (*hs)[S + 1],
                                              int rmax) {
  constexpr int NW = kWyThreads / 64;
  constexpr int NT = S / 16, NU = (NT + NW - 1) / NW, CR = kWyRows(S);
  constexpr int PER = (CR * S + kWyThreads - 1) / kWyThreads;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  double pre[PER];
  auto fetch = [&](const double *M, int r0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * kWyThreads;
      pre[q] = e < CR * S ? M[(size_t)(r0 + e / S) * S + (e % S)] : 0.0;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * kWyThreads;
      if (e < CR * S) hs[e / S][e % S] = pre[q];
    }
  };
  // m1 = Y^T seg over rows r < rmax.  Two accumulators per tile (even and
  // odd 4-row steps) keep two independent matrix-core chains in flight.
  dbl4 acc[NU], acc2[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) acc[u] = acc2[u] = dbl4{0.0, 0.0, 0.0, 0.0};
  fetch(H, 0);
  for (int r0 = 0; r0 < rmax; r0 += CR) {
    __syncthreads();
    stash();
    __syncthreads();
    if (r0 + CR < rmax) fetch(H, r0 + CR);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int jt = wave + NW * u;
      if (jt < NT && jt * 16 < r0 + CR) {
        const int j = jt * 16 + lc;
#pragma unroll 4
        for (int kk = 0; kk < CR; kk += 8) {
          const int r = r0 + kk + lr, r2 = r + 4;
          const double h = hs[kk + lr][j], h2 = hs[kk + 4 + lr][j];
          acc[u] = mfma4(r == j ? 1.0 : (r > j ? h : 0.0), seg[r][lc], acc[u]);
          acc2[u] = mfma4(r2 == j ? 1.0 : (r2 > j ? h2 : 0.0), seg[r2][lc], acc2[u]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) acc[u] += acc2[u];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int jt = wave + NW * u;
    if (jt < NT)
#pragma unroll
      for (int q = 0; q < 4; ++q) m1[jt * 16 + lr + 4 * q][lc] = acc[u][q];
  }
  // m2 = T^T m1: M2[j][c] = sum_{q <= j} T[q][j] M1[q][c], T streamed by rows q
#pragma unroll
  for (int u = 0; u < NU; ++u) acc[u] = acc2[u] = dbl4{0.0, 0.0, 0.0, 0.0};
  fetch(T, 0);
  for (int q0 = 0; q0 < S; q0 += CR) {
    __syncthreads();
    stash();
    __syncthreads();
    if (q0 + CR < S) fetch(T, q0 + CR);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int jt = wave + NW * u;
      if (jt < NT && q0 < jt * 16 + 16) {
        const int j = jt * 16 + lc;
#pragma unroll 4
        for (int kk = 0; kk < CR; kk += 8) {
          const int q = q0 + kk + lr, q2 = q + 4;
          acc[u] = mfma4(q <= j ? hs[kk + lr][j] : 0.0, m1[q][lc], acc[u]);
          acc2[u] = mfma4(q2 <= j ? hs[kk + 4 + lr][j] : 0.0, m1[q2][lc], acc2[u]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) acc[u] += acc2[u];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int jt = wave + NW * u;
    if (jt < NT)
#pragma unroll
      for (int q = 0; q < 4; ++q) m1[jt * 16 + lr + 4 * q][lc] = acc[u][q];
  }
  // seg -= Y m2, Y streamed by rows; wavefront w < CR/16 owns row tile w of
  // the chunk
  fetch(H, 0);
  for (int r0 = 0; r0 < 2 * S; r0 += CR) {
    __syncthreads();
    stash();
    __syncthreads();
    if (r0 + CR < 2 * S) fetch(H, r0 + CR);
    if (wave < CR / 16) {
      const int rt = r0 / 16 + wave;
      dbl4 a4 = {0.0, 0.0, 0.0, 0.0}, b4 = {0.0, 0.0, 0.0, 0.0};
      const int r = rt * 16 + lc;
      const int jend = min(S, rt * 16 + 16);   // a multiple of 16
#pragma unroll 4
      for (int j0 = 0; j0 < jend; j0 += 8) {
        const int j = j0 + lr, j2 = j + 4;
        const double h = hs[r - r0][j], h2 = hs[r - r0][j2];
        a4 = mfma4(r == j ? 1.0 : (r > j ? h : 0.0), m1[j][lc], a4);
        b4 = mfma4(r == j2 ? 1.0 : (r > j2 ? h2 : 0.0), m1[j2][lc], b4);
      }
      a4 += b4;
#pragma unroll
      for (int q = 0; q < 4; ++q) seg[rt * 16 + lr + 4 * q][lc] -= a4[q];
    }
  }
  __syncthreads();
}

// Grid (S/16 column tiles, groups).  A workgroup owns 16 columns of the new
// block column [Psi_0; ...; Psi_{k+1}] and applies Omega_i^T, i = 0..k-1, to
// rows [iS, iS + 2S): a two-block window slides down the column in LDS.
template <int S>
__global__ __launch_bounds__(kWyThreads) void hess_apply_kernel(HessArgs a) {
  __shared__ double seg[2 * S][17];
  __shared__ double m1[S][17];
  __shared__ double hs[kWyRows(S)][S + 1];
  const int b0 = blockIdx.x * 16, g = blockIdx.y;
  const GroupDesc gd = a.groups[g];
  const int k = a.k;
  if (b0 >= gd.w || !gd.active || k <= 0) return;
  const int tid = threadIdx.x;
  auto xfer = [&](int i, int half, bool load) {
    double *blk = hblock(a.cols, k, g, i, S);
    for (int e = tid; e < S * 16; e += kWyThreads) {
      const int r = e >> 4, c = e & 15;
      double *gp = blk + (size_t)r * S + b0 + c;
      if (load) seg[half * S + r][c] = *gp;
      else *gp = seg[half * S + r][c];
    }
  };
  xfer(0, 0, true);
  xfer(1, 1, true);
  __syncthreads();
  // Block i + 2 of the column is untouched until step i + 1: its loads are
  // issued before step i's products and land in the window afterwards.
  constexpr int NE = (S * 16 + kWyThreads - 1) / kWyThreads;
  double nxt[NE];
  auto prefetch = [&](int i) {
    const double *blk = hblock(a.cols, k, g, i, S);
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int e = tid + q * kWyThreads;
      nxt[q] = e < S * 16 ? blk[(size_t)(e >> 4) * S + b0 + (e & 15)] : 0.0;
    }
  };
  for (int i = 0; i < k; ++i) {
    if (i + 2 <= k) prefetch(i + 2);
    wy_apply_tile<S>(hblock(a.cols, i, g, i, S), htmat(a.cols, i, g, a.ng, S), seg, m1, hs, 2 * S);
    if (i + 1 < k) {
      // block i is final; slide the window down by one block
      xfer(i, 0, false);
      for (int e = tid; e < S * 16; e += kWyThreads) seg[e >> 4][e & 15] = seg[S + (e >> 4)][e & 15];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        const int e = tid + q * kWyThreads;
        if (e < S * 16) seg[S + (e >> 4)][e & 15] = nxt[q];
      }
      __syncthreads();
    }
  }
  xfer(k - 1, 0, false);
  xfer(k, 1, false);
}

// One workgroup (16 wavefronts) per group: D = [Psi_kk; C] (2S x S, blocks
// k and k+1 of block column k) is factored in place by blocked Householder
// QR (LAPACK geqrf convention: reflector j is v = e_j + D[j+1.., j], H_j =
// I - tau_j v v^T).  Each 16-column panel is factored in LDS; its block
// reflector updates the trailing columns on matrix cores.  The triangular
// factor T of Omega_k = H_0 ... H_{w-1} = I - Y T Y^T is accumulated panel by
// panel (T12 = -T11 (Y1^T Y2) T22, Schreiber & Van Loan 1989) into htmat(k);
// columns and rows of T past w stay zero.
template <int S>
__global__ __launch_bounds__(1024) void hess_qr_kernel(HessQrArgs a) {
  constexpr int S2 = 2 * S, PB = 16;
  const int g = blockIdx.x;
  const GroupDesc gd = a.groups[g];
  if (!gd.active) return;
  const int w = gd.w;
  double *D = hblock(a.cols, a.k, g, a.k, S);
  double *tau = htau(a.cols, a.k, g, a.ng, S);
  double *Tg = htmat(a.cols, a.k, g, a.ng, S);
  __shared__ double pan[S2][PB + 1];       // panel columns j0..j0+15, rows j0..2S-1
  __shared__ double Tp[PB][PB + 1];        // panel T (upper)
  __shared__ double scr[16][16][17];       // per-wavefront 16 x 16 scratch
  __shared__ double zs[S][PB + 1];         // T-merge scratch (j0 x 16)
  __shared__ double red[16];
  __shared__ double s_beta, s_tau, s_scale;
  const int tid = threadIdx.x, nt = 1024;
  const int lane = tid & 63, wid = tid >> 6, nw = 16;
  const int lr = lane >> 4, lc = lane & 15;
  for (int e = tid; e < S * S; e += nt) Tg[e] = 0.0;
  for (int j = tid; j < S; j += nt) tau[j] = 0.0;
  __syncthreads();
  for (int j0 = 0; j0 < w; j0 += PB) {
    const int pw = min(PB, w - j0);
    for (int e = tid; e < S2 * PB; e += nt) {
      const int r = e / PB, c = e % PB;
      pan[r][c] = (c < pw && r >= j0) ? D[(size_t)r * S + j0 + c] : 0.0;
    }
    __syncthreads();
    // unblocked Householder on the panel
    for (int jj = 0; jj < pw; ++jj) {
      const int j = j0 + jj;
      double part = 0.0;
      for (int r = j + 1 + tid; r < S2; r += nt) part += pan[r][jj] * pan[r][jj];
      for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
      if (lane == 0) red[wid] = part;
      __syncthreads();
      if (tid == 0) {
        double sig = 0.0;
        for (int q = 0; q < nw; ++q) sig += red[q];
        const double alpha = pan[j][jj];
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
      for (int r = j + 1 + tid; r < S2; r += nt) pan[r][jj] *= s_scale;
      __syncthreads();
      if (tid == 0) pan[j][jj] = s_beta;
      const double t = s_tau;
      for (int c = jj + 1 + wid; c < pw; c += nw) {
        double dot = 0.0;
        for (int r = j + lane; r < S2; r += 64) dot += (r == j ? 1.0 : pan[r][jj]) * pan[r][c];
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
        const double f = t * dot;
        for (int r = j + lane; r < S2; r += 64) pan[r][c] -= f * (r == j ? 1.0 : pan[r][jj]);
      }
      __syncthreads();
    }
    // Panel reflector Yp (2S x 16): Yp[r][q] = 1 at r == j0 + q, pan[r][q] below.
    auto ypan = [&](int r, int q) -> double {
      return r == j0 + q ? 1.0 : (r > j0 + q ? pan[r][q] : 0.0);
    };
    // panel T (pw x pw upper) by the forward recurrence T[jj][jj] = tau_jj,
    // T[0:jj, jj] = -tau_jj T[0:jj, 0:jj] Gp[0:jj, jj] with the panel Gram
    // matrix Gp = Yp^T Yp (one wavefront, matrix cores).
    double (*gp)[17] = scr[0];
    for (int e = tid; e < PB * (PB + 1); e += nt) Tp[e / (PB + 1)][e % (PB + 1)] = 0.0;
    if (wid == 0) {
      dbl4 a4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
      for (int r0 = j0; r0 < S2; r0 += 4) a4 = mfma4(ypan(r0 + lr, lc), ypan(r0 + lr, lc), a4);
#pragma unroll
      for (int q = 0; q < 4; ++q) gp[lr + 4 * q][lc] = a4[q];
    }
    __syncthreads();
    for (int jj = 0; jj < pw; ++jj) {
      const double tj = tau[j0 + jj];
      if (tid < jj) {
        double acc = 0.0;
        for (int m = tid; m < jj; ++m) acc += Tp[tid][m] * gp[m][jj];
        Tp[tid][jj] = -tj * acc;
      }
      if (tid == 0) Tp[jj][jj] = tj;
      __syncthreads();
    }
    // factored panel back to D (R above, reflectors below the diagonal)
    for (int e = tid; e < S2 * PB; e += nt) {
      const int r = e / PB, c = e % PB;
      if (c < pw && r >= j0) D[(size_t)r * S + j0 + c] = pan[r][c];
    }
    // Trailing update, 16-column tiles, one tile per wavefront:
    // D[:, tile] -= Yp Tp^T (Yp^T D[:, tile]).
    const int c0 = j0 + PB;
    double (*sw)[17] = scr[wid];
    for (int tt = wid; c0 + tt * 16 < w; tt += nw) {
      const int tc = c0 + tt * 16 + lc;
      const bool okc = tc < w;
      dbl4 a4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
      for (int r0 = j0; r0 < S2; r0 += 4) {
        const int r = r0 + lr;
        a4 = mfma4(ypan(r, lc), okc ? D[(size_t)r * S + tc] : 0.0, a4);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) sw[lr + 4 * q][lc] = a4[q];
      // W' = Tp^T W  (16 x 16)
      dbl4 b4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m0 = 0; m0 < PB; m0 += 4) {
        const int m = m0 + lr;
        b4 = mfma4(m <= lc ? Tp[m][lc] : 0.0, sw[m][lc], b4);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) sw[lr + 4 * q][lc] = b4[q];
      // D[r][tile] -= sum_q Yp[r][q] W'[q][tile] for rows j0..2S-1
      for (int rt = j0 / 16; rt < S2 / 16; ++rt) {
        dbl4 d4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q0 = 0; q0 < PB; q0 += 4) d4 = mfma4(ypan(rt * 16 + lc, q0 + lr), sw[q0 + lr][lc], d4);
        if (okc)
#pragma unroll
          for (int q = 0; q < 4; ++q) D[(size_t)(rt * 16 + lr + 4 * q) * S + tc] -= d4[q];
      }
    }
    // T accumulation: T[j0:j0+16, j0:j0+16] = Tp; T[0:j0, j0:j0+16] =
    // -T[0:j0, 0:j0] (Y[:, 0:j0]^T Yp) Tp.  Y[r][m] for m < j0 <= r is D[r][m].
    for (int e = tid; e < PB * PB; e += nt) {
      const int q = e / PB, c = e % PB;
      if (q < pw && c < pw) Tg[(size_t)(j0 + q) * S + j0 + c] = Tp[q][c];
    }
    if (j0 > 0) {
      for (int mt = wid; mt < j0 / 16; mt += nw) {
        dbl4 a4 = {0.0, 0.0, 0.0, 0.0};
        const int m = mt * 16 + lc;
#pragma unroll 8
        for (int r0 = j0; r0 < S2; r0 += 4) {
          const int r = r0 + lr;
          a4 = mfma4(D[(size_t)r * S + m], ypan(r, lc), a4);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) zs[mt * 16 + lr + 4 * q][lc] = a4[q];
      }
      __syncthreads();
      dbl4 z2[(S / 16 + 15) / 16];
#pragma unroll
      for (int u = 0; u < (S / 16 + 15) / 16; ++u) {
        const int mt = wid + nw * u;
        dbl4 a4 = {0.0, 0.0, 0.0, 0.0};
        if (mt < j0 / 16) {
          const int m = mt * 16 + lc;
#pragma unroll 8
          for (int k0 = mt * 16; k0 < j0; k0 += 4) {
            const int kk = k0 + lr;
            a4 = mfma4(kk >= m ? Tg[(size_t)m * S + kk] : 0.0, zs[kk][lc], a4);
          }
        }
        z2[u] = a4;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < (S / 16 + 15) / 16; ++u) {
        const int mt = wid + nw * u;
        if (mt < j0 / 16)
#pragma unroll
          for (int q = 0; q < 4; ++q) zs[mt * 16 + lr + 4 * q][lc] = z2[u][q];
      }
      __syncthreads();
      for (int mt = wid; mt < j0 / 16; mt += nw) {
        dbl4 a4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q0 = 0; q0 < PB; q0 += 4) a4 = mfma4(zs[mt * 16 + lc][q0 + lr], Tp[q0 + lr][lc], a4);
        if (lc < pw)
#pragma unroll
          for (int q = 0; q < 4; ++q) Tg[(size_t)(mt * 16 + lr + 4 * q) * S + j0 + lc] = -a4[q];
      }
    }
    __syncthreads();
  }
}

// Grid (S/16 column tiles, groups): E = [Z_k; 0] <- Omega_k^T E on the tile's
// 16 columns, the residual estimates ||Lambda[:, c]|| of the new Lambda =
// E[S:2S], then zfin(k) = E[0:S], E[0:S] = Lambda, E[S:2S] = 0.
template <int S>
__global__ __launch_bounds__(kWyThreads) void hess_rot_kernel(HessQrArgs a) {
  __shared__ double seg[2 * S][17];
  __shared__ double m1[S][17];
  __shared__ double hs[kWyRows(S)][S + 1];
  constexpr int RG = kWyThreads / 16;
  __shared__ double part[RG][17];
  const int b0 = blockIdx.x * 16, g = blockIdx.y;
  const GroupDesc gd = a.groups[g];
  if (b0 >= gd.w || !gd.active) return;
  const int tid = threadIdx.x;
  double *E = a.E + (size_t)g * 2 * S * S;
  for (int e = tid; e < 2 * S * 16; e += kWyThreads) seg[e >> 4][e & 15] = E[(size_t)(e >> 4) * S + b0 + (e & 15)];
  __syncthreads();
  wy_apply_tile<S>(hblock(a.cols, a.k, g, a.k, S), htmat(a.cols, a.k, g, a.ng, S), seg, m1, hs, S);
  {
    const int c = tid & 15, rg = tid >> 4;
    double sq = 0.0;
    for (int r = S + rg; r < 2 * S; r += RG) sq += seg[r][c] * seg[r][c];
    part[rg][c] = sq;
  }
  __syncthreads();
  if (tid < 16 && b0 + tid < gd.w) {
    double sq = 0.0;
    for (int q = 0; q < RG; ++q) sq += part[q][tid];
    a.rnorm[gd.slot + b0 + tid] = sqrt(sq);
  }
  double *zf = hzfin(a.cols, a.k, g, a.ng, S);
  for (int e = tid; e < S * 16; e += kWyThreads) {
    const int r = e >> 4, c = e & 15;
    zf[(size_t)r * S + b0 + c] = seg[r][c];
    E[(size_t)r * S + b0 + c] = seg[S + r][c];
    E[(size_t)(S + r) * S + b0 + c] = 0.0;
  }
}

// ---------------------------------------------------- back substitution
// Block upper-triangular solve R Y = Zfin (R_ij = block i of column region j;
// R_jj is the upper triangle of the factored block).  Two stages:
//  1. trinv_kernel, one workgroup per (j, g), all at once: Rinv_j = R_jj^{-1}
//     (column-oriented back substitution in LDS) and Z_j = zfin(j);
//  2. for j = K-1 .. 0, backsub_step_kernel: y_j = Rinv_j Z_j on matrix cores
//     (written to Y_j by the row-j workgroups), and Z_i -= R_ij y_j for i < j
//     by the other workgroups, each recomputing y_j for its 16-column tile.
// A group uses its first min(K, K_g) block columns; its other Y blocks are 0.
__global__ __launch_bounds__(1024) void trinv_kernel(BacksubArgs a) {
  __shared__ double G[128 * 128];
  __shared__ double dinv[128];
  const int j = blockIdx.x, g = blockIdx.y;
  const int S = a.S;
  const int K = min(a.K, a.groups[g].K);
  const int w = a.groups[g].w;
  const size_t SS = (size_t)S * S;
  double *Ri = a.Rinv + ((size_t)j * a.ng + g) * SS;
  double *Z = a.Z + ((size_t)j * a.ng + g) * SS;
  if (j >= K) return;
  const double *zf = hzfin(a.cols, j, g, a.ng, S);
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) Z[e] = zf[e];
  const double *D = hblock(a.cols, j, g, j, S);
  for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
    const int r = e / w, c = e % w;
    G[r * S + c] = c >= r ? D[(size_t)r * S + c] : 0.0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < w; c += blockDim.x) dinv[c] = 1.0 / G[c * S + c];
  __syncthreads();
  // column c of R^{-1}: x[c] = dinv[c]; x[r] (r < c) kept at G[c*S + r]
  for (int c = threadIdx.x; c < w; c += blockDim.x) {
    for (int r = c - 1; r >= 0; --r) {
      double v = G[r * S + c] * dinv[c];
      for (int m = r + 1; m < c; ++m) v += G[r * S + m] * G[c * S + m];
      G[c * S + r] = -v * dinv[r];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) {
    const int r = e / S, c = e % S;
    double v = 0.0;
    if (r < w && c < w) v = c > r ? G[c * S + r] : (c == r ? dinv[r] : 0.0);
    Ri[e] = v;
  }
}

// Grid (S/16 column tiles, j + 1 block rows, groups), 4 wavefronts.
template <int S>
__global__ __launch_bounds__(256) void backsub_step_kernel(BacksubArgs a, int j) {
  constexpr int NT = S / 16, NU = (NT + 3) / 4;
  __shared__ double ys[S][17];
  const int t = blockIdx.x, i = blockIdx.y, g = blockIdx.z;
  const GroupDesc gd = a.groups[g];
  const int b0 = t * 16;
  if (b0 >= gd.w) return;
  const int K = min(a.K, gd.K);
  const size_t SS = (size_t)S * S;
  double *Yj = a.Y + ((size_t)j * a.ng + g) * SS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  if (j >= K) {
    if (i == j)
      for (int e = tid; e < S * 16; e += 256) Yj[(size_t)(e >> 4) * S + b0 + (e & 15)] = 0.0;
    return;
  }
  // y = Rinv_j Z_j[:, tile]
  const double *Ri = a.Rinv + ((size_t)j * a.ng + g) * SS;
  const double *Zj = a.Z + ((size_t)j * a.ng + g) * SS;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int rt = wave + 4 * u;
    if (rt < NT) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      const int r = rt * 16 + lc;
#pragma unroll 8
      for (int k0 = rt * 16; k0 < S; k0 += 4)      // Rinv_j is upper triangular
        acc = mfma4(Ri[(size_t)r * S + k0 + lr], Zj[(size_t)(k0 + lr) * S + b0 + lc], acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) ys[rt * 16 + lr + 4 * q][lc] = acc[q];
    }
  }
  __syncthreads();
  if (i == j) {
    for (int e = tid; e < S * 16; e += 256) Yj[(size_t)(e >> 4) * S + b0 + (e & 15)] = ys[e >> 4][e & 15];
    return;
  }
  // Z_i[:, tile] -= R_ij y
  const double *Rij = hblock(a.cols, j, g, i, S);
  double *Zi = a.Z + ((size_t)i * a.ng + g) * SS;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int rt = wave + 4 * u;
    if (rt < NT) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      const int r = rt * 16 + lc;
#pragma unroll 8
      for (int k0 = 0; k0 < S; k0 += 4) acc = mfma4(Rij[(size_t)r * S + k0 + lr], ys[k0 + lr][lc], acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) Zi[(size_t)(rt * 16 + lr + 4 * q) * S + b0 + lc] -= acc[q];
    }
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

// Row-major n x p: per-column inf-norm (mode 0: atomic max into out[p],
// zeroed by the caller, exact in any order) or sum of squares (mode 1: the
// block's sum, its 4 wavefronts combined in a fixed order, into
// part[blockIdx.x * p + c]; sum_parts_kernel adds the blocks in a fixed
// order, so the result does not depend on scheduling).
template <typename T>
__global__ __launch_bounds__(kB) void colstat_rm_kernel(int n, int p, const T *X, int ld, int mode,
                                                        double *out, double *part) {
  __shared__ double red[kB / 64][64];
  const int c = blockIdx.y * 64 + (threadIdx.x & 63);
  const int rstride = (kB / 64) * gridDim.x;
  double m = 0.0;
  if (c < p) {
    for (int r = blockIdx.x * (kB / 64) + (threadIdx.x >> 6); r < n; r += rstride) {
      const double v = to_d(X[(size_t)r * ld + c]);
      m = mode == 0 ? fmax(m, fabs(v)) : m + v * v;
    }
    if (mode == 0) atomic_max_nonneg(out + c, m);
  }
  if (mode == 0) return;
  red[threadIdx.x >> 6][threadIdx.x & 63] = m;
  __syncthreads();
  if (threadIdx.x < 64 && c < p)
    part[(size_t)blockIdx.x * p + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                       (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// out[c] = sum_b part[b * stride + c] for b < nparts, c < ncol: one block
// per column, a fixed split over its threads and a fixed reduction tree.
__global__ __launch_bounds__(kB) void sum_parts_kernel(const double *part, int nparts, size_t stride,
                                                       double *out) {
  __shared__ double red[kB / 64];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kB) s += part[(size_t)b * stride + c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

// R = B - A*X for row-major FP64 blocks (residual in working precision).
__global__ __launch_bounds__(kB) void axpby_kernel(size_t m, double alpha, const double *x,
                                                   double beta, double *y) {
  const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
  if (i < m) y[i] = alpha * x[i] + beta * y[i];
}

__global__ __launch_bounds__(kB) void add_groups_kernel(size_t count, size_t stride, const double *x,
                                                       double *y) {
  const size_t e = (size_t)blockIdx.x * kB + threadIdx.x;
  if (e >= count) return;
  const size_t o = (size_t)blockIdx.y * stride + e;
  y[o] += x[o];
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
  int nchunk;
  const int nw16 = std::min(4, a.nb);   // S = 16: wavefronts per workgroup
  if (S == 16) {
    // One workgroup per row chunk and group (all blocks): ~8192 wavefronts
    // over all of them whatever the number of blocks.
    const long rows = a.rows_sum > 0 ? a.rows_sum : (long)a.n * a.ng;
    const long target = 2048L * 4 / nw16;
    const int cr = (int)std::max<long>(64, (rows + target - 1) / target);
    a.chunk_rows = ((cr + 15) / 16) * 16;
  } else if (a.rows_sum > 0) {
    // Stacked groups of uneven height: ~1024 non-empty workgroups over all
    // groups' rows (a.n is the tallest group; shorter ones leave chunks empty).
    const long target = std::max(64, 1024 / std::max(1, a.nb));
    int cr = (int)std::max<long>(64, (a.rows_sum + target - 1) / target);
    a.chunk_rows = ((cr + 15) / 16) * 16;
  } else {
    // Chunk rows so that the grid has ~512 workgroups (two per CU) while the
    // partial sums (nchunk * nb * ng * S^2 doubles) stay small.
    int target = std::max(1, 512 / std::max(1, a.nb * a.ng));
    nchunk = std::max(1, std::min(target, (a.n + 255) / 256));
    a.chunk_rows = (a.n + nchunk - 1) / nchunk;
    a.chunk_rows = ((a.chunk_rows + 15) / 16) * 16;
  }
  nchunk = std::max(1, (a.n + a.chunk_rows - 1) / a.chunk_rows);
  part.ensure((size_t)nchunk * a.nb * a.ng * S * S * sizeof(double));
  a.part = part.as<double>();
  dim3 grid(a.nb, nchunk, a.ng);
  const int tok = prof::begin(s);
  switch (S) {
    case 16:
      if constexpr (std::is_same<TVV, float>::value && std::is_same<TW, float>::value) {
        if (a.f32acc) {
          hipLaunchKernelGGL((proj16_kernel<TVV, TW, true>), dim3(xcd_grid(nchunk, a.ng)), dim3(64 * nw16), 0, s, a);
          break;
        }
      }
      hipLaunchKernelGGL((proj16_kernel<TVV, TW, false>), dim3(xcd_grid(nchunk, a.ng)), dim3(64 * nw16), 0, s, a);
      break;
    case 32: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 32>), grid, dim3(kB), 0, s, a); break;
    case 64: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 64>), grid, dim3(kB), 0, s, a); break;
    case 128: hipLaunchKernelGGL((proj_partial_kernel<TVV, TW, 128>), grid, dim3(kB), 0, s, a); break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", S);
  }
  const size_t total = (size_t)a.nb * a.ng * S * S;
  if (a.rows_sum > 0)
    hipLaunchKernelGGL(proj_reduce_stacked_kernel, dim3((unsigned)((total + kB - 1) / kB)), dim3(kB), 0, s,
                       a.part, nchunk, a.nb, a.ng, S * S, out, o_istride, o_gstride, accumulate ? 1 : 0,
                       a.groups, a.chunk_rows);
  else
    hipLaunchKernelGGL(proj_reduce_kernel, dim3((unsigned)((total + 63) / 64)), dim3(64 * kRedWaves), 0, s,
                       a.part, nchunk, a.nb, a.ng, S * S, out, o_istride, o_gstride,
                       accumulate ? 1 : 0, a.groups, 0);
  BPF_CHECK_HIP(hipGetLastError());
  // Algorithmic traffic: each basis block and W read once; the s x s
  // products are 2 n w^2 flops per (block, group).
  // Stacked groups: rows_sum rows of width S; column groups: n rows of ng*S.
  const double cells = a.cells >= 0 ? a.cells : a.rows_sum > 0 ? (double)a.rows_sum * S : (double)a.n * a.ng * S;
  prof::end(tok, s, S <= 32 ? prof::PROJ_NARROW : prof::PROJ, cells * ((a.nb) * sizeof(TVV) + sizeof(TW)),
            2.0 * cells * S * a.nb);
}
template void launch_proj<double, double>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);
template void launch_proj<float, float>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);
template void launch_proj<float, double>(hipStream_t, ProjArgs, int, double *, size_t, size_t, bool, DevBuf &);

template <typename TVV, typename TY>
void launch_block_gemm(hipStream_t s, GemmArgs a, int S, int ng) {
  if (a.n <= 0 || ng <= 0) return;
  dim3 grid((a.n + 127) / 128, ng);   // a.n: largest group row count
  if (a.nb <= 0) {
    if (a.beta == 1.0) return;
  }
  const int tok = prof::begin(s);
  const dim3 blk(kGemmThreads);
  if (S == 16 && std::is_same<TVV, float>::value && std::is_same<TY, float>::value) {
    // FP32 basis, width 16: the inner solve's updates on FP32 matrix cores
    const long nchunk = (a.n + kG16Rows - 1) / kG16Rows;
    BPF_REQUIRE((nchunk + kXcds) * ng < (1L << 31), "update grid too large");
    const dim3 grid16(xcd_grid(nchunk, ng));
    if (a.vec4) hipLaunchKernelGGL(gemm16f_kernel<true>, grid16, dim3(256), 0, s, a, ng);
    else hipLaunchKernelGGL(gemm16f_kernel<false>, grid16, dim3(256), 0, s, a, ng);
  } else if (S == 128 && a.vec4 && std::is_same<TVV, double>::value && std::is_same<TY, double>::value) {
    BPF_REQUIRE(((a.n + kG128Rows - 1) / kG128Rows) < (1L << 31), "update grid too large");
    hipLaunchKernelGGL(gemm128d_kernel, dim3((a.n + kG128Rows - 1) / kG128Rows, ng), dim3(64 * kG128Waves), 0, s, a);
  } else switch (S) {
    case 16: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 16>), grid, blk, 0, s, a); break;
    case 32: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 32>), grid, blk, 0, s, a); break;
    case 64: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 64>), grid, blk, 0, s, a); break;
    case 128: hipLaunchKernelGGL((block_gemm_kernel<TVV, TY, 128>), grid, blk, 0, s, a); break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", S);
  }
  BPF_CHECK_HIP(hipGetLastError());
  const double cells = a.cells >= 0 ? a.cells : a.rows_sum > 0 ? (double)a.rows_sum * S : (double)a.n * ng * S;
  prof::end(tok, s, S <= 32 ? prof::GEMM_NARROW : prof::GEMM,
            cells * (a.nb * sizeof(TVV) + (a.beta == 0.0 ? 1.0 : 2.0) * sizeof(TY)),
            2.0 * cells * S * a.nb);
}
template void launch_block_gemm<double, double>(hipStream_t, GemmArgs, int, int);
template void launch_block_gemm<float, float>(hipStream_t, GemmArgs, int, int);
template void launch_block_gemm<float, double>(hipStream_t, GemmArgs, int, int);


void launch_chol(hipStream_t s, CholArgs a, int ng) {
  if (a.S > 128) BPF_THROW(BPF_E_UNSUPPORTED, "block width %d too large for Cholesky", a.S);
  if (a.S == 16) {
    hipLaunchKernelGGL(chol16_kernel, dim3(ng), dim3(64), 0, s, a);
    BPF_CHECK_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(chol_kernel, dim3(ng), dim3(1024), 0, s, a);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_split_groups(hipStream_t s, const GroupDesc *groups, const int *gflags, int ng, GroupDesc *out) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(split_groups_kernel, dim3((ng + kB - 1) / kB), dim3(kB), 0, s, groups, gflags, ng, out);
  BPF_CHECK_HIP(hipGetLastError());
}

template <typename TV>
void launch_fill_deflated(hipStream_t s, TV *Q, int ld, const double *R, int S, size_t stride,
                          const GroupDesc *groups, int ng, int gmax_rows, const int *flags) {
  if (ng <= 0 || gmax_rows <= 0) return;
  dim3 grid((gmax_rows + kB - 1) / kB, ng);
  hipLaunchKernelGGL(fill_deflated_kernel<TV>, grid, dim3(kB), 0, s, Q, ld, R, S, stride, groups, flags);
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_fill_deflated<double>(hipStream_t, double *, int, const double *, int, size_t,
                                           const GroupDesc *, int, int, const int *);
template void launch_fill_deflated<float>(hipStream_t, float *, int, const double *, int, size_t,
                                          const GroupDesc *, int, int, const int *);

void launch_trmm_upper(hipStream_t s, const double *A, const double *B, double *C, int S,
                       size_t stride, size_t cstride, const GroupDesc *groups, int ng) {
  dim3 grid((S * S + kB - 1) / kB, ng);
  hipLaunchKernelGGL(trmm_upper_kernel, grid, dim3(kB), 0, s, A, B, C, S, stride, cstride, groups);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_hess_apply(hipStream_t s, HessArgs a, int ng) {
  if (a.k <= 0) return;
  const dim3 grid(a.S / 16, ng);
  switch (a.S) {
    case 16: hipLaunchKernelGGL(hess_apply_kernel<16>, grid, dim3(kWyThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL(hess_apply_kernel<32>, grid, dim3(kWyThreads), 0, s, a); break;
    case 64: hipLaunchKernelGGL(hess_apply_kernel<64>, grid, dim3(kWyThreads), 0, s, a); break;
    case 128: hipLaunchKernelGGL(hess_apply_kernel<128>, grid, dim3(kWyThreads), 0, s, a); break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", a.S);
  }
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_hess_qr(hipStream_t s, HessQrArgs a, int ng) {
  const dim3 grid(a.S / 16, ng);
  switch (a.S) {
    case 16:
      hipLaunchKernelGGL(hess_qr_kernel<16>, dim3(ng), dim3(1024), 0, s, a);
      hipLaunchKernelGGL(hess_rot_kernel<16>, grid, dim3(kWyThreads), 0, s, a);
      break;
    case 32:
      hipLaunchKernelGGL(hess_qr_kernel<32>, dim3(ng), dim3(1024), 0, s, a);
      hipLaunchKernelGGL(hess_rot_kernel<32>, grid, dim3(kWyThreads), 0, s, a);
      break;
    case 64:
      hipLaunchKernelGGL(hess_qr_kernel<64>, dim3(ng), dim3(1024), 0, s, a);
      hipLaunchKernelGGL(hess_rot_kernel<64>, grid, dim3(kWyThreads), 0, s, a);
      break;
    case 128:
      hipLaunchKernelGGL(hess_qr_kernel<128>, dim3(ng), dim3(1024), 0, s, a);
      hipLaunchKernelGGL(hess_rot_kernel<128>, grid, dim3(kWyThreads), 0, s, a);
      break;
    default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", a.S);
  }
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_backsub(hipStream_t s, BacksubArgs a, int ng) {
  if (a.K <= 0) return;
  BPF_REQUIRE(a.Z && a.Rinv, "back substitution needs its scratch blocks");
  hipLaunchKernelGGL(trinv_kernel, dim3(a.K, ng), dim3(1024), 0, s, a);
  for (int j = a.K - 1; j >= 0; --j) {
    const dim3 grid(a.S / 16, j + 1, ng);
    switch (a.S) {
      case 16: hipLaunchKernelGGL(backsub_step_kernel<16>, grid, dim3(256), 0, s, a, j); break;
      case 32: hipLaunchKernelGGL(backsub_step_kernel<32>, grid, dim3(256), 0, s, a, j); break;
      case 64: hipLaunchKernelGGL(backsub_step_kernel<64>, grid, dim3(256), 0, s, a, j); break;
      case 128: hipLaunchKernelGGL(backsub_step_kernel<128>, grid, dim3(256), 0, s, a, j); break;
      default: BPF_THROW(BPF_E_UNSUPPORTED, "block width %d", a.S);
    }
  }
  BPF_CHECK_HIP(hipGetLastError());
}

void colmax_colmajor(hipStream_t s, int n, int p, const double *X, int ld, double *out) {
  if (p <= 0) return;
  hipLaunchKernelGGL((colmax_cm_kernel<double>), dim3(p), dim3(kB), 0, s, n, X, ld, out);
  BPF_CHECK_HIP(hipGetLastError());
}

int colstat_blocks(int n) { return std::max(1, std::min(1024, (n + 255) / 256)); }

template <typename T>
void colstat_rowmajor(hipStream_t s, int n, int p, const T *X, int ld, int mode, double *out, double *part) {
  if (p <= 0) return;
  fill_double(s, out, 0.0, p);
  if (n <= 0) return;
  BPF_REQUIRE(mode == 0 || part, "column sums need a partial-sum buffer");
  const int gx = colstat_blocks(n);
  dim3 grid(gx, (p + 63) / 64);
  hipLaunchKernelGGL((colstat_rm_kernel<T>), grid, dim3(kB), 0, s, n, p, X, ld, mode, out, part);
  BPF_CHECK_HIP(hipGetLastError());
  if (mode == 1) {
    hipLaunchKernelGGL(sum_parts_kernel, dim3(p), dim3(kB), 0, s, part, gx, (size_t)p, out);
    BPF_CHECK_HIP(hipGetLastError());
  }
}
template void colstat_rowmajor<double>(hipStream_t, int, int, const double *, int, int, double *, double *);
template void colstat_rowmajor<float>(hipStream_t, int, int, const float *, int, int, double *, double *);

void axpby_groups(hipStream_t s, int ng, size_t count, size_t stride, const double *x, double *y) {
  if (count == 0 || ng <= 0) return;
  hipLaunchKernelGGL(add_groups_kernel, dim3((unsigned)((count + kB - 1) / kB), ng), dim3(kB), 0, s, count, stride,
                     x, y);
  BPF_CHECK_HIP(hipGetLastError());
}

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
  if (b < groups[g].w) Y[e] *= scale[groups[g].slot + b];
}

__global__ __launch_bounds__(256) void scale_cols_kernel(int n, int p, double *X, int ld,
                                                        const double *scale, int invert) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * p) return;
  const int r = (int)(e / p), c = (int)(e % p);
  const double sc = scale[c];
  X[(size_t)r * ld + c] = invert ? X[(size_t)r * ld + c] / sc : X[(size_t)r * ld + c] * sc;
}

// Grid (row chunks, group, 64-column slice); threads: 64 columns x 4 rows.
// Mode 1 writes the block's sum (wavefronts combined in a fixed order) to
// part[blockIdx.x * nslots + slot], reduced by sum_parts_kernel.
template <typename T>
__global__ __launch_bounds__(kB) void colstat_groups_kernel(const GroupDesc *groups, const T *X, int ld,
                                                            int mode, double *out, double *part, int nslots) {
  __shared__ double red[kB / 64][64];
  const GroupDesc gd = groups[blockIdx.y];
  const int c = blockIdx.z * 64 + (threadIdx.x & 63);
  const int rend = gd.r0 + gd.nr;
  const int rstride = (kB / 64) * gridDim.x;
  double m = 0.0;
  if (c < gd.w) {
    for (int r = gd.r0 + blockIdx.x * (kB / 64) + (threadIdx.x >> 6); r < rend; r += rstride) {
      const double v = to_d(X[(size_t)r * ld + gd.c0 + c]);
      m = mode == 0 ? fmax(m, fabs(v)) : m + v * v;
    }
    if (mode == 0) atomic_max_nonneg(out + gd.slot + c, m);
  }
  if (mode == 0) return;
  red[threadIdx.x >> 6][threadIdx.x & 63] = m;
  __syncthreads();
  if (threadIdx.x < 64 && c < gd.w)
    part[(size_t)blockIdx.x * nslots + gd.slot + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                                      (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ __launch_bounds__(kB) void scale_groups_kernel(const GroupDesc *groups, double *X, int ld,
                                                          const double *scale, int invert) {
  const GroupDesc gd = groups[blockIdx.y];
  const int c = blockIdx.z * 64 + (threadIdx.x & 63);
  if (c >= gd.w) return;
  const int rend = gd.r0 + gd.nr;
  const int rstride = (kB / 64) * gridDim.x;
  const double sc = scale[gd.slot + c];
  for (int r = gd.r0 + blockIdx.x * (kB / 64) + (threadIdx.x >> 6); r < rend; r += rstride) {
    double *x = X + (size_t)r * ld + gd.c0 + c;
    *x = invert ? *x / sc : *x * sc;
  }
}

// Element e of group g's nr x w block: row e / w, column e % w (row-major
// side coalesced).
__global__ __launch_bounds__(kB) void pack_groups_kernel(const GroupDesc *groups, double *const *cm,
                                                         const int *ldc, double *rm, int ld, int to_rm) {
  const int g = blockIdx.y;
  const GroupDesc gd = groups[g];
  const size_t m = (size_t)gd.nr * gd.w;
  double *C = cm[g];
  const int lc = ldc[g];
  for (size_t e = (size_t)blockIdx.x * kB + threadIdx.x; e < m; e += (size_t)kB * gridDim.x) {
    const int i = (int)(e / gd.w), c = (int)(e % gd.w);
    double *r = rm + (size_t)(gd.r0 + i) * ld + gd.c0 + c;
    double *x = C + (size_t)c * lc + i;
    if (to_rm) *r = *x;
    else *x = *r;
  }
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

// Per-block partial dot products (fixed reduction tree) into part[blockIdx.x].
template <typename T>
__global__ __launch_bounds__(256) void dot_kernel(size_t m, const T *x, const T *y, double *part) {
  __shared__ double red[4];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (size_t)gridDim.x * 256)
    s += static_cast<double>(x[i]) * static_cast<double>(y[i]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
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

namespace {
dim3 group_grid(int ng, int gmax_rows, int wmax, int rows_per_wg) {
  const int gx = std::max(1, std::min(256, (gmax_rows + rows_per_wg - 1) / rows_per_wg));
  return dim3(gx, ng, (wmax + 63) / 64);
}
}  // namespace

int colstat_groups_blocks(int gmax_rows) { return group_grid(1, gmax_rows, 1, 256).x; }

template <typename T>
void colstat_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                    const T *X, int ld, int mode, double *out, int nslots, double *part) {
  if (nslots <= 0) return;
  fill_double(s, out, 0.0, nslots);
  if (ng <= 0 || gmax_rows <= 0 || wmax <= 0) return;
  BPF_REQUIRE(mode == 0 || part, "column sums need a partial-sum buffer");
  const dim3 grid = group_grid(ng, gmax_rows, wmax, 256);
  if (mode == 1) fill_double(s, part, 0.0, (size_t)grid.x * nslots);   // slots of empty columns
  hipLaunchKernelGGL((colstat_groups_kernel<T>), grid, dim3(kB), 0, s, groups, X, ld, mode, out, part,
                     nslots);
  BPF_CHECK_HIP(hipGetLastError());
  if (mode == 1) {
    hipLaunchKernelGGL(sum_parts_kernel, dim3(nslots), dim3(kB), 0, s, part, (int)grid.x, (size_t)nslots, out);
    BPF_CHECK_HIP(hipGetLastError());
  }
}
template void colstat_groups<double>(hipStream_t, const GroupDesc *, int, int, int, const double *, int,
                                     int, double *, int, double *);

void launch_scale_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                         double *X, int ld, const double *scale, bool invert) {
  if (ng <= 0 || gmax_rows <= 0 || wmax <= 0) return;
  hipLaunchKernelGGL(scale_groups_kernel, group_grid(ng, gmax_rows, wmax, 256), dim3(kB), 0, s, groups,
                     X, ld, scale, invert ? 1 : 0);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_pack_groups(hipStream_t s, const GroupDesc *groups, int ng, int gmax_rows, int wmax,
                        double *const *cm, const int *ldc, double *rm, int ld, bool to_rm) {
  if (ng <= 0 || gmax_rows <= 0 || wmax <= 0) return;
  const size_t m = (size_t)gmax_rows * wmax;
  const int gx = (int)std::max<size_t>(1, std::min<size_t>(1024, (m + kB * 4 - 1) / (kB * 4)));
  hipLaunchKernelGGL(pack_groups_kernel, dim3(gx, ng), dim3(kB), 0, s, groups, cm, ldc, rm, ld,
                     to_rm ? 1 : 0);
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
void launch_dot(hipStream_t s, size_t m, const T *x, const T *y, double *out, double *part) {
  fill_double(s, out, 0.0, 1);
  if (m == 0) return;
  const unsigned nb = (unsigned)std::min<size_t>(kDotBlocks, (m + 255) / 256);
  hipLaunchKernelGGL((dot_kernel<T>), dim3(nb), dim3(256), 0, s, m, x, y, part);
  BPF_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(kB), 0, s, part, (int)nb, (size_t)1, out);
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_dot<double>(hipStream_t, size_t, const double *, const double *, double *, double *);
template void launch_dot<float>(hipStream_t, size_t, const float *, const float *, double *, double *);

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
