// Sparse kernels for gfx950: CSR SpMM over n x p row-major blocks, level
// scheduled multi-RHS triangular solves, and IKJ ILU / IC factorization.
//
// Layout.  Inside the library every n x p block is row-major (element (i, c)
// at X[i*ld + c]): a nonzero a_ij then multiplies the contiguous row X[j, :],
// so a row gather is one coalesced load of p elements instead of p strided
// loads.  A group of G lanes (G = power of two <= 64) owns one matrix row and
// spreads the p columns over its lanes (CPL columns per lane).
#include <hip/hip_ext.h>

#include "bpf_internal.h"
#include "profile.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace bpf {

namespace {

constexpr int kBlock = 256;

inline int next_pow2(int v) {
  int g = 1;
  while (g < v) g <<= 1;
  return g;
}

// Lanes per row and columns per lane for a p-wide block.
inline void lane_shape(int p, int &glog, int &cpl) {
  if (p <= 64) {
    int g = next_pow2(std::max(p, 1));
    glog = 0;
    while ((1 << glog) < g) ++glog;
    cpl = 1;
  } else {
    glog = 6;
    cpl = p <= 128 ? 2 : 4;
  }
}

// Rows of a row-major block start on cpl-element boundaries (vector loads).
template <typename TV>
inline bool vec_ok(const void *ptr, int ld, int cpl) {
  return ld % cpl == 0 && reinterpret_cast<uintptr_t>(ptr) % (sizeof(TV) * cpl) == 0;
}

template <typename T> __device__ __forceinline__ T ld_val(const void *v, int k);
template <> __device__ __forceinline__ double ld_val<double>(const void *v, int k) {
  return static_cast<const double *>(v)[k];
}
template <> __device__ __forceinline__ float ld_val<float>(const void *v, int k) {
  return static_cast<const float *>(v)[k];
}

// ------------------------------------------------------------------ SpMM
// CPL elements of a row-major block, CPL consecutive columns of one row,
// loaded or stored as one vector access.
template <typename TV, int CPL>
struct alignas(sizeof(TV) * CPL) VecT {
  TV v[CPL];
};

// Column of element m of a lane: with VEC the lane owns CPL consecutive
// columns (one vector load per gathered row); otherwise columns lane + m*G.
template <int CPL, bool VEC>
__device__ __forceinline__ int lane_col(int cbase, int lane, int m, int G) {
  return VEC ? cbase + lane * CPL + m : cbase + lane + m * G;
}

// Y = alpha A X + beta Y.  Each row's non-zeros are taken eight at a time,
// the last group padded (its column indices and values clamped to the row's
// last entry, the padded products skipped): one round trip for the indices
// and one for all gathers per eight non-zeros, the products added in k
// order, so the result does not depend on the unrolling, the launch shape or
// VEC.
template <typename TA, typename TV, int CPL, bool VEC>
__global__ __launch_bounds__(kBlock) void spmm_rm_kernel(
    int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
    const TA *__restrict__ val, int p, const TV *__restrict__ X, int ldx,
    TV *__restrict__ Y, int ldy, TV alpha, TV beta, int glog) {
  const int G = 1 << glog;
  const int rpb = kBlock >> glog;
  const int row = blockIdx.x * rpb + (threadIdx.x >> glog);
  const int lane = threadIdx.x & (G - 1);
  if (row >= n) return;
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int cbase = 0; cbase < p; cbase += G * CPL) {
    TV acc[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) acc[m] = TV(0);
    for (int k = k0; k < k1; k += 8) {
      const int last = k1 - 1 - k;   // >= 0
      int j[8];
      TV a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k + min(u, last);
        j[u] = colidx[kk];
        a[u] = static_cast<TV>(val[kk]);
      }
      TV xv[8][CPL];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const TV *xr = X + (size_t)j[u] * ldx;
        if (VEC) {
          const VecT<TV, CPL> w = *reinterpret_cast<const VecT<TV, CPL> *>(xr + cbase + lane * CPL);
#pragma unroll
          for (int m = 0; m < CPL; ++m) xv[u][m] = w.v[m];
        } else {
#pragma unroll
          for (int m = 0; m < CPL; ++m) {
            const int c = lane_col<CPL, VEC>(cbase, lane, m, G);
            xv[u][m] = c < p ? xr[c] : TV(0);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u <= last)
#pragma unroll
          for (int m = 0; m < CPL; ++m) acc[m] = fma(a[u], xv[u][m], acc[m]);
    }
    TV *yr = Y + (size_t)row * ldy;
    if (VEC) {
      VecT<TV, CPL> w;
      if (beta != TV(0)) w = *reinterpret_cast<const VecT<TV, CPL> *>(yr + cbase + lane * CPL);
#pragma unroll
      for (int m = 0; m < CPL; ++m) w.v[m] = beta == TV(0) ? alpha * acc[m] : fma(beta, w.v[m], alpha * acc[m]);
      *reinterpret_cast<VecT<TV, CPL> *>(yr + cbase + lane * CPL) = w;
    } else {
#pragma unroll
      for (int m = 0; m < CPL; ++m) {
        const int c = lane_col<CPL, VEC>(cbase, lane, m, G);
        if (c < p) yr[c] = beta == TV(0) ? alpha * acc[m] : fma(beta, yr[c], alpha * acc[m]);
      }
    }
  }
}

// Column-major SpMM used at the library boundary: one thread per (row, col).
template <typename TA, typename TV>
__global__ __launch_bounds__(kBlock) void spmm_cm_kernel(
    int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
    const TA *__restrict__ val, int p, const TV *__restrict__ X, int ldx,
    TV *__restrict__ Y, int ldy, TV alpha, TV beta) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || c >= p) return;
  const TV *x = X + (size_t)c * ldx;
  TV acc = TV(0);
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
    acc = fma(static_cast<TV>(val[k]), x[colidx[k]], acc);
  TV *y = Y + (size_t)c * ldy + i;
  *y = beta == TV(0) ? alpha * acc : fma(beta, *y, alpha * acc);
}

// ------------------------------------------------------ triangular solve
struct TriDev {
  const int *order, *ptr, *col, *lvl;
  const void *val, *diag;
};

// Solves one stored row t of a level-scheduled triangular operator for all
// p columns owned by this lane group.  Non-zeros are taken eight at a time
// as in the SpMM (the last group padded, padded products skipped), so the
// products are added in k order whatever the shape.
template <typename TF, typename TV, bool UNIT, int CPL, bool VEC>
__device__ __forceinline__ void tri_row(const TriDev &a, int t, int lane, int G,
                                        const TV *x, int ldx, TV *y, int ldy, int p) {
  const int i = a.order[t];
  const int k0 = a.ptr[t], k1 = a.ptr[t + 1];
  const TF *val = static_cast<const TF *>(a.val);
  TV dinv = TV(1);
  if (!UNIT) dinv = TV(1) / static_cast<TV>(static_cast<const TF *>(a.diag)[t]);
  for (int cbase = 0; cbase < p; cbase += G * CPL) {
    TV acc[CPL];
    if (VEC) {
      const VecT<TV, CPL> w = *reinterpret_cast<const VecT<TV, CPL> *>(x + (size_t)i * ldx + cbase + lane * CPL);
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] = w.v[m];
    } else {
#pragma unroll
      for (int m = 0; m < CPL; ++m) {
        const int c = lane_col<CPL, VEC>(cbase, lane, m, G);
        acc[m] = c < p ? x[(size_t)i * ldx + c] : TV(0);
      }
    }
    for (int k = k0; k < k1; k += 8) {
      const int last = k1 - 1 - k;
      int j[8];
      TV v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k + min(u, last);
        j[u] = a.col[kk];
        v[u] = static_cast<TV>(val[kk]);
      }
      TV av[8][CPL];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const TV *yu = y + (size_t)j[u] * ldy;
        if (VEC) {
          const VecT<TV, CPL> w = *reinterpret_cast<const VecT<TV, CPL> *>(yu + cbase + lane * CPL);
#pragma unroll
          for (int m = 0; m < CPL; ++m) av[u][m] = w.v[m];
        } else {
#pragma unroll
          for (int m = 0; m < CPL; ++m) {
            const int c = lane_col<CPL, VEC>(cbase, lane, m, G);
            av[u][m] = c < p ? yu[c] : TV(0);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u <= last)
#pragma unroll
          for (int m = 0; m < CPL; ++m) acc[m] = fma(-v[u], av[u][m], acc[m]);
    }
    if (VEC) {
      VecT<TV, CPL> w;
#pragma unroll
      for (int m = 0; m < CPL; ++m) w.v[m] = UNIT ? acc[m] : acc[m] * dinv;
      *reinterpret_cast<VecT<TV, CPL> *>(y + (size_t)i * ldy + cbase + lane * CPL) = w;
    } else {
#pragma unroll
      for (int m = 0; m < CPL; ++m) {
        const int c = lane_col<CPL, VEC>(cbase, lane, m, G);
        if (c < p) y[(size_t)i * ldy + c] = UNIT ? acc[m] : acc[m] * dinv;
      }
    }
  }
}

// One launch per wide level: rows [t0, t1) of the level-ordered list.
template <typename TF, typename TV, bool UNIT, int CPL, bool VEC>
__global__ __launch_bounds__(kBlock) void tri_level_kernel(TriDev a, int t0, int t1,
                                                           const TV *x, int ldx, TV *y,
                                                           int ldy, int p, int glog) {
  const int G = 1 << glog;
  const int rpb = kBlock >> glog;
  const int t = t0 + blockIdx.x * rpb + (threadIdx.x >> glog);
  if (t >= t1) return;
  tri_row<TF, TV, UNIT, CPL, VEC>(a, t, threadIdx.x & (G - 1), G, x, ldx, y, ldy, p);
}

// ------------------------------------------------------------- sync-free
// One launch for the whole operator (Liu et al., "A synchronization-free
// algorithm for parallel sparse triangular solves", 2016): one wavefront per
// row in level order; a row waits on the completion flags of the rows it
// depends on, then publishes its own.  Rows are dispatched in level order, so
// every row waited on belongs to a workgroup dispatched earlier.  The wait is
// bounded: when it runs out (or another wavefront already reported that) the
// status word is set and the row still publishes, so every wavefront exits.
constexpr int kSpinLimit = 1 << 20;

__device__ __forceinline__ bool sf_wait(const int *flags, const int *cols, int k0, int k1, int epoch,
                                        int lane, int *status) {
  // Relaxed polls (coherent at device scope, no cache invalidation); one
  // acquire fence once every dependency is complete.
  for (int kb = k0; kb < k1; kb += 64) {
    const int k = kb + lane;
    const int j = k < k1 ? cols[k] : -1;
    for (int spins = 0;; ++spins) {
      const bool ready =
          j < 0 || __hip_atomic_load(flags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
      if (__all(ready)) break;
      if (spins >= kSpinLimit ||
          __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        if (lane == 0) atomicOr(status, 1);
        return false;
      }
      __builtin_amdgcn_s_sleep(4);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

__device__ __forceinline__ void sf_publish(int *flags, int i, int epoch, int lane) {
  // every lane's stores of row i reach device scope before the flag does
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) __hip_atomic_store(flags + i, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename TF, typename TV, bool UNIT, int CPL, bool VEC>
__global__ __launch_bounds__(kBlock) void tri_syncfree_kernel(TriDev a, int n, const TV *x, int ldx, TV *y,
                                                              int ldy, int p, int *flags, int epoch,
                                                              int *status) {
  const int t = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (t >= n) return;
  const int lane = threadIdx.x & 63;
  if (sf_wait(flags, a.col, a.ptr[t], a.ptr[t + 1], epoch, lane, status))
    tri_row<TF, TV, UNIT, CPL, VEC>(a, t, lane, 64, x, ldx, y, ldy, p);
  sf_publish(flags, a.order[t], epoch, lane);
}

// One launch walks many narrow levels [l0, l1), with a workgroup barrier
// between levels.  Its workgroups split the columns into slabs of cw: each
// solves its own slab through every level (the columns of a triangular solve
// are independent, so no workgroup waits on another).
constexpr int kTailBlock = 1024;
// With one row per wavefront (G = 64) and a single column chunk (p <= 64
// CPL), a level with fewer rows than wavefronts splits each row's non-zeros
// over several wavefronts; the partial sums meet in LDS.  Deep elimination
// tails (one long row per level, as in fill-reducing LU orderings) then run
// at the speed of the whole workgroup instead of one wavefront.
// The segment's level bounds, row pointers, row numbers and diagonals are
// staged in LDS first when they fit (at most kTailRows rows and kTailLevels
// levels), and each row's right-hand side and diagonal are read before the
// barrier that precedes the row's final combination.
constexpr int kTailRows = 2048, kTailLevels = 1024;
template <typename TF, typename TV, bool UNIT, int CPL, bool VEC>
__global__ __launch_bounds__(kTailBlock) void tri_tail_kernel(TriDev a, int l0, int l1,
                                                              const TV *x, int ldx, TV *y,
                                                              int ldy, int p, int glog, int cw) {
  {
    const int c0 = blockIdx.x * cw;
    x += c0;
    y += c0;
    p = min(cw, p - c0);
  }
  constexpr int NW = kTailBlock / 64;
  __shared__ TV part[NW][64 * CPL];
  __shared__ int s_lvl[kTailLevels + 1];
  __shared__ int s_ptr[kTailRows + 1], s_ord[kTailRows];
  __shared__ TF s_dg[kTailRows];
  const TF *val = static_cast<const TF *>(a.val);
  const TF *dg = static_cast<const TF *>(a.diag);
  const int T0 = a.lvl[l0], T1 = a.lvl[l1];
  const bool staged = T1 - T0 <= kTailRows && l1 - l0 <= kTailLevels;
  if (staged) {
    for (int e = threadIdx.x; e <= l1 - l0; e += kTailBlock) s_lvl[e] = a.lvl[l0 + e] - T0;
    for (int e = threadIdx.x; e <= T1 - T0; e += kTailBlock) s_ptr[e] = a.ptr[T0 + e];
    for (int e = threadIdx.x; e < T1 - T0; e += kTailBlock) {
      s_ord[e] = a.order[T0 + e];
      if (!UNIT) s_dg[e] = dg[T0 + e];
    }
    __syncthreads();
  }
  auto lvl_at = [&](int l) { return staged ? T0 + s_lvl[l - l0] : a.lvl[l]; };
  auto ptr_at = [&](int t) { return staged ? s_ptr[t - T0] : a.ptr[t]; };
  auto ord_at = [&](int t) { return staged ? s_ord[t - T0] : a.order[t]; };
  auto dg_at = [&](int t) { return staged ? s_dg[t - T0] : dg[t]; };
  const int G = 1 << glog;
  const int ngroups = kTailBlock >> glog;
  const int gid = threadIdx.x >> glog;
  const int lane = threadIdx.x & (G - 1);
  const bool can_split = G == 64 && p <= 64 * CPL;
  for (int l = l0; l < l1; ++l) {
    const int t0 = lvl_at(l), t1 = lvl_at(l + 1);
    const int rows = t1 - t0;
    if (!can_split || rows >= NW) {
      for (int t = t0 + gid; t < t1; t += ngroups)
        tri_row<TF, TV, UNIT, CPL, VEC>(a, t, lane, G, x, ldx, y, ldy, p);
    } else {
      const int wpr = NW / rows;          // wavefronts per row
      const int r = gid / wpr, sub = gid % wpr;
      int i = 0;
      TV xr[CPL];
      TV dinv = TV(1);
      if (r < rows) {
        const int t = t0 + r;
        const int k0 = ptr_at(t), k1 = ptr_at(t + 1);
        if (sub == 0) {
          i = ord_at(t);
          if (!UNIT) dinv = TV(1) / static_cast<TV>(dg_at(t));
#pragma unroll
          for (int m = 0; m < CPL; ++m) {
            const int c = lane + m * 64;
            xr[m] = c < p ? x[(size_t)i * ldx + c] : TV(0);
          }
        }
        TV acc[CPL];
#pragma unroll
        for (int m = 0; m < CPL; ++m) acc[m] = TV(0);
        TV acc2[CPL];
#pragma unroll
        for (int m = 0; m < CPL; ++m) acc2[m] = TV(0);
        // four of the wavefront's non-zeros per step (every wpr-th of the
        // row): their column indices and values, then their y rows, loaded
        // together; the last step's missing entries read the row's last
        // entry and are skipped; two accumulator sets
        for (int k = k0 + sub; k < k1; k += 4 * wpr) {
          int j[4];
          TV v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int kk = min(k + u * wpr, k1 - 1);
            j[u] = a.col[kk];
            v[u] = static_cast<TV>(val[kk]);
          }
          const int nv = (k1 - k + wpr - 1) / wpr;   // entries of this step that exist
#pragma unroll
          for (int m = 0; m < CPL; ++m) {
            const int c = lane + m * 64;
            if (c < p) {
              TV av[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) av[u] = y[(size_t)j[u] * ldy + c];
              acc[m] = fma(v[0], av[0], acc[m]);
              if (nv > 1) acc2[m] = fma(v[1], av[1], acc2[m]);
              if (nv > 2) acc[m] = fma(v[2], av[2], acc[m]);
              if (nv > 3) acc2[m] = fma(v[3], av[3], acc2[m]);
            }
          }
        }
#pragma unroll
        for (int m = 0; m < CPL; ++m) part[gid][lane + m * 64] = acc[m] + acc2[m];
      }
      __syncthreads();
      if (r < rows && sub == 0) {
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
          const int c = lane + m * 64;
          if (c < p) {
            TV s = xr[m];
            for (int q = 0; q < wpr; ++q) s -= part[gid + q][c];
            y[(size_t)i * ldy + c] = UNIT ? s : s * dinv;
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------- factorization
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lds_find(const int *s_col, int lo, int hi, int j) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int c = s_col[mid];
    if (c == j) return mid;
    if (c < j) lo = mid + 1; else hi = mid;
  }
  return -1;
}

struct FacDev {
  const int *rowptr, *colidx, *diagpos, *order, *lvl;
  const PivInfo *pinfo;   // ILU / LU elimination plan: pivot rows per lower
  const int *map;         // entry, and the update positions of each pivot
  double *rdiag;          // 1 / U[k][k] per row k, written once row k is eliminated
  double *fac;
  int *status;
  int cap;  // LDS entries per wave
};

// IKJ elimination of row i (one wave).  Row i's pattern and values are staged
// in this wave's LDS slice; rows k < i referenced by it are final.
// kmax < i: only pivot rows k < kmax are eliminated (the rest of the row is
// left as the Schur complement of the leading kmax rows).
__device__ void ilu_row(const FacDev &f, int i, int lane, int *s_col, double *s_val, int kmax = 0x7fffffff) {
  // (row scalars are wave-uniform: one row per wavefront)
  i = __builtin_amdgcn_readfirstlane(i);
  const int base = __builtin_amdgcn_readfirstlane(f.rowptr[i]);
  const int len = __builtin_amdgcn_readfirstlane(f.rowptr[i + 1]) - base;
  const int dloc = __builtin_amdgcn_readfirstlane(f.diagpos[i]) - base;
  for (int q = lane; q < len; q += 64) {
    s_col[q] = f.colidx[base + q];
    s_val[q] = f.fac[base + q];
  }
  wave_sync();
  // Pivot rows are final, and the plan (pinfo: where pivot row k starts and
  // ends and where its update positions start in map; read 64 pivots at a
  // time, one entry per lane) comes from the pattern alone: the loads of the
  // next four pivots (diagonal, first 128 upper entries and their positions
  // in this row) are in flight while the current one is applied, and no
  // step searches this row's pattern.  The loop is unrolled by the four
  // slots, each refilled in place once applied, with every load issued at a
  // clamped address.  Each position of the row is updated at most once per
  // pivot, in pivot order, so the result is that of the IKJ recurrence.
  if (dloc > 0) {
    struct Piv {
      int ulen, moff, d, pos0, pos1;
      double ukk, rukk, val0, val1;
    };
    const PivInfo *pinfo = f.pinfo + base;
    int wbase = 0;
    PivInfo wpi = pinfo[min(lane, dloc - 1)];
    auto fetch = [&](int kk) {
      kk = min(kk, dloc - 1);
      int d, end, moff, k;
      if (kk >= wbase + 64) {
        // next 64 pivots (reached exactly at the window's end: kk == wbase)
        wbase += 64;
        wpi = pinfo[min(wbase + lane, dloc - 1)];
        d = __builtin_amdgcn_readlane(wpi.diag, 0);
        end = __builtin_amdgcn_readlane(wpi.end, 0);
        moff = __builtin_amdgcn_readlane(wpi.moff, 0);
        k = __builtin_amdgcn_readlane(wpi.row, 0);
      } else {
        d = __builtin_amdgcn_readlane(wpi.diag, kk - wbase);
        end = __builtin_amdgcn_readlane(wpi.end, kk - wbase);
        moff = __builtin_amdgcn_readlane(wpi.moff, kk - wbase);
        k = __builtin_amdgcn_readlane(wpi.row, kk - wbase);
      }
      Piv q;
      q.d = d;
      q.ulen = end - d - 1;
      q.moff = moff;
      q.ukk = f.fac[d];
      q.rukk = f.rdiag[k];
      // (an empty upper part reads entry 0: row k + 1's first value and the
      // next pivot's first position, or the map's final padding entry)
      const int last = max(q.ulen - 1, 0);
      const int u0 = min(lane, last), u1 = min(lane + 64, last);
      q.pos0 = f.map[moff + u0];
      q.val0 = f.fac[d + 1 + u0];
      q.pos1 = f.map[moff + u1];
      q.val1 = f.fac[d + 1 + u1];
      return q;
    };
    // Applies pivot kk (held in cur); false when k >= kmax ends the row.
    auto apply = [&](int kk, const Piv &cur) {
      const int k = s_col[kk];
      if (k >= kmax) return false;
      const double ukk = cur.ukk;
      const double l = s_val[kk] * cur.rukk;   // the multiplier: a product with the pivot's reciprocal
      if (lane < cur.ulen && cur.pos0 >= 0) s_val[cur.pos0] -= l * cur.val0;
      if (lane + 64 < cur.ulen && cur.pos1 >= 0) s_val[cur.pos1] -= l * cur.val1;
      for (int u = 128 + lane; u < cur.ulen; u += 64) {
        const int pos = f.map[cur.moff + u];
        if (pos >= 0) s_val[pos] -= l * f.fac[cur.d + 1 + u];
      }
      // (the updates only reach positions after kk)
      if (lane == 0) {
        s_val[kk] = l;
        if (ukk == 0.0) atomicMin(f.status, k);
      }
      wave_sync();
      return true;
    };
    Piv p0 = fetch(0), p1 = fetch(1), p2 = fetch(2), p3 = fetch(3);
    for (int kk = 0; kk < dloc; kk += 4) {
      if (!apply(kk, p0)) break;
      p0 = fetch(kk + 4);
      if (kk + 1 >= dloc || !apply(kk + 1, p1)) break;
      p1 = fetch(kk + 5);
      if (kk + 2 >= dloc || !apply(kk + 2, p2)) break;
      p2 = fetch(kk + 6);
      if (kk + 3 >= dloc || !apply(kk + 3, p3)) break;
      p3 = fetch(kk + 7);
    }
  }
  for (int q = lane; q < len; q += 64) f.fac[base + q] = s_val[q];
  if (lane == 0 && kmax > i) {
    if (s_val[dloc] == 0.0) atomicMin(f.status, i);
    f.rdiag[i] = 1.0 / s_val[dloc];
  }
  wave_sync();
}

// IC(0), row-oriented: l_ik = (a_ik - sum_{j<k} l_ij l_kj) / l_kk on the
// lower pattern, l_ii = sqrt(a_ii - sum_j l_ij^2).
__device__ void ic_row(const FacDev &f, int i, int lane, int *s_col, double *s_val) {
  const int base = f.rowptr[i];
  const int dloc = f.diagpos[i] - base;
  for (int q = lane; q <= dloc; q += 64) {
    s_col[q] = f.colidx[base + q];
    s_val[q] = f.fac[base + q];
  }
  wave_sync();
  for (int kk = 0; kk < dloc; ++kk) {
    const int k = s_col[kk];
    const int kb = f.rowptr[k], kd = f.diagpos[k];
    double part = 0.0;
    for (int jj = kb + lane; jj < kd; jj += 64) {
      const int pos = lds_find(s_col, 0, kk, f.colidx[jj]);
      if (pos >= 0) part += s_val[pos] * f.fac[jj];
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    const double lkk = f.fac[kd];
    wave_sync();
    if (lane == 0) s_val[kk] = (s_val[kk] - part) / lkk;
    wave_sync();
  }
  double sq = 0.0;
  for (int q = lane; q < dloc; q += 64) sq += s_val[q] * s_val[q];
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  wave_sync();
  if (lane == 0) {
    const double d = s_val[dloc] - sq;
    if (!(d > 0.0)) { atomicMin(f.status, i); s_val[dloc] = 1.0; }
    else s_val[dloc] = sqrt(d);
  }
  wave_sync();
  for (int q = lane; q <= dloc; q += 64) f.fac[base + q] = s_val[q];
  wave_sync();
}

template <bool IC>
__global__ void fac_level_kernel(FacDev f, int t0, int t1) {
  extern __shared__ unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = t0 + blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= t1) return;
  int *s_col = reinterpret_cast<int *>(smem) + (size_t)wave * f.cap;
  double *s_val = reinterpret_cast<double *>(smem + (size_t)(blockDim.x >> 6) * f.cap * sizeof(int)) +
                  (size_t)wave * f.cap;
  if (IC) ic_row(f, f.order[t], lane, s_col, s_val);
  else ilu_row(f, f.order[t], lane, s_col, s_val);
}

// Sync-free factorization: one wavefront per row in level order; row i waits
// for the rows of its strictly lower part (dependency list dptr/dcol of the
// lower operator), eliminates, then publishes.
template <bool IC>
__global__ void fac_syncfree_kernel(FacDev f, int n, const int *dptr, const int *dcol, int *flags,
                                    int epoch, int *status) {
  extern __shared__ unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= n) return;
  int *s_col = reinterpret_cast<int *>(smem) + (size_t)wave * f.cap;
  double *s_val = reinterpret_cast<double *>(smem + (size_t)(blockDim.x >> 6) * f.cap * sizeof(int)) +
                  (size_t)wave * f.cap;
  const int i = f.order[t];
  if (sf_wait(flags, dcol, dptr[t], dptr[t + 1], epoch, lane, status)) {
    if (IC) ic_row(f, i, lane, s_col, s_val);
    else ilu_row(f, i, lane, s_col, s_val);
  }
  sf_publish(flags, i, epoch, lane);
}

// Rows [r0, r1) (mutually independent) eliminated against pivot rows
// k < kmax, one wavefront per row.
__global__ void fac_partial_kernel(FacDev f, int r0, int r1, int kmax) {
  extern __shared__ unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = r0 + blockIdx.x * (blockDim.x >> 6) + wave;
  if (i >= r1) return;
  int *s_col = reinterpret_cast<int *>(smem) + (size_t)wave * f.cap;
  double *s_val = reinterpret_cast<double *>(smem + (size_t)(blockDim.x >> 6) * f.cap * sizeof(int)) +
                  (size_t)wave * f.cap;
  ilu_row(f, i, lane, s_col, s_val, kmax);
}

template <bool IC>
__global__ void fac_tail_kernel(FacDev f, int l0, int l1) {
  extern __shared__ unsigned char smem[];
  const int nw = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int *s_col = reinterpret_cast<int *>(smem) + (size_t)wave * f.cap;
  double *s_val = reinterpret_cast<double *>(smem + (size_t)nw * f.cap * sizeof(int)) +
                  (size_t)wave * f.cap;
  for (int l = l0; l < l1; ++l) {
    for (int t = f.lvl[l] + wave; t < f.lvl[l + 1]; t += nw) {
      if (IC) ic_row(f, f.order[t], lane, s_col, s_val);
      else ilu_row(f, f.order[t], lane, s_col, s_val);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ conversions
template <typename TD>
__global__ __launch_bounds__(kBlock) void cm_to_rm_kernel(int n, int p, const double *src,
                                                         int lds, TD *dst, int ldd) {
  // 32 x 32 tile transpose through LDS.
  __shared__ double tile[32][33];
  const int i0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, c = c0 + r;
    tile[r][tx] = (i < n && c < p) ? src[(size_t)c * lds + i] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, c = c0 + tx;
    if (i < n && c < p) dst[(size_t)i * ldd + c] = static_cast<TD>(tile[tx][r]);
  }
}

template <typename TS>
__global__ __launch_bounds__(kBlock) void rm_to_cm_kernel(int n, int p, const TS *src, int lds,
                                                         double *dst, int ldd) {
  __shared__ double tile[32][33];
  const int i0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, c = c0 + tx;
    tile[r][tx] = (i < n && c < p) ? static_cast<double>(src[(size_t)i * lds + c]) : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, c = c0 + r;
    if (i < n && c < p) dst[(size_t)c * ldd + i] = tile[tx][r];
  }
}

// Gathers level-ordered triangular values (and diagonals) from the factor.
__global__ __launch_bounds__(kBlock) void tri_gather_kernel(int m, const int *gather,
                                                           const double *fac, double *v64,
                                                           float *v32) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m) return;
  const double v = fac[gather[k]];
  v64[k] = v;
  v32[k] = static_cast<float>(v);
}

}  // namespace

// ------------------------------------------------------------------ hosts
std::atomic<bool> g_autotune{true};

namespace {

template <typename TV>
void spmm_shape(hipStream_t s, int n, const int *rowptr, const int *colidx, const void *val, bool val32,
                int p, const TV *X, int ldx, TV *Y, int ldy, TV a, TV b, int glog, int cpl) {
  const int rpb = kBlock >> glog;
  dim3 grid((n + rpb - 1) / rpb);
  // cpl > 100: vector loads of cpl - 100 consecutive columns per lane
  const bool vec = cpl > 100;
  if (vec) cpl -= 100;
#define BPF_SPMM(TA, C)                                                                                \
  do {                                                                                                  \
    if (vec)                                                                                            \
      hipLaunchKernelGGL((spmm_rm_kernel<TA, TV, C, true>), grid, dim3(kBlock), 0, s, n, rowptr, colidx,  \
                         static_cast<const TA *>(val), p, X, ldx, Y, ldy, a, b, glog);                  \
    else                                                                                                \
      hipLaunchKernelGGL((spmm_rm_kernel<TA, TV, C, false>), grid, dim3(kBlock), 0, s, n, rowptr, colidx, \
                         static_cast<const TA *>(val), p, X, ldx, Y, ldy, a, b, glog);                  \
  } while (0)
  if (val32) {
    if (cpl == 1) BPF_SPMM(float, 1); else if (cpl == 2) BPF_SPMM(float, 2); else BPF_SPMM(float, 4);
  } else {
    if (cpl == 1) BPF_SPMM(double, 1); else if (cpl == 2) BPF_SPMM(double, 2); else BPF_SPMM(double, 4);
  }
#undef BPF_SPMM
}

// SpMM launch-shape autotuner.  On the first product of a matrix (keyed by
// its row-pointer array), width and element types, each candidate shape --
// G lanes per row (a power of two <= 64) with CPL = 1, 2 or 4 columns per
// lane -- runs on the actual operands into a scratch block and is timed with
// HIP events; the fastest shape is cached for the process.
struct SpmmKey {
  const int *rowptr;
  int n, p, ldx, tv, val32;
  bool operator<(const SpmmKey &o) const {
    return std::tie(rowptr, n, p, ldx, tv, val32) < std::tie(o.rowptr, o.n, o.p, o.ldx, o.tv, o.val32);
  }
};
std::map<SpmmKey, std::pair<int, int>> g_spmm_shapes;
std::mutex g_spmm_mu;

template <typename TV>
std::pair<int, int> spmm_autotune(hipStream_t s, int n, const int *rowptr, const int *colidx, const void *val,
                                  bool val32, int p, const TV *X, int ldx) {
  std::vector<std::pair<int, int>> cand;
  for (int cpl : {1, 2, 4}) {
    int g = next_pow2((p + cpl - 1) / cpl);
    g = std::max(1, std::min(64, g));
    if (cpl > 1 && g * (cpl / 2) >= p) continue;   // the narrower CPL already covers p
    int glog = 0;
    while ((1 << glog) < g) ++glog;
    cand.push_back({glog, cpl});
    // the same shape with vector loads when its lanes cover the columns
    // exactly and the rows are aligned (the result is the same bit for bit)
    if (cpl > 1 && g * cpl == p && vec_ok<TV>(X, ldx, cpl) && vec_ok<TV>(X, p, cpl))
      cand.push_back({glog, cpl + 100});
  }
  if (cand.size() == 1) return cand[0];
  DevBuf scratch;
  scratch.ensure((size_t)n * p * sizeof(TV));
  hipEvent_t e0, e1;
  BPF_CHECK_HIP(hipEventCreate(&e0));
  BPF_CHECK_HIP(hipEventCreate(&e1));
  std::pair<int, int> best = cand[0];
  float best_ms = 1e30f;
  for (const auto &c : cand) {
    spmm_shape<TV>(s, n, rowptr, colidx, val, val32, p, X, ldx, scratch.as<TV>(), p, TV(1), TV(0), c.first, c.second);
    BPF_CHECK_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; ++r)
      spmm_shape<TV>(s, n, rowptr, colidx, val, val32, p, X, ldx, scratch.as<TV>(), p, TV(1), TV(0), c.first,
                     c.second);
    BPF_CHECK_HIP(hipEventRecord(e1, s));
    BPF_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    BPF_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

void spmm_forget(const int *rowptr) {
  std::lock_guard<std::mutex> lk(g_spmm_mu);
  for (auto it = g_spmm_shapes.begin(); it != g_spmm_shapes.end();)
    it = it->first.rowptr == rowptr ? g_spmm_shapes.erase(it) : std::next(it);
}

template <typename TV>
void launch_spmm(hipStream_t s, int n, int nnz, const int *rowptr, const int *colidx, const void *val,
                 bool val32, int p, const TV *X, int ldx, TV *Y, int ldy, double alpha,
                 double beta) {
  if (n <= 0 || p <= 0) return;
  int glog, cpl;
  lane_shape(p, glog, cpl);
  if (g_autotune) {
    const SpmmKey key{rowptr, n, p, ldx, (int)sizeof(TV), val32 ? 1 : 0};
    std::pair<int, int> shape;
    bool have = false;
    {
      std::lock_guard<std::mutex> lk(g_spmm_mu);
      auto it = g_spmm_shapes.find(key);
      if (it != g_spmm_shapes.end()) { shape = it->second; have = true; }
    }
    if (!have) {
      shape = spmm_autotune<TV>(s, n, rowptr, colidx, val, val32, p, X, ldx);
      std::lock_guard<std::mutex> lk(g_spmm_mu);
      g_spmm_shapes[key] = shape;
    }
    glog = shape.first;
    cpl = shape.second;
  } else if (cpl > 1 && (1 << glog) * cpl == p) {
    cpl += 100;   // vector loads (checked against the operands below)
  }
  // a vector shape needs rows on cpl-element boundaries in X and Y
  if (cpl > 100 && !(vec_ok<TV>(X, ldx, cpl - 100) && vec_ok<TV>(Y, ldy, cpl - 100))) cpl -= 100;
  const int tok = prof::begin(s);
  spmm_shape<TV>(s, n, rowptr, colidx, val, val32, p, X, ldx, Y, ldy, static_cast<TV>(alpha), static_cast<TV>(beta),
                 glog, cpl);
  BPF_CHECK_HIP(hipGetLastError());
  const double wv = val32 ? 4.0 : 8.0, wx = sizeof(TV);
  prof::end(tok, s, prof::SPMM, nnz * (wv + 4.0) + 4.0 * (n + 1) + 2.0 * n * p * wx,
            2.0 * nnz * (double)p);
}
template void launch_spmm<double>(hipStream_t, int, int, const int *, const int *, const void *, bool,
                                  int, const double *, int, double *, int, double, double);
template void launch_spmm<float>(hipStream_t, int, int, const int *, const int *, const void *, bool,
                                 int, const float *, int, float *, int, double, double);

void launch_spmm_colmajor(hipStream_t s, int n, const int *rowptr, const int *colidx,
                          const void *val, bool val32, int p, const void *X, int ldx, void *Y,
                          int ldy, double alpha, double beta, bool f32) {
  if (n <= 0 || p <= 0) return;
  dim3 grid((n + kBlock - 1) / kBlock, p);
  if (f32) {
    if (val32)
      hipLaunchKernelGGL((spmm_cm_kernel<float, float>), grid, dim3(kBlock), 0, s, n, rowptr, colidx,
                         static_cast<const float *>(val), p, static_cast<const float *>(X), ldx,
                         static_cast<float *>(Y), ldy, (float)alpha, (float)beta);
    else
      hipLaunchKernelGGL((spmm_cm_kernel<double, float>), grid, dim3(kBlock), 0, s, n, rowptr, colidx,
                         static_cast<const double *>(val), p, static_cast<const float *>(X), ldx,
                         static_cast<float *>(Y), ldy, (float)alpha, (float)beta);
  } else {
    hipLaunchKernelGGL((spmm_cm_kernel<double, double>), grid, dim3(kBlock), 0, s, n, rowptr, colidx,
                       static_cast<const double *>(val), p, static_cast<const double *>(X), ldx,
                       static_cast<double *>(Y), ldy, alpha, beta);
  }
  BPF_CHECK_HIP(hipGetLastError());
}

std::atomic<bool> g_syncfree{false};
std::atomic<int> g_tail_slab{64};
std::atomic<int> g_tail_rounds{2};

void triop_check(TriOp &op, hipStream_t s) {
  // only a sync-free launch since the last check can have set the word
  if (op.n <= 0 || op.epoch == op.checked) return;
  op.checked = op.epoch;
  int st = 0;
  BPF_CHECK_HIP(hipMemcpyAsync(&st, op.status.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  if (st != 0) {
    const int zero = 0;
    BPF_CHECK_HIP(hipMemcpy(op.status.ptr, &zero, sizeof(int), hipMemcpyHostToDevice));
    g_syncfree = false;
    BPF_THROW(BPF_E_HIP,
              "sync-free triangular solve/factorization: a dependency wait ran out; results of the "
              "last call are invalid (level-by-level launches are used from now on)");
  }
}

template <typename TV>
void trisolve(hipStream_t s, TriOp &op, int p, const TV *x, int ldx, TV *y, int ldy) {
  if (op.n <= 0 || p <= 0) return;
  constexpr bool F32 = std::is_same<TV, float>::value;
  TriDev a{op.order.as<int>(), op.ptr.as<int>(), op.col.as<int>(), op.lvl.as<int>(),
           F32 ? op.val32.ptr : op.val.ptr, F32 ? op.diag32.ptr : op.diag.ptr};
  if (g_syncfree) {
    // one wavefront per row, columns spread over the 64 lanes
    const int cpl = p <= 64 ? 1 : (p <= 128 ? 2 : 4);
    const int epoch = ++op.epoch;
    dim3 grid((op.n + kBlock / 64 - 1) / (kBlock / 64));
    const int tok = prof::begin(s);
#define BPF_SF(U, C)                                                                                      \
  hipLaunchKernelGGL((tri_syncfree_kernel<TV, TV, U, C, false>), grid, dim3(kBlock), 0, s, a, op.n, x, ldx, y, \
                     ldy, p, op.flags.as<int>(), epoch, op.status.as<int>())
    if (op.unit) {
      if (cpl == 1) BPF_SF(true, 1); else if (cpl == 2) BPF_SF(true, 2); else BPF_SF(true, 4);
    } else {
      if (cpl == 1) BPF_SF(false, 1); else if (cpl == 2) BPF_SF(false, 2); else BPF_SF(false, 4);
    }
#undef BPF_SF
    BPF_CHECK_HIP(hipGetLastError());
    const double rows = op.n, nz = (double)op.h_col.size(), w = sizeof(TV);
    prof::end(tok, s, prof::TRSV, nz * (w + 4.0) + rows * (12.0 + (op.unit ? 0.0 : w)) + 2.0 * rows * p * w,
              2.0 * nz * p);
    return;
  }
  int glog, cpl;
  lane_shape(p, glog, cpl);
  const int rpb = kBlock >> glog;
  const auto plan = make_plan(op, g_tail_rounds * (kTailBlock >> glog));
  // Vector loads when the lanes of a row cover its columns exactly and the
  // rows are aligned (each lane then owns cpl consecutive columns).
  const bool vec = cpl > 1 && ((1 << glog) * cpl == p) && vec_ok<TV>(x, ldx, cpl) && vec_ok<TV>(y, ldy, cpl);
  for (const auto &seg : plan) {
    if (seg.wide) {
      for (int l = seg.l0; l < seg.l1; ++l) {
        const int t0 = op.h_lvl[l], t1 = op.h_lvl[l + 1];
        dim3 grid((t1 - t0 + rpb - 1) / rpb);
        // short launches: kernel-stamped timer events (no packets of their own)
        const prof::ExtPair ep = prof::ext_begin();
#define BPF_TRI1(U, C, VV)                                                                              \
  do {                                                                                                   \
    if (ep.a)                                                                                            \
      hipExtLaunchKernelGGL((tri_level_kernel<TV, TV, U, C, VV>), grid, dim3(kBlock), 0, s, ep.a, ep.b,   \
                            0, a, t0, t1, x, ldx, y, ldy, p, glog);                                      \
    else                                                                                                 \
      hipLaunchKernelGGL((tri_level_kernel<TV, TV, U, C, VV>), grid, dim3(kBlock), 0, s, a, t0, t1, x,    \
                         ldx, y, ldy, p, glog);                                                          \
  } while (0)
#define BPF_TRI(U, C)                                                                                   \
  do {                                                                                                   \
    if (vec) BPF_TRI1(U, C, true);                                                                       \
    else BPF_TRI1(U, C, false);                                                                          \
  } while (0)
        if (op.unit) {
          if (cpl == 1) BPF_TRI(true, 1); else if (cpl == 2) BPF_TRI(true, 2); else BPF_TRI(true, 4);
        } else {
          if (cpl == 1) BPF_TRI(false, 1); else if (cpl == 2) BPF_TRI(false, 2); else BPF_TRI(false, 4);
        }
#undef BPF_TRI
#undef BPF_TRI1
        const double rows = t1 - t0, nz = op.h_ptr[t1] - op.h_ptr[t0], w = sizeof(TV);
        prof::ext_end(ep, prof::TRSV, nz * (w + 4.0) + rows * (12.0 + (op.unit ? 0.0 : w)) + 2.0 * rows * p * w,
                      2.0 * nz * p);
      }
    } else {
      const prof::ExtPair ep = prof::ext_begin();
      // More columns than a slab: one workgroup per slab of 64 columns (one
      // column per lane, one row per wavefront, the same per-column
      // arithmetic as one workgroup over all columns).
      const int slab = g_tail_slab;
      const bool split = slab > 0 && p > slab;
      const int nslab = split ? (p + slab - 1) / slab : 1;
      const int tglog = split ? 6 : glog;
      const int cw = split ? slab : p;
#define BPF_TAIL1(U, C, VV)                                                                                 \
  do {                                                                                                       \
    if (ep.a)                                                                                                \
      hipExtLaunchKernelGGL((tri_tail_kernel<TV, TV, U, C, VV>), dim3(nslab), dim3(kTailBlock), 0, s, ep.a,   \
                            ep.b, 0, a, seg.l0, seg.l1, x, ldx, y, ldy, p, tglog, cw);                       \
    else                                                                                                     \
      hipLaunchKernelGGL((tri_tail_kernel<TV, TV, U, C, VV>), dim3(nslab), dim3(kTailBlock), 0, s, a,         \
                         seg.l0, seg.l1, x, ldx, y, ldy, p, tglog, cw);                                      \
  } while (0)
#define BPF_TAIL(U, C)                                                                                      \
  do {                                                                                                       \
    if (vec) BPF_TAIL1(U, C, true);                                                                          \
    else BPF_TAIL1(U, C, false);                                                                             \
  } while (0)
      if (split) {
        if (op.unit) BPF_TAIL1(true, 1, false); else BPF_TAIL1(false, 1, false);
      } else if (op.unit) {
        if (cpl == 1) BPF_TAIL(true, 1); else if (cpl == 2) BPF_TAIL(true, 2); else BPF_TAIL(true, 4);
      } else {
        if (cpl == 1) BPF_TAIL(false, 1); else if (cpl == 2) BPF_TAIL(false, 2); else BPF_TAIL(false, 4);
      }
#undef BPF_TAIL
#undef BPF_TAIL1
      const int t0 = op.h_lvl[seg.l0], t1 = op.h_lvl[seg.l1];
      const double rows = t1 - t0, nz = op.h_ptr[t1] - op.h_ptr[t0], w = sizeof(TV);
      prof::ext_end(ep, prof::TRSV, nz * (w + 4.0) + rows * (12.0 + (op.unit ? 0.0 : w)) + 2.0 * rows * p * w,
                    2.0 * nz * p);
    }
  }
  BPF_CHECK_HIP(hipGetLastError());
}
template void trisolve<double>(hipStream_t, TriOp &, int, const double *, int, double *, int);
template void trisolve<float>(hipStream_t, TriOp &, int, const float *, int, float *, int);

static void launch_factor(bool ic, hipStream_t s, TriOp &lower, const int *rowptr,
                          const int *colidx, const int *diagpos, const ElimPlan *eplan, double *fac,
                          int max_row_nnz, int *d_status, double bytes) {
  BPF_REQUIRE(ic || (eplan && eplan->ready()), "LU elimination without its elimination plan");
  const PivInfo *pinfo = eplan ? eplan->pinfo.as<PivInfo>() : nullptr;
  const int *map = eplan ? eplan->map.as<int>() : nullptr;
  const int cap = std::max(64, max_row_nnz);
  const size_t per_wave = (size_t)cap * (sizeof(int) + sizeof(double));
  if (per_wave > 64 * 1024)
    BPF_THROW(BPF_E_UNSUPPORTED, "factor row with %d entries exceeds the LDS capacity", max_row_nnz);
  int wpb = (int)std::min<size_t>(4, std::max<size_t>(1, (64 * 1024) / per_wave));
  double *rdiag = eplan ? eplan->rdiag.as<double>() : nullptr;
  FacDev f{rowptr, colidx, diagpos, lower.order.as<int>(), lower.lvl.as<int>(), pinfo, map, rdiag, fac, d_status,
           cap};
  const int tok = prof::begin(s);
  if (g_syncfree && lower.n > 0) {
    const size_t shm = per_wave * wpb;
    const int epoch = ++lower.epoch;
    dim3 grid((lower.n + wpb - 1) / wpb);
    if (ic)
      hipLaunchKernelGGL((fac_syncfree_kernel<true>), grid, dim3(64 * wpb), shm, s, f, lower.n,
                         lower.ptr.as<int>(), lower.col.as<int>(), lower.flags.as<int>(), epoch,
                         lower.status.as<int>());
    else
      hipLaunchKernelGGL((fac_syncfree_kernel<false>), grid, dim3(64 * wpb), shm, s, f, lower.n,
                         lower.ptr.as<int>(), lower.col.as<int>(), lower.flags.as<int>(), epoch,
                         lower.status.as<int>());
    BPF_CHECK_HIP(hipGetLastError());
    prof::end(tok, s, prof::FACTOR, bytes, 0.0);
    return;
  }
  // Levels of at most 8 rows go to the single-workgroup tail kernel (one
  // round of its 16 wavefronts); wider ones get a launch each.  On the
  // 70,000-bus LU pattern 8 rows beat 32 (the 17-32-row levels took two
  // rounds there): 206 -> 197 ms of factorization per 32 value sets.
  const auto plan = make_plan(lower, 8);
  for (const auto &seg : plan) {
    if (seg.wide) {
      const size_t shm = per_wave * wpb;
      for (int l = seg.l0; l < seg.l1; ++l) {
        const int t0 = lower.h_lvl[l], t1 = lower.h_lvl[l + 1];
        dim3 grid((t1 - t0 + wpb - 1) / wpb);
        if (ic) hipLaunchKernelGGL((fac_level_kernel<true>), grid, dim3(64 * wpb), shm, s, f, t0, t1);
        else hipLaunchKernelGGL((fac_level_kernel<false>), grid, dim3(64 * wpb), shm, s, f, t0, t1);
      }
    } else {
      int tw = (int)std::min<size_t>(16, std::max<size_t>(1, (64 * 1024) / per_wave));
      const size_t shm = per_wave * tw;
      if (ic) hipLaunchKernelGGL((fac_tail_kernel<true>), dim3(1), dim3(64 * tw), shm, s, f, seg.l0, seg.l1);
      else hipLaunchKernelGGL((fac_tail_kernel<false>), dim3(1), dim3(64 * tw), shm, s, f, seg.l0, seg.l1);
    }
  }
  BPF_CHECK_HIP(hipGetLastError());
  prof::end(tok, s, prof::FACTOR, bytes, 0.0);
}

void launch_partial_factor(hipStream_t s, int r0, int r1, int kmax, const int *rowptr, const int *colidx,
                           const int *diagpos, const ElimPlan &plan, double *fac, int max_row_nnz,
                           int *d_status, double bytes) {
  if (r1 <= r0) return;
  BPF_REQUIRE(plan.ready(), "LU elimination without its elimination plan");
  const int cap = std::max(64, max_row_nnz);
  const size_t per_wave = (size_t)cap * (sizeof(int) + sizeof(double));
  if (per_wave > 64 * 1024)
    BPF_THROW(BPF_E_UNSUPPORTED, "factor row with %d entries exceeds the LDS capacity", max_row_nnz);
  const int wpb = (int)std::min<size_t>(4, std::max<size_t>(1, (64 * 1024) / per_wave));
  FacDev f{rowptr, colidx, diagpos, nullptr, nullptr, plan.pinfo.as<PivInfo>(), plan.map.as<int>(),
           plan.rdiag.as<double>(), fac, d_status, cap};
  const int tok = prof::begin(s);
  hipLaunchKernelGGL(fac_partial_kernel, dim3((r1 - r0 + wpb - 1) / wpb), dim3(64 * wpb), per_wave * wpb, s, f,
                     r0, r1, kmax);
  BPF_CHECK_HIP(hipGetLastError());
  prof::end(tok, s, prof::FACTOR, bytes, 0.0);
}

void launch_ilu_factor(hipStream_t s, TriOp &lower, int, const int *rowptr,
                       const int *colidx, const int *diagpos, const ElimPlan &plan, double *fac,
                       int max_row_nnz, int *d_status, double bytes) {
  launch_factor(false, s, lower, rowptr, colidx, diagpos, &plan, fac, max_row_nnz, d_status, bytes);
}

void launch_ic_factor(hipStream_t s, TriOp &lower, int, const int *rowptr,
                      const int *colidx, const int *diagpos, double *fac, int max_row_nnz,
                      int *d_status, double bytes) {
  launch_factor(true, s, lower, rowptr, colidx, diagpos, nullptr, fac, max_row_nnz, d_status, bytes);
}

void launch_cm_to_rm(hipStream_t s, int n, int p, const double *src, int lds, void *dst, int ldd,
                     bool dst32) {
  if (n <= 0 || p <= 0) return;
  dim3 grid((n + 31) / 32, (p + 31) / 32);
  if (dst32)
    hipLaunchKernelGGL((cm_to_rm_kernel<float>), grid, dim3(kBlock), 0, s, n, p, src, lds,
                       static_cast<float *>(dst), ldd);
  else
    hipLaunchKernelGGL((cm_to_rm_kernel<double>), grid, dim3(kBlock), 0, s, n, p, src, lds,
                       static_cast<double *>(dst), ldd);
  BPF_CHECK_HIP(hipGetLastError());
}

void launch_rm_to_cm(hipStream_t s, int n, int p, const void *src, int lds, bool src32,
                     double *dst, int ldd) {
  if (n <= 0 || p <= 0) return;
  dim3 grid((n + 31) / 32, (p + 31) / 32);
  if (src32)
    hipLaunchKernelGGL((rm_to_cm_kernel<float>), grid, dim3(kBlock), 0, s, n, p,
                       static_cast<const float *>(src), lds, dst, ldd);
  else
    hipLaunchKernelGGL((rm_to_cm_kernel<double>), grid, dim3(kBlock), 0, s, n, p,
                       static_cast<const double *>(src), lds, dst, ldd);
  BPF_CHECK_HIP(hipGetLastError());
}

void triop_refresh(TriOp &op, const double *d_fac, hipStream_t s) {
  const int m = (int)op.h_col.size();
  if (m > 0) {
    hipLaunchKernelGGL(tri_gather_kernel, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m,
                       op.gather.as<int>(), d_fac, op.val.as<double>(), op.val32.as<float>());
  }
  if (!op.unit && op.n > 0) {
    hipLaunchKernelGGL(tri_gather_kernel, dim3((op.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       op.n, op.dgather.as<int>(), d_fac, op.diag.as<double>(),
                       op.diag32.as<float>());
  }
  BPF_CHECK_HIP(hipGetLastError());
}

}  // namespace bpf

// ------------------------------------------------- permutations, scatter
namespace bpf {
namespace {

__global__ __launch_bounds__(kBlock) void scatter_values_kernel(int nnz, const int *amap,
                                                               const double *aval, double *fac) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < nnz) fac[amap[k]] = aval[k];
}

// dst (row-major, row k) = src (column-major) row rowmap[k]
template <typename TD>
__global__ __launch_bounds__(kBlock) void perm_cm_to_rm_kernel(int n, int p, const double *src, int lds,
                                                              const int *rowmap, TD *dst, int ldd) {
  const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (size_t)n * p) return;
  const int c = (int)(e / n), k = (int)(e % n);   // consecutive threads: consecutive k
  dst[(size_t)k * ldd + c] = static_cast<TD>(src[(size_t)c * lds + rowmap[k]]);
}

// dst (column-major) row rowmap[k] = src (row-major) row k
template <typename TS>
__global__ __launch_bounds__(kBlock) void perm_rm_to_cm_kernel(int n, int p, const TS *src, int lds,
                                                              const int *rowmap, double *dst, int ldd) {
  const size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (size_t)n * p) return;
  const int c = (int)(e / n), k = (int)(e % n);
  dst[(size_t)c * ldd + rowmap[k]] = static_cast<double>(src[(size_t)k * lds + c]);
}

// row-major n x p: gather dst[k] = src[map[k]] or scatter dst[map[k]] = src[k]
template <typename TS, typename TD>
// Row-major n x p rows moved by a row map, one wavefront per row: mode 0
// dst[k] = src[map[k]], 1 dst[map[k]] = src[k], 2 dst[map[k]] += src[k].
__global__ __launch_bounds__(kBlock) void perm_rows_kernel(int n, int p, const TS *src, const int *map,
                                                          TD *dst, int mode) {
  const int k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (k >= n) return;
  const int lane = threadIdx.x & 63;
  const int mk = map[k];
  const TS *sr = src + (size_t)(mode == 0 ? mk : k) * p;
  TD *dr = dst + (size_t)(mode == 0 ? k : mk) * p;
  for (int c = lane; c < p; c += 64) {
    const TD v = static_cast<TD>(sr[c]);
    dr[c] = mode == 2 ? dr[c] + v : v;
  }
}

}  // namespace

void launch_scatter_values(hipStream_t s, int nnz, const int *amap, const double *aval, double *fac) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(scatter_values_kernel, dim3((nnz + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nnz,
                     amap, aval, fac);
  BPF_CHECK_HIP(hipGetLastError());
}

template <typename TD>
void launch_perm_cm_to_rm(hipStream_t s, int n, int p, const double *src, int lds, const int *rowmap,
                          TD *dst, int ldd) {
  const size_t m = (size_t)n * p;
  if (m == 0) return;
  hipLaunchKernelGGL((perm_cm_to_rm_kernel<TD>), dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, s, n, p, src, lds, rowmap, dst, ldd);
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_perm_cm_to_rm<double>(hipStream_t, int, int, const double *, int, const int *, double *, int);
template void launch_perm_cm_to_rm<float>(hipStream_t, int, int, const double *, int, const int *, float *, int);

template <typename TS>
void launch_perm_rm_to_cm(hipStream_t s, int n, int p, const TS *src, int lds, const int *rowmap,
                          double *dst, int ldd) {
  const size_t m = (size_t)n * p;
  if (m == 0) return;
  hipLaunchKernelGGL((perm_rm_to_cm_kernel<TS>), dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, s, n, p, src, lds, rowmap, dst, ldd);
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_perm_rm_to_cm<double>(hipStream_t, int, int, const double *, int, const int *, double *, int);
template void launch_perm_rm_to_cm<float>(hipStream_t, int, int, const float *, int, const int *, double *, int);

template <typename TS, typename TD>
void launch_perm_rows_rm(hipStream_t s, int n, int p, const TS *src, const int *rowmap, TD *dst,
                         bool scatter, bool add) {
  if (n <= 0 || p <= 0) return;
  const int rpb = kBlock / 64;
  hipLaunchKernelGGL((perm_rows_kernel<TS, TD>), dim3((unsigned)((n + rpb - 1) / rpb)), dim3(kBlock), 0, s, n, p,
                     src, rowmap, dst, add ? 2 : (scatter ? 1 : 0));
  BPF_CHECK_HIP(hipGetLastError());
}
template void launch_perm_rows_rm<double, float>(hipStream_t, int, int, const double *, const int *, float *, bool,
                                                  bool);
template void launch_perm_rows_rm<float, double>(hipStream_t, int, int, const float *, const int *, double *, bool,
                                                  bool);
template void launch_perm_rows_rm<double, double>(hipStream_t, int, int, const double *, const int *, double *, bool,
                                                   bool);

}  // namespace bpf
