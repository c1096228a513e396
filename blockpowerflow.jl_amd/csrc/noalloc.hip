// Per-RHS backward error with no device allocation: the variant of
// bpf_backward_error for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).  Its only scratch is the file-scope device
// array below (3 x kNoAllocMaxRhs doubles per GPU, held until the process
// ends); the residual B - A X is never stored, each workgroup folds its rows'
// |r|, |x| and |b| maxima into the array with atomics.
#include <mutex>

#include "bpf_internal.h"

namespace {
constexpr int kNoAllocMaxRhs = 4096;
}  // namespace

// [0, p): max |B - A X| per column, [p, 2p): max |X|, [2p, 3p): max |B|,
// as the bit patterns of non-negative doubles (which order like integers).
// Host code uses the per-device address from hipGetSymbolAddress, never
// this name.
static __device__ __align__(256) unsigned long long scratch[3 * kNoAllocMaxRhs];

namespace bpf {
namespace {

constexpr int kMaxDevices = 64;
constexpr int kNaBlock = 256;

std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

__global__ __launch_bounds__(kNaBlock) void na_clear_kernel(unsigned long long *s, int m) {
  const int i = blockIdx.x * kNaBlock + threadIdx.x;
  if (i < m) s[i] = 0ull;
}

// One workgroup: kNaBlock consecutive rows of one column c (blockIdx.y).
__global__ __launch_bounds__(kNaBlock) void na_resid_max_kernel(int n, int p, const int *rowptr,
                                                                const int *colidx, const double *val,
                                                                const double *X, int ldx, const double *B,
                                                                int ldb, unsigned long long *s) {
  __shared__ double red[3][kNaBlock / 64];
  const int c = blockIdx.y;
  const int i = blockIdx.x * kNaBlock + threadIdx.x;
  double r = 0.0, xa = 0.0, ba = 0.0;
  if (i < n) {
    const double *x = X + (size_t)c * ldx;
    double acc = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += val[k] * x[colidx[k]];
    const double b = B[(size_t)c * ldb + i];
    r = fabs(b - acc);
    xa = fabs(x[i]);
    ba = fabs(b);
  }
  for (int off = 32; off > 0; off >>= 1) {
    r = fmax(r, __shfl_xor(r, off));
    xa = fmax(xa, __shfl_xor(xa, off));
    ba = fmax(ba, __shfl_xor(ba, off));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = r;
    red[1][wave] = xa;
    red[2][wave] = ba;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double m = 0.0;
    for (int w = 0; w < kNaBlock / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
    atomicMax(s + (size_t)threadIdx.x * p + c, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ __launch_bounds__(kNaBlock) void na_finish_kernel(int p, double norm_a, const unsigned long long *s,
                                                             double *berr) {
  const int c = blockIdx.x * kNaBlock + threadIdx.x;
  if (c >= p) return;
  const double r = __longlong_as_double((long long)s[c]);
  const double x = __longlong_as_double((long long)s[p + c]);
  const double b = __longlong_as_double((long long)s[2 * p + c]);
  const double den = norm_a * x + b;
  berr[c] = den > 0.0 ? r / den : (r > 0.0 ? INFINITY : 0.0);
}

}  // namespace

// Resolves the device address of `scratch` on `device` (current), once per
// device.  Called by bpf_ctx_create after it has made the device current.
void noalloc_setup(int device) {
  BPF_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    BPF_THROW(BPF_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
}

}  // namespace bpf

using namespace bpf;

extern "C" int bpf_backward_error_noalloc(bpf_ctx ctx, bpf_csr A, int p, const double *d_X, int ldx,
                                          const double *d_B, int ldb, double *d_berr) {
  return guard([&] {
    BPF_REQUIRE(ctx && A, "null handle");
    if (p <= 0) return;
    BPF_REQUIRE(d_X && d_B && d_berr, "null block");
    if (p > kNoAllocMaxRhs)
      BPF_THROW(BPF_E_UNSUPPORTED, "bpf_backward_error_noalloc takes at most %d right-hand sides (p = %d)",
                kNoAllocMaxRhs, p);
    const int n = A->n;
    BPF_REQUIRE(ldx >= n && ldb >= n, "leading dimension < n");
    void *sp = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      sp = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
    }
    BPF_REQUIRE(sp, "scratch of device %d not set up", ctx->device);
    auto *s = static_cast<unsigned long long *>(sp);
    hipStream_t st = ctx->stream;
    const int m = 3 * p;
    hipLaunchKernelGGL(na_clear_kernel, dim3((m + kNaBlock - 1) / kNaBlock), dim3(kNaBlock), 0, st, s, m);
    BPF_CHECK_HIP(hipGetLastError());
    if (n > 0) {
      hipLaunchKernelGGL(na_resid_max_kernel, dim3((n + kNaBlock - 1) / kNaBlock, p), dim3(kNaBlock), 0, st, n,
                         p, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.as<double>(), d_X, ldx, d_B, ldb, s);
      BPF_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(na_finish_kernel, dim3((p + kNaBlock - 1) / kNaBlock), dim3(kNaBlock), 0, st, p,
                       A->norm_inf, s, d_berr);
    BPF_CHECK_HIP(hipGetLastError());
  });
}
