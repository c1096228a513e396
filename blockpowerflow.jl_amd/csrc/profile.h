// In-process kernel timing: HIP events around instrumented launches, with the
// algorithmic bytes / flops of each launch, summed per kernel family.
#pragma once

#include <hip/hip_runtime.h>

namespace bpf {
namespace prof {

// PROJ / GEMM: block widths of 64 and more (matrix-core bound); PROJ_NARROW /
// GEMM_NARROW: widths up to 32 (bandwidth bound), e.g. an inner solve's
// 16-column groups.
enum Family {
  SPMM = 0, TRSV = 1, PROJ = 2, GEMM = 3, FACTOR = 4, DENSE_LU = 5, PROJ_NARROW = 6, GEMM_NARROW = 7, NFAM = 8
};

bool enabled();
// Records a start event on s when profiling is on; returns a token for end().
int begin(hipStream_t s);
void end(int token, hipStream_t s, Family f, double bytes, double flops);
// A start / stop event pair for hipExtLaunchKernelGGL, which stamps them
// with the kernel's own start and end (no event packets of their own, whose
// dispatch a separate pair would add to a short launch); both null when
// profiling is off.  ext_end() registers the pair like end().
struct ExtPair {
  hipEvent_t a = nullptr, b = nullptr;
};
ExtPair ext_begin();
void ext_end(const ExtPair &p, Family f, double bytes, double flops);

}  // namespace prof
}  // namespace bpf
