// Batched systems -- filled in below.
#include "krylov.h"

using namespace bpf;

extern "C" {
int bpf_batch_create(bpf_ctx, int, const int32_t *, const int32_t *, const int32_t *const *,
                     const int32_t *const *, const double *const *, const int32_t *, bpf_batch *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "batched solver not built"); });
}
int bpf_batch_solve(bpf_batch, const double *const *, double *const *, const bpf_solver_params *,
                    bpf_solver_stats *, int32_t *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "batched solver not built"); });
}
int bpf_batch_destroy(bpf_batch) { return 0; }
}
