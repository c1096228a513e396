// Fixed-pattern LU refactorization (CUSOLVERRF replacement) -- filled in below.
#include "krylov.h"

using namespace bpf;

extern "C" {
int bpf_rf_create(bpf_ctx, bpf_csr, int, double, bpf_rf *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
int bpf_rf_refactor(bpf_rf, bpf_csr) { return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); }); }
int bpf_rf_solve(bpf_rf, int, int, const double *, int, double *, int) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
int bpf_rf_solve_refined(bpf_rf, bpf_csr, int, const double *, int, double *, int, int, double *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
int bpf_rf_info(bpf_rf, int *, int *, int *, int *) { return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); }); }
int bpf_rf_get_factor(bpf_rf, int32_t *, int32_t *, int32_t *, int32_t *, double *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
int bpf_rf_destroy(bpf_rf) { return 0; }
int bpf_rf_batch_create(bpf_ctx, bpf_csr, int, int, double, bpf_rf *) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
int bpf_rf_batch_refactor(bpf_rf, const double *, int) { return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); }); }
int bpf_rf_batch_solve(bpf_rf, const double *, int, double *, int) {
  return guard([&] { BPF_THROW(BPF_E_UNSUPPORTED, "refactorization not built"); });
}
}
