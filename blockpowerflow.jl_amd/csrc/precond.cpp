// Preconditioner application on row-major n x p blocks: the one place that
// turns "apply the preconditioner" (mulorldiv!, src/block_utils.jl:113) into
// kernel launches for each kind of handle.
#include "bpf_internal.h"

namespace bpf {

template <typename TV>
void prec_apply_rm(hipStream_t s, bpf_prec P, int p, const TV *x, int ldx, TV *y, int ldy) {
  if (P->kind == BPF_PREC_BLOCK_JACOBI) {
    bj_apply<TV>(s, P, p, x, ldx, y, ldy);
    return;
  }
  // ILU(0): y = U^{-1} L^{-1} x (src/preconditioners_nvidia.jl:55-58);
  // IC(0): y = L^{-T} L^{-1} x (:19-22).
  trisolve<TV>(s, P->lo, p, x, ldx, y, ldy);
  trisolve<TV>(s, P->up, p, y, ldy, y, ldy);
}

void prec_check(bpf_prec P, hipStream_t s) {
  if (P->kind == BPF_PREC_BLOCK_JACOBI) return;
  triop_check(P->lo, s);
  triop_check(P->up, s);
}

template void prec_apply_rm<double>(hipStream_t, bpf_prec, int, const double *, int, double *, int);
template void prec_apply_rm<float>(hipStream_t, bpf_prec, int, const float *, int, float *, int);

}  // namespace bpf
