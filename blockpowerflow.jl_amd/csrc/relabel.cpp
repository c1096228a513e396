// Locality relabelling of a system for the block-Krylov solvers.
//
// The synthetic and real power-flow Jacobians number their buses without
// regard to the network's geometry, so the rows an SpMM row or a triangular-
// solve row gathers lie anywhere in the n x p block: at 70,000 buses every
// gathered 512-byte row comes from beyond the gathering XCD's L2 (the SpMM
// reads 3.6x its algorithmic bytes).  A reverse Cuthill-McKee ordering of
// the symmetrized pattern (Cuthill & McKee 1969, George 1971) puts coupled
// unknowns near each other.  The solver then works in that relabelling:
//   - A's relabelled copy: storage row perm[i] holds row i with columns
//     perm[j], in the row's original entry order (so every product is
//     summed in the same order as before);
//   - relabelled views of an ILU(0)/IC(0) handle's triangular operators: the
//     same levels and rows, dependency indices mapped, each level's rows in
//     storage order, values gathered from the handle's own factor;
//   - right-hand sides and solutions permuted at the ABI's layout
//     conversion.
// The factorization is the handle's, in the original order (an ILU(0)
// computed in the relabelled order would be a different preconditioner with
// far more levels).
#include <algorithm>
#include <memory>

#include "bpf_internal.h"
#include "dense.h"
#include "ordering.h"

namespace bpf {

std::atomic<int> g_relabel_min{32768};

void launch_scatter_values(hipStream_t s, int nnz, const int *amap, const double *aval, double *fac);

CsrRelabel *relabel_csr(bpf_csr A, hipStream_t s) {
  if (!A->rl) {
    const int n = A->n, nnz = A->nnz;
    auto rl = std::make_unique<CsrRelabel>();
    const std::vector<int> order = rcm_order(symmetric_graph(n, A->h_rowptr, A->h_colidx));
    BPF_REQUIRE((int)order.size() == n, "ordering of %zu rows for %d", order.size(), n);
    rl->perm.assign(n, -1);
    for (int k = 0; k < n; ++k) rl->perm[order[k]] = k;
    std::vector<int> rp(n + 1, 0), ci(nnz), emap(std::max(1, nnz));
    for (int k = 0; k < n; ++k) {
      const int i = order[k];
      rp[k + 1] = rp[k] + (A->h_rowptr[i + 1] - A->h_rowptr[i]);
    }
    for (int k = 0; k < n; ++k) {
      const int i = order[k];
      int o = rp[k];
      for (int e = A->h_rowptr[i]; e < A->h_rowptr[i + 1]; ++e, ++o) {
        ci[o] = rl->perm[A->h_colidx[e]];
        emap[e] = o;
      }
    }
    auto *Ar = new bpf_csr_s();
    Ar->ctx = A->ctx;
    Ar->n = n;
    Ar->nnz = nnz;
    Ar->h_rowptr = rp;
    Ar->h_colidx = ci;
    upload(Ar->rowptr, rp, s);
    upload(Ar->colidx, ci, s);
    Ar->val.ensure(std::max(1, nnz) * sizeof(double));
    Ar->val32.ensure(std::max(1, nnz) * sizeof(float));
    upload(rl->order, order, s);
    upload(rl->emap, emap, s);
    rl->Ar = Ar;
    A->rl = rl.release();
  }
  CsrRelabel *rl = A->rl;
  if (rl->version != A->version) {
    bpf_csr_s *Ar = rl->Ar;
    launch_scatter_values(s, A->nnz, rl->emap.as<int>(), A->val.as<double>(), Ar->val.as<double>());
    if (A->nnz > 0)
      copy_cast<double, float>(s, 1, A->nnz, Ar->val.as<double>(), A->nnz, Ar->val32.as<float>(), A->nnz, 1.0);
    Ar->norm_inf = A->norm_inf;
    Ar->val32_valid = true;
    ++Ar->version;
    rl->version = A->version;
  }
  return rl;
}

bpf_prec relabel_prec(bpf_prec P, const CsrRelabel *rl, hipStream_t s) {
  if (P->kind != BPF_PREC_ILU0 && P->kind != BPF_PREC_IC0) return nullptr;
  if (P->rl && P->rl->key != rl) relabel_free(P);
  if (!P->rl) {
    auto pr = std::make_unique<PrecRelabel>();
    auto *Nr = new bpf_prec_s();
    Nr->ctx = P->ctx;
    Nr->kind = P->kind;
    Nr->n = P->n;
    Nr->nnz = P->nnz;
    Nr->max_row_nnz = P->max_row_nnz;
    TriWindow w;
    w.relabel = &rl->perm;
    const bool ilu = P->kind == BPF_PREC_ILU0;
    try {
      build_triop(Nr->lo, P->n, P->h_rowptr, P->h_colidx, ilu ? 0 : 2, s, w);
      build_triop(Nr->up, P->n, P->h_rowptr, P->h_colidx, ilu ? 1 : 3, s, w);
    } catch (...) {
      delete Nr;
      throw;
    }
    pr->key = rl;
    pr->Nr = Nr;
    P->rl = pr.release();
  }
  PrecRelabel *pr = P->rl;
  if (pr->fac_version != P->fac_version) {
    triop_refresh(pr->Nr->lo, P->fac.as<double>(), s);
    triop_refresh(pr->Nr->up, P->fac.as<double>(), s);
    pr->fac_version = P->fac_version;
  }
  return pr->Nr;
}

void relabel_free(bpf_csr A) {
  if (!A || !A->rl) return;
  if (A->rl->Ar) {
    spmm_forget(A->rl->Ar->rowptr.as<int>());
    delete A->rl->Ar;
  }
  delete A->rl;
  A->rl = nullptr;
}

void relabel_free(bpf_prec P) {
  if (!P || !P->rl) return;
  delete P->rl->Nr;
  delete P->rl;
  P->rl = nullptr;
}

}  // namespace bpf
