// Device side of the host analysis: level-ordered triangular operators and
// elimination plans uploaded for the solve and factorization kernels.
#include <algorithm>
#include <numeric>

#include "bpf_internal.h"

namespace bpf {

void build_elim_plan(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
                     const std::vector<int> &diagpos, hipStream_t s, ElimPlan &out) {
  std::vector<PivInfo> pinfo;
  std::vector<int> map;
  elim_plan_host(n, rowptr, colidx, diagpos, pinfo, map);
  upload(out.pinfo, pinfo, s);
  upload(out.map, map, s);
  out.rdiag.ensure(std::max(1, n) * sizeof(double));
}

std::vector<Seg> make_plan(const TriOp &op, int tail_rows) {
  std::vector<Seg> plan;
  int l = 0;
  while (l < op.nlevels) {
    const bool wide = (op.h_lvl[l + 1] - op.h_lvl[l]) > tail_rows;
    int e = l + 1;
    while (e < op.nlevels && ((op.h_lvl[e + 1] - op.h_lvl[e]) > tail_rows) == wide) ++e;
    plan.push_back({l, e, wide});
    l = e;
  }
  return plan;
}

void build_triop(TriOp &op, int n, const std::vector<int> &rowptr,
                 const std::vector<int> &colidx, int kind, hipStream_t s, const TriWindow &win) {
  analyze_triop(op, n, rowptr, colidx, kind, win);
  upload(op.order, op.h_order, s);
  upload(op.ptr, op.h_ptr, s);
  upload(op.col, op.h_col, s);
  upload(op.gather, op.h_gather, s);
  upload(op.dgather, op.h_dgather, s);
  upload(op.lvl, op.h_lvl, s);
  const size_t m = std::max<size_t>(1, op.h_col.size());
  op.val.ensure(m * sizeof(double));
  op.val32.ensure(m * sizeof(float));
  op.diag.ensure(std::max(1, op.n) * sizeof(double));
  op.diag32.ensure(std::max(1, op.n) * sizeof(float));
  upload(op.flags, std::vector<int>(std::max(1, op.ndim), 0), s);
  upload(op.status, std::vector<int>(1, 0), s);
  op.epoch = 0;
}


}  // namespace bpf
