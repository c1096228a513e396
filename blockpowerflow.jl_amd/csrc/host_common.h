// Host-only parts of libbpf_amd that use no HIP: error plumbing, the level
// analysis of triangular operators, the elimination plan of an IKJ
// factorization on a fixed pattern, and the host thread pool of a batched
// solve.  These files also build with g++ alone, under address, undefined-
// behaviour and thread sanitizers (host_check.cpp).
#pragma once

#include <algorithm>
#include <atomic>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bpf.h"

namespace bpf {

// ---------------------------------------------------------------- errors
// The message of the last error, per host thread.
void set_error(const char *fmt, ...);
const char *last_error();

struct Error {
  int code;
};

#define BPF_THROW(code, ...)          \
  do {                                \
    ::bpf::set_error(__VA_ARGS__);    \
    throw ::bpf::Error{code};         \
  } while (0)

#define BPF_REQUIRE(cond, ...)                 \
  do {                                         \
    if (!(cond)) BPF_THROW(BPF_E_INVALID, __VA_ARGS__); \
  } while (0)

// ------------------------------------------------- triangular level sets
// The host description of a triangular operator in level-scheduled form.
// Rows are stored in level order; row t of the stored CSR is original row
// `h_order[t]`.  Off-diagonal entries only; the diagonal (if not unit) is
// stored separately.  `h_gather` maps stored entry -> position in the source
// value array (combined factor), `h_dgather` maps stored row -> position of
// its diagonal.
struct TriLevels {
  int n = 0;                   // rows of the operator (entries of order)
  int ndim = 0;                // dimension of the vectors it acts on
  bool unit = true;            // implicit unit diagonal
  bool lower = true;           // dependencies on smaller (true) / larger indices
  int nlevels = 0;
  std::vector<int> h_lvl;      // level start offsets into order (nlevels+1)
  std::vector<int> h_order;
  std::vector<int> h_ptr, h_col, h_gather, h_dgather;
  int max_row_nnz = 0;
};

// Restriction of an operator built from a CSR pattern.  A window restricts
// the operator to rows [row_lo, row_hi) and entries with columns in
// [col_lo, col_hi) (-1 = n); dependencies on rows outside it are taken as
// solved.  unit: no diagonal scaling; skip_empty: rows without entries are
// left out (their y is not written).
struct TriWindow {
  int row_lo = 0, row_hi = -1, col_lo = 0, col_hi = -1;
  bool unit = false, skip_empty = false;
  // Relabelling of the vectors the operator acts on (original row i is
  // stored at row relabel[i]): the same levels, rows and arithmetic, with
  // row and dependency indices mapped and each level's rows in storage
  // order.
  const std::vector<int> *relabel = nullptr;
};

// Level analysis (Anderson & Saad 1989 level scheduling) of a CSR pattern.
// kind: 0 = strict lower part (unit diag), 1 = upper part (diag separate),
// 2 = lower part with diag separate, 3 = transpose of the lower part (upper
// operator, diag separate).
void analyze_triop(TriLevels &op, int n, const std::vector<int> &rowptr,
                   const std::vector<int> &colidx, int kind, const TriWindow &win = TriWindow());

// ------------------------------------------------- elimination plan
// Elimination plan of an IKJ factorization on a fixed pattern (ILU(0) on
// A's pattern, LU on the fill pattern), built once per pattern: for every
// strictly lower entry e = (i, k), PivInfo at e holds pivot row k, its
// diagonal position and end and where its update positions start in `map`: for each
// upper entry (k, j) of row k, in order, the position of (i, j) relative to
// row i's start, or -1 where row i has no entry (a fill ILU(0) drops).  So
// no pivot step searches row i's pattern and the loads of a pivot depend on
// nothing computed during the elimination.  `map` ends with one -1 (a pivot
// row without upper entries still reads one entry).
struct PivInfo {
  int diag, end, moff, row;
};
void elim_plan_host(int n, const std::vector<int> &rowptr, const std::vector<int> &colidx,
                    const std::vector<int> &diagpos, std::vector<PivInfo> &pinfo,
                    std::vector<int> &map);

// ------------------------------------------------- host thread pool
// Runs run(i) for i in [0, n) on nthr host threads (nthr <= 1: in order on
// the calling thread).  Each thread first calls start() (e.g. to make the
// device current).  Items are handed out in increasing order; the first
// exception of each thread stops that thread, and after all threads have
// joined the exception of the lowest-numbered failing thread is rethrown
// with that thread's error message.
template <typename Run, typename Start>
void run_pool(int n, int nthr, Run &&run, Start &&start) {
  if (nthr <= 1 || n <= 1) {
    for (int i = 0; i < n; ++i) run(i);
    return;
  }
  nthr = std::min(nthr, n);
  std::atomic<int> next{0};
  std::vector<std::exception_ptr> err(nthr);
  std::vector<std::string> msg(nthr);
  std::vector<std::thread> pool;
  pool.reserve(nthr);
  for (int t = 0; t < nthr; ++t)
    pool.emplace_back([&, t] {
      try {
        start();
        for (int i; (i = next.fetch_add(1)) < n;) run(i);
      } catch (...) {
        err[t] = std::current_exception();
        msg[t] = last_error();   // the message is per thread
      }
    });
  for (auto &th : pool) th.join();
  for (int t = 0; t < nthr; ++t)
    if (err[t]) {
      set_error("%s", msg[t].c_str());
      std::rethrow_exception(err[t]);
    }
}

}  // namespace bpf
