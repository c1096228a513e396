// GMRES-polynomial preconditioner built on ILU(0) (see poly.cpp).
#pragma once

#include <complex>
#include <vector>

#include "bpf_internal.h"

namespace bpf {

std::vector<std::complex<double>> hessenberg_eigenvalues(const double *H, int d);

struct Poly {
  std::vector<std::complex<double>> roots;   // modified Leja order, conjugates adjacent
  DevBuf V, Wb, hcoef, tab, part, grp;       // root computation
  DevBuf prod, q, t;                         // application work blocks
  int degree() const { return (int)roots.size(); }
  // Harmonic Ritz values of `degree` Arnoldi steps on A M^{-1} (M = N).
  void build(bpf_ctx ctx, bpf_csr A, bpf_prec N, int degree);
  // Z = M^{-1} p(A M^{-1}) V for a row-major n x p block (Z must not alias V).
  template <typename TV>
  void apply(hipStream_t s, bpf_csr A, bpf_prec N, int n, int p, const TV *V, TV *Z);
};

// y = alpha x + beta y elementwise over m entries (beta == 0 ignores y).
template <typename T>
void lincomb(hipStream_t s, size_t m, double alpha, const T *x, double beta, T *y);

}  // namespace bpf
