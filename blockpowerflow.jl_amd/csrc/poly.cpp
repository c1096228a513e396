// GMRES-polynomial preconditioning on top of ILU(0) (Loe & Morgan, "Toward
// efficient polynomial preconditioning for GMRES", NLAA 2022).
//
// The roots of the polynomial are the harmonic Ritz values of d Arnoldi steps
// on A M^{-1} (M = ILU(0)) from a fixed random start vector; they are put in
// modified Leja order with conjugate pairs adjacent.  With pi(z) the GMRES
// residual polynomial, p(z) = (1 - pi(z)) / z, the flexible preconditioner
// applied to a block V is Z = M^{-1} p(A M^{-1}) V, formed by the product
// recurrence
//     q_j = M^{-1} prod_j,  Z += q_j / theta_j,  prod_{j+1} = prod_j - A q_j / theta_j
// and, for a conjugate pair (a +- ib, m2 = a^2 + b^2),
//     q = M^{-1} prod, r = M^{-1} A q, w = 2a q - r, Z += w / m2, prod -= A w / m2,
// so all arithmetic stays real.
#include "poly.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <random>

#include "dense.h"
#include "poly.h"

namespace bpf {

using cplx = std::complex<double>;

// Eigenvalues of a d x d upper Hessenberg matrix (row-major, ld d) by the
// complex shifted QR iteration with Wilkinson shifts and deflation.
std::vector<cplx> hessenberg_eigenvalues(const double *Hin, int d) {
  std::vector<cplx> H(d * d);
  for (int i = 0; i < d * d; ++i) H[i] = Hin[i];
  auto at = [&](int r, int c) -> cplx & { return H[r * d + c]; };
  std::vector<cplx> ev(d);
  int hi = d - 1;
  int iter = 0;
  const double eps = 2.220446049250313e-16;
  while (hi >= 0) {
    if (hi == 0) { ev[0] = at(0, 0); break; }
    // find small subdiagonal
    int lo = hi;
    while (lo > 0) {
      const double s = std::abs(at(lo - 1, lo - 1)) + std::abs(at(lo, lo));
      if (std::abs(at(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) { at(lo, lo - 1) = 0.0; break; }
      --lo;
    }
    if (lo == hi) { ev[hi] = at(hi, hi); --hi; iter = 0; continue; }
    if (++iter > 300 * d) {   // give up: take the diagonal
      for (int i = lo; i <= hi; ++i) ev[i] = at(i, i);
      break;
    }
    // Wilkinson shift from the trailing 2x2 of the active block.
    const cplx a = at(hi - 1, hi - 1), b = at(hi - 1, hi), c = at(hi, hi - 1), dd = at(hi, hi);
    const cplx tr = a + dd, det = a * dd - b * c;
    const cplx disc = std::sqrt(tr * tr / 4.0 - det);
    cplx mu1 = tr / 2.0 + disc, mu2 = tr / 2.0 - disc;
    cplx mu = std::abs(mu1 - dd) < std::abs(mu2 - dd) ? mu1 : mu2;
    if (iter % 11 == 10) mu += cplx(std::abs(at(hi, hi - 1)), 0.0);   // exceptional shift
    // QR step on the active block lo..hi: H - mu I = QR, H <- RQ + mu I (Givens).
    std::vector<cplx> cs(hi - lo), sn(hi - lo);
    for (int i = lo; i <= hi; ++i) at(i, i) -= mu;
    for (int k = lo; k < hi; ++k) {
      const cplx x = at(k, k), y = at(k + 1, k);
      const double nr = std::sqrt(std::norm(x) + std::norm(y));
      cplx cc = 1.0, ss = 0.0;
      if (nr > 0) { cc = x / nr; ss = y / nr; }
      cs[k - lo] = cc;
      sn[k - lo] = ss;
      // apply G^H to rows k, k+1 (columns k..d-1)
      for (int j = k; j < d; ++j) {
        const cplx u = at(k, j), v = at(k + 1, j);
        at(k, j) = std::conj(cc) * u + std::conj(ss) * v;
        at(k + 1, j) = -ss * u + cc * v;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const cplx cc = cs[k - lo], ss = sn[k - lo];
      // apply G to columns k, k+1 (rows 0..min(k+2, hi))
      const int rmax = std::min(k + 2, hi);
      for (int i = 0; i <= rmax; ++i) {
        const cplx u = at(i, k), v = at(i, k + 1);
        at(i, k) = u * cc + v * ss;
        at(i, k + 1) = -u * std::conj(ss) + v * std::conj(cc);
      }
    }
    for (int i = lo; i <= hi; ++i) at(i, i) += mu;
  }
  return ev;
}

// Modified Leja ordering; conjugate pairs (a+ib, a-ib) stay adjacent with the
// positive imaginary part first.  Values with |im| <= tol*|z| become real.
static std::vector<cplx> leja_order(std::vector<cplx> z) {
  const int d = (int)z.size();
  for (auto &v : z)
    if (std::abs(v.imag()) <= 1e-10 * std::abs(v)) v = cplx(v.real(), 0.0);
  // Pair conjugates: keep only the members with im >= 0, remember pairs.
  std::vector<cplx> pos;
  std::vector<bool> used(d, false);
  for (int i = 0; i < d; ++i) {
    if (used[i]) continue;
    used[i] = true;
    if (z[i].imag() == 0.0) { pos.push_back(z[i]); continue; }
    // find the closest conjugate partner
    int best = -1;
    double bd = 1e300;
    for (int j = 0; j < d; ++j)
      if (!used[j] && z[j].imag() != 0.0) {
        const double dist = std::abs(z[j] - std::conj(z[i]));
        if (dist < bd) { bd = dist; best = j; }
      }
    if (best >= 0) {
      used[best] = true;
      const cplx m = 0.5 * (z[i] + std::conj(z[best]));
      pos.push_back(cplx(m.real(), std::abs(m.imag())));
    } else {
      pos.push_back(cplx(z[i].real(), 0.0));   // unpaired: project to the real axis
    }
  }
  std::vector<cplx> out;
  std::vector<bool> taken(pos.size(), false);
  std::vector<cplx> chosen;
  for (size_t step = 0; step < pos.size(); ++step) {
    int bi = -1;
    double bv = -1e300;
    for (size_t i = 0; i < pos.size(); ++i) {
      if (taken[i]) continue;
      double v;
      if (chosen.empty()) {
        v = std::abs(pos[i]);
      } else {
        v = 0.0;
        for (const auto &c : chosen) v += std::log(std::abs(pos[i] - c) + 1e-300);
      }
      if (v > bv) { bv = v; bi = (int)i; }
    }
    taken[bi] = true;
    chosen.push_back(pos[bi]);
    if (pos[bi].imag() != 0.0) chosen.push_back(std::conj(pos[bi]));
    out.push_back(pos[bi]);
    if (pos[bi].imag() != 0.0) out.push_back(std::conj(pos[bi]));
  }
  return out;
}

void Poly::build(bpf_ctx ctx, bpf_csr A, bpf_prec N, int degree) {
  roots.clear();
  for (DevBuf *b : {&V, &Wb, &hcoef, &tab, &part, &grp, &prod, &q, &t}) b->pool = &ctx->pool;
  if (degree <= 0 || !N || A->n < 2 * degree) return;
  const int n = A->n;
  const int d = degree;
  hipStream_t s = ctx->stream;
  // d Arnoldi steps (classical Gram-Schmidt, two passes) on A M^{-1}, p = 1.
  V.ensure((size_t)(d + 1) * n * sizeof(double));
  Wb.ensure((size_t)n * sizeof(double));
  double *Vp = V.as<double>(), *w = Wb.as<double>();
  std::vector<double> x(n);
  std::mt19937_64 rng(12345);
  std::normal_distribution<double> nd(0.0, 1.0);
  double nrm = 0.0;
  for (auto &v : x) { v = nd(rng); nrm += v * v; }
  nrm = std::sqrt(nrm);
  for (auto &v : x) v /= nrm;
  BPF_CHECK_HIP(hipMemcpyAsync(Vp, x.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<double> H((size_t)(d + 1) * d, 0.0);
  std::vector<GroupDesc> hg{{0, 1, 0, n, 0, 1 << 30, 1, 0}};
  upload(grp, hg, s);
  hcoef.ensure((size_t)(d + 1) * 16 * 16 * sizeof(double));
  DevBuf dpart;
  dpart.ensure(kDotBlocks * sizeof(double));
  tab.ensure((size_t)(d + 1) * sizeof(void *));
  std::vector<const void *> ptrs(d + 1);
  for (int i = 0; i <= d; ++i) ptrs[i] = Vp + (size_t)i * n;
  BPF_CHECK_HIP(hipMemcpyAsync(tab.ptr, ptrs.data(), (d + 1) * sizeof(void *), hipMemcpyHostToDevice, s));
  BPF_CHECK_HIP(hipStreamSynchronize(s));
  ctx->ensure_small(4 * (d + 2));
  double *hs = ctx->h_small;
  int steps = d;
  for (int j = 0; j < d; ++j) {
    double *vj = Vp + (size_t)j * n;
    prec_apply_rm<double>(s, N, 1, vj, 1, w, 1);
    double *t = Vp + (size_t)(j + 1) * n;   // A M^{-1} v_j lands in v_{j+1}
    launch_spmm<double>(s, n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), A->val.ptr, false, 1, w,
                        1, t, 1, 1.0, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      ProjArgs pa{tab.as<const void *>(), nullptr, 0, j + 1, 1, t, 1, n, grp.as<GroupDesc>(), 1, 0, nullptr};
      launch_proj<double, double>(s, pa, 16, hcoef.as<double>(), 256, 0, false, part);
      GemmArgs ga{tab.as<const void *>(), nullptr, 0, j + 1, 1, hcoef.as<double>(), 256, 0, 16, t, 1, n,
                  -1.0, 1.0, grp.as<GroupDesc>()};
      launch_block_gemm<double, double>(s, ga, 16, 1);
      for (int i = 0; i <= j; ++i)
        BPF_CHECK_HIP(hipMemcpyAsync(hs + i, hcoef.as<double>() + (size_t)i * 256, sizeof(double),
                                     hipMemcpyDeviceToHost, s));
      BPF_CHECK_HIP(hipStreamSynchronize(s));
      for (int i = 0; i <= j; ++i) H[(size_t)i * d + j] += hs[i];
    }
    launch_dot<double>(s, n, t, t, hcoef.as<double>(), dpart.as<double>());
    BPF_CHECK_HIP(hipMemcpyAsync(hs, hcoef.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    const double h = std::sqrt(hs[0]);
    H[(size_t)(j + 1) * d + j] = h;
    if (!(h > 1e-14 * std::abs(H[(size_t)j * d + j])) || !std::isfinite(h)) {   // invariant subspace
      steps = j + 1;
      break;
    }
    std::vector<double> sc{1.0 / h};
    // t /= h
    axpby(s, n, 1.0 / h - 1.0, t, 1.0, t);
  }
  // Harmonic Ritz values: eig(H_m + h_{m+1,m}^2 f e_m^T), H_m^T f = e_m.
  const int m = steps;
  std::vector<double> Hm((size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Hm[(size_t)i * m + j] = H[(size_t)i * d + j];
  const double hlast = m < d + 1 ? H[(size_t)m * d + (m - 1)] : 0.0;
  if (hlast != 0.0) {
    // Solve H_m^T f = e_m by Gaussian elimination with partial pivoting.
    std::vector<double> M((size_t)m * m), f(m, 0.0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) M[(size_t)i * m + j] = Hm[(size_t)j * m + i];
    f[m - 1] = 1.0;
    std::vector<int> piv(m);
    bool ok = true;
    for (int k = 0; k < m; ++k) {
      int p = k;
      for (int i = k + 1; i < m; ++i)
        if (std::abs(M[(size_t)i * m + k]) > std::abs(M[(size_t)p * m + k])) p = i;
      if (M[(size_t)p * m + k] == 0.0) { ok = false; break; }
      if (p != k) {
        for (int j = 0; j < m; ++j) std::swap(M[(size_t)k * m + j], M[(size_t)p * m + j]);
        std::swap(f[k], f[p]);
      }
      for (int i = k + 1; i < m; ++i) {
        const double l = M[(size_t)i * m + k] / M[(size_t)k * m + k];
        for (int j = k; j < m; ++j) M[(size_t)i * m + j] -= l * M[(size_t)k * m + j];
        f[i] -= l * f[k];
      }
    }
    if (ok) {
      for (int i = m - 1; i >= 0; --i) {
        double v = f[i];
        for (int j = i + 1; j < m; ++j) v -= M[(size_t)i * m + j] * f[j];
        f[i] = v / M[(size_t)i * m + i];
      }
      for (int i = 0; i < m; ++i) Hm[(size_t)i * m + (m - 1)] += hlast * hlast * f[i];
    }
  }
  auto ev = hessenberg_eigenvalues(Hm.data(), m);
  // Stability control (Loe & Morgan, Sec. 3.2): a root whose product of other
  // factors pof(k) = prod_{j != k} |1 - theta_k / theta_j| exceeds 1e4 gets
  // ceil((log10 pof(k) - 4) / 14) extra copies.
  std::vector<cplx> aug = ev;
  for (int k = 0; k < m; ++k) {
    double lp = 0.0;
    for (int j = 0; j < m; ++j)
      if (j != k && std::abs(ev[j]) > 0) lp += std::log10(std::abs(1.0 - ev[k] / ev[j]) + 1e-300);
    if (lp > 4.0) {
      const int extra = (int)std::ceil((lp - 4.0) / 14.0);
      for (int e = 0; e < extra; ++e) aug.push_back(ev[k]);
    }
  }
  roots = leja_order(aug);
}

template <typename TV>
void Poly::apply(hipStream_t s, bpf_csr A, bpf_prec N, int n, int p, const TV *Vin, TV *Z) {
  const bool f32 = std::is_same<TV, float>::value;
  const size_t blk = (size_t)n * p * sizeof(TV);
  prod.ensure(blk);
  q.ensure(blk);
  t.ensure(blk);
  TV *PR = prod.as<TV>(), *Q = q.as<TV>(), *T = t.as<TV>();
  BPF_CHECK_HIP(hipMemcpyAsync(PR, Vin, blk, hipMemcpyDeviceToDevice, s));
  const void *aval = f32 ? A->val32.ptr : A->val.ptr;
  auto spmm = [&](const TV *x, TV *y, double alpha, double beta) {
    launch_spmm<TV>(s, n, A->nnz, A->rowptr.as<int>(), A->colidx.as<int>(), aval, f32, p, x, p, y, p,
                    alpha, beta);
  };
  auto minv = [&](const TV *x, TV *y) {
    prec_apply_rm<TV>(s, N, p, x, p, y, p);
  };
  const int d = (int)roots.size();
  bool first = true;
  int i = 0;
  while (i < d) {
    const cplx th = roots[i];
    minv(PR, Q);                                   // q = M^{-1} prod
    if (th.imag() == 0.0) {
      const double inv = 1.0 / th.real();
      lincomb<TV>(s, (size_t)n * p, inv, Q, first ? 0.0 : 1.0, Z);   // Z += q / theta
      first = false;
      if (i < d - 1) spmm(Q, PR, -inv, 1.0);       // prod -= A q / theta
      i += 1;
    } else {
      const double a = th.real(), m2 = std::norm(th);
      spmm(Q, T, 1.0, 0.0);                        // t = A q
      minv(T, T);                                  // r = M^{-1} t
      lincomb<TV>(s, (size_t)n * p, -1.0, T, 2.0 * a, Q);   // w = 2a q - r
      lincomb<TV>(s, (size_t)n * p, 1.0 / m2, Q, first ? 0.0 : 1.0, Z);
      first = false;
      if (i < d - 2) spmm(Q, PR, -1.0 / m2, 1.0);  // prod -= A w / m2
      i += 2;
    }
  }
}

template void Poly::apply<float>(hipStream_t, bpf_csr, bpf_prec, int, int, const float *, float *);
template void Poly::apply<double>(hipStream_t, bpf_csr, bpf_prec, int, int, const double *, double *);

}  // namespace bpf

extern "C" int bpf_hessenberg_eigenvalues(const double *h_H, int d, double *h_re, double *h_im) {
  return bpf::guard([&] {
    BPF_REQUIRE(h_H && h_re && h_im && d > 0, "bad argument");
    auto ev = bpf::hessenberg_eigenvalues(h_H, d);
    for (int i = 0; i < d; ++i) {
      h_re[i] = ev[i].real();
      h_im[i] = ev[i].imag();
    }
  });
}
