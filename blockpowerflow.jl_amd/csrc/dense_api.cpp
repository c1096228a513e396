// C ABI of the dense building blocks of block GMRES, so they can be driven
// on their own: the thin QR of an n x p block (householder!,
// src/block_utils.jl:38-53) and the block-Hessenberg least-squares update
// with back substitution (src/block_gmres.jl:318-375).
#include <cmath>
#include <cstring>

#include "krylov.h"

using namespace bpf;

extern "C" {

int bpf_householder(bpf_ctx ctx, int n, int p, double *d_Q, int ldq, double *d_R, int ldr, int *shifted) {
  return guard([&] {
    BPF_REQUIRE(ctx && d_Q && d_R, "null argument");
    BPF_REQUIRE(n > 0 && p > 0 && p <= 128, "need n > 0 and 0 < p <= 128");
    BPF_REQUIRE(n >= p, "householder needs n >= p");
    BPF_REQUIRE(ldq >= n && ldr >= p, "leading dimension too small");
    hipStream_t s = ctx->stream;
    const int S = pad_width(p);
    std::vector<GroupDesc> hg{{0, p, 0, n, 0, 1 << 30, 1, 0}};
    DevBuf grp, W, V, C;
    upload(grp, hg, s);
    const size_t blk = (size_t)n * p * sizeof(double);
    W.ensure(blk);
    V.ensure(blk);
    C.ensure((size_t)S * S * sizeof(double));
    CholQrWork w;
    w.init(1, S);
    fill_int(s, w.flags.as<int>(), 0, 1);
    launch_cm_to_rm(s, n, p, d_Q, ldq, W.ptr, p, false);
    cholqr2_groups<double>(s, grp.as<GroupDesc>(), 1, S, p, n, 0, n, W.ptr, V.ptr, C.as<double>(), 0, w);
    launch_rm_to_cm(s, n, p, V.ptr, p, false, d_Q, ldq);
    // R (row-major S x S on the device) -> column-major p x p at d_R.
    std::vector<double> h((size_t)S * S), r((size_t)ldr * p, 0.0);
    BPF_CHECK_HIP(hipMemcpyAsync(h.data(), C.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    int fl = 0;
    BPF_CHECK_HIP(hipMemcpyAsync(&fl, w.flags.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j) r[(size_t)j * ldr + i] = j >= i ? h[(size_t)i * S + j] : 0.0;
    BPF_CHECK_HIP(hipMemcpyAsync(d_R, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    if (shifted) *shifted = fl != 0;
  });
}

int bpf_block_hessenberg_lsq(bpf_ctx ctx, int p, int K, const double *h_H, int ldh, const double *h_G,
                             int ldg, double *h_Y, int ldy, double *h_rnorm) {
  return guard([&] {
    BPF_REQUIRE(ctx && h_H && h_G && h_Y, "null argument");
    BPF_REQUIRE(p > 0 && p <= 128 && K > 0, "need 0 < p <= 128 and K > 0");
    BPF_REQUIRE(ldh >= (K + 1) * p && ldg >= p && ldy >= K * p, "leading dimension too small");
    hipStream_t s = ctx->stream;
    const int S = pad_width(p);
    const size_t SS = (size_t)S * S;
    std::vector<GroupDesc> hg{{0, p, 0, 1, 0, 1 << 30, 1, 0}};
    DevBuf grp, E, rnorm, Y;
    upload(grp, hg, s);
    const GroupDesc *dg = grp.as<GroupDesc>();
    // E = [Gamma; 0] (2S x S row-major).
    std::vector<double> e(2 * SS, 0.0);
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j) e[(size_t)i * S + j] = h_G[(size_t)j * ldg + i];
    upload(E, e, s);
    rnorm.ensure(S * sizeof(double));
    std::vector<DevBuf> cols(K);
    PtrTable tab;
    std::vector<double> rn(S);
    for (int k = 0; k < K; ++k) {
      // Block column k: Psi_i = H[i p : (i+1) p, k p : (k+1) p], i = 0..k+1.
      std::vector<double> c(hcol_elems(k, 1, S), 0.0);
      for (int i = 0; i <= k + 1; ++i)
        for (int a = 0; a < p; ++a)
          for (int b = 0; b < p; ++b)
            c[(size_t)i * SS + (size_t)a * S + b] = h_H[(size_t)(k * p + b) * ldh + i * p + a];
      upload(cols[k], c, s);
      tab.set(k, cols[k].ptr, s);
      HessArgs ha{tab.dev<double>(), k, S, 1, dg};
      launch_hess_apply(s, ha, 1);
      HessQrArgs qa{tab.dev<double>(), k, S, 1, E.as<double>(), dg, rnorm.as<double>()};
      launch_hess_qr(s, qa, 1);
      BPF_CHECK_HIP(hipMemcpyAsync(rn.data(), rnorm.ptr, S * sizeof(double), hipMemcpyDeviceToHost, s));
      BPF_CHECK_HIP(hipStreamSynchronize(s));
      if (h_rnorm) std::memcpy(h_rnorm + (size_t)k * p, rn.data(), p * sizeof(double));
    }
    Y.ensure((size_t)K * SS * sizeof(double));
    DevBuf Zb, Rinvb;
    Zb.ensure((size_t)K * SS * sizeof(double));
    Rinvb.ensure((size_t)K * SS * sizeof(double));
    BacksubArgs ba{tab.dev<double>(), K, S, 1, Y.as<double>(), dg, Zb.as<double>(), Rinvb.as<double>()};
    launch_backsub(s, ba, 1);
    std::vector<double> y((size_t)K * SS);
    BPF_CHECK_HIP(hipMemcpyAsync(y.data(), Y.ptr, y.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    BPF_CHECK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < K; ++i)
      for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) h_Y[(size_t)b * ldy + i * p + a] = y[(size_t)i * SS + (size_t)a * S + b];
  });
}

}  // extern "C"
