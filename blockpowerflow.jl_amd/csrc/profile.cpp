// Kernel-family timers behind bpf_profile_enable / bpf_profile_read.
#include "profile.h"

#include <mutex>
#include <vector>

#include "bpf_internal.h"

namespace bpf {
namespace prof {

namespace {
struct Rec {
  hipEvent_t a, b;
  int fam;
  double bytes, flops;
};
struct State {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<Rec> open;   // started, not yet ended (indexed by token)
  std::vector<Rec> done;
  long long launches[NFAM] = {};
  double ms[NFAM] = {}, bytes[NFAM] = {}, flops[NFAM] = {};
};
State g;
std::mutex mu;

hipEvent_t take() {
  if (!g.pool.empty()) {
    hipEvent_t e = g.pool.back();
    g.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  BPF_CHECK_HIP(hipEventCreate(&e));
  return e;
}

void drain() {
  for (auto &r : g.done) {
    BPF_CHECK_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    BPF_CHECK_HIP(hipEventElapsedTime(&t, r.a, r.b));
    g.launches[r.fam] += 1;
    g.ms[r.fam] += t;
    g.bytes[r.fam] += r.bytes;
    g.flops[r.fam] += r.flops;
    g.pool.push_back(r.a);
    g.pool.push_back(r.b);
  }
  g.done.clear();
}
}  // namespace

bool enabled() { return g.on; }

int begin(hipStream_t s) {
  if (!g.on) return -1;
  std::lock_guard<std::mutex> lk(mu);
  Rec r{take(), take(), 0, 0.0, 0.0};
  BPF_CHECK_HIP(hipEventRecord(r.a, s));
  g.open.push_back(r);
  return (int)g.open.size() - 1;
}

ExtPair ext_begin() {
  ExtPair p;
  if (!g.on) return p;
  std::lock_guard<std::mutex> lk(mu);
  p.a = take();
  p.b = take();
  return p;
}

void ext_end(const ExtPair &p, Family f, double bytes, double flops) {
  if (!p.a || !g.on) return;
  std::lock_guard<std::mutex> lk(mu);
  g.done.push_back(Rec{p.a, p.b, f, bytes, flops});
  if (g.done.size() > 4096) drain();
}

void end(int token, hipStream_t s, Family f, double bytes, double flops) {
  if (token < 0 || !g.on) return;
  std::lock_guard<std::mutex> lk(mu);
  Rec r = g.open[token];
  BPF_CHECK_HIP(hipEventRecord(r.b, s));
  r.fam = f;
  r.bytes = bytes;
  r.flops = flops;
  g.done.push_back(r);
  if (token == (int)g.open.size() - 1) g.open.pop_back();
  if (g.done.size() > 4096) drain();
}

}  // namespace prof
}  // namespace bpf

using namespace bpf;

extern "C" {

int bpf_profile_enable(bpf_ctx ctx, int on) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    std::lock_guard<std::mutex> lk(prof::mu);
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    prof::drain();
    for (int f = 0; f < prof::NFAM; ++f) {
      prof::g.launches[f] = 0;
      prof::g.ms[f] = prof::g.bytes[f] = prof::g.flops[f] = 0.0;
    }
    prof::g.open.clear();
    prof::g.on = on != 0;
  });
}

int bpf_profile_read(bpf_ctx ctx, int family, long long *launches, double *ms, double *bytes,
                     double *flops) {
  return guard([&] {
    BPF_REQUIRE(ctx, "null context");
    BPF_REQUIRE(family >= 0 && family < prof::NFAM, "bad kernel family");
    std::lock_guard<std::mutex> lk(prof::mu);
    BPF_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    prof::drain();
    if (launches) *launches = prof::g.launches[family];
    if (ms) *ms = prof::g.ms[family];
    if (bytes) *bytes = prof::g.bytes[family];
    if (flops) *flops = prof::g.flops[family];
  });
}

}  // extern "C"
