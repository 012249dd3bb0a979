// Halton tables (device, per context) and per-resolution sampler constants.
#include <vector>

#include "halton.hpp"

namespace flg {

namespace {
struct TableStore {
  int32_t* prime = nullptr;
  float* inv = nullptr;
  int32_t* overflow = nullptr;
  int32_t n = 0;
};
std::map<flg_ctx*, TableStore>& stores() {
  static std::map<flg_ctx*, TableStore> s;
  return s;
}

// Linear sieve below 16384 * 16 (src/common/sampler.h:39-57).
std::vector<int32_t> sieve() {
  const int N = 16384 * 16;
  std::vector<int32_t> prime;
  std::vector<char> flag(N, 0);
  for (int i = 2; i < N; ++i) {
    if (!flag[i]) prime.push_back(i);
    for (size_t j = 0; j < prime.size(); ++j) {
      int v = i * prime[j];
      if (v >= N) break;
      flag[v] = 1;
      if (i % prime[j] == 0) break;
    }
  }
  return prime;
}

// ExGCD / MulModInverse (sampler.h:77-96)
void exgcd(uint64_t a, uint64_t b, int64_t& x, int64_t& y) {
  if (b == 0) {
    x = 1;
    y = 0;
    return;
  }
  exgcd(b, a % b, y, x);
  y -= (int64_t)(a / b) * x;
}
uint64_t mul_mod_inverse(int64_t a, int64_t b) {
  int64_t x, y;
  exgcd((uint64_t)a, (uint64_t)b, x, y);
  int64_t r = x - (x / b) * b;
  return (uint64_t)((r < 0) ? r + b : r);
}
}  // namespace

int get_halton_tables(flg_ctx* ctx, HaltonTables* out) {
  auto& st = stores()[ctx];
  if (!st.prime) {
    std::vector<int32_t> pr = sieve();
    std::vector<float> inv(pr.size());
    for (size_t i = 0; i < pr.size(); ++i) inv[i] = 1.0f / (float)pr[i];
    size_t n = pr.size();
    FLG_CHECK_HIP(ctx, hipMalloc(&st.prime, n * sizeof(int32_t)));
    FLG_CHECK_HIP(ctx, hipMalloc(&st.inv, n * sizeof(float)));
    FLG_CHECK_HIP(ctx, hipMalloc(&st.overflow, sizeof(int32_t)));
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(st.prime, pr.data(), n * sizeof(int32_t),
                                      hipMemcpyHostToDevice, ctx->stream));
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(st.inv, inv.data(), n * sizeof(float),
                                      hipMemcpyHostToDevice, ctx->stream));
    FLG_CHECK_HIP(ctx, hipMemsetAsync(st.overflow, 0, sizeof(int32_t), ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    st.n = (int32_t)n;
  }
  out->prime = st.prime;
  out->inv = st.inv;
  out->nprimes = st.n;
  out->overflow = st.overflow;
  return FLG_OK;
}

bool halton_tables_ready(flg_ctx* ctx, HaltonTables* out) {
  auto it = stores().find(ctx);
  if (it == stores().end() || !it->second.prime) return false;
  out->prime = it->second.prime;
  out->inv = it->second.inv;
  out->nprimes = it->second.n;
  out->overflow = it->second.overflow;
  return true;
}

void release_halton_tables(flg_ctx* ctx) {
  auto it = stores().find(ctx);
  if (it == stores().end()) return;
  (void)hipFree(it->second.prime);
  (void)hipFree(it->second.inv);
  (void)hipFree(it->second.overflow);
  stores().erase(it);
}

HaltonParams make_halton_params(int resx, int resy) {
  HaltonParams h{};
  int interval[2] = {resx, resy};
  int scales[2], exps[2];
  for (int i = 0; i < 2; ++i) {
    int base = i == 0 ? 2 : 3;
    int scale = 1, e = 0;
    int lim = interval[i] < 128 ? interval[i] : 128;
    while (scale < lim) {
      scale *= base;
      ++e;
    }
    scales[i] = scale;
    exps[i] = e;
  }
  h.bs0 = scales[0];
  h.bs1 = scales[1];
  h.exp0 = exps[0];
  h.exp1 = exps[1];
  h.stride = (uint32_t)(scales[0] * scales[1]);
  h.mulinv0 = (uint32_t)mul_mod_inverse(scales[1], scales[0]);
  h.mulinv1 = (uint32_t)mul_mod_inverse(scales[0], scales[1]);
  return h;
}

}  // namespace flg
