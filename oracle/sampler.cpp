// TEST INFRASTRUCTURE — CPU restatement of fledge's Halton sampler
// (src/common/sampler.h) and pinhole camera (src/common/camera.h).
#include <cmath>

#include "oracle_impl.hpp"

namespace orc {

// GetPrimeList sampler.h:39-57: linear sieve below 16384*16.
const std::vector<int>& prime_list() {
  static std::vector<int> primes = [] {
    const int N = 16384 * 16;
    std::vector<int> prime;
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
  }();
  return primes;
}

// RadicalInverse_ sampler.h:59-72
float radical_inverse_(int base, uint64_t a) {
  float inv_base = 1.0f / (float)base;
  uint64_t reversed = 0;
  float inv_base_n = 1.0f;
  while (a) {
    uint64_t next = a / (uint64_t)base;
    uint64_t digit = a - next * (uint64_t)base;
    reversed = reversed * (uint64_t)base + digit;
    inv_base_n *= inv_base;
    a = next;
  }
  return (float)reversed * inv_base_n;
}

// ExGCD / MulModInverse sampler.h:77-96 ; Mod<int64> math_utils.h:155-162
static void exgcd(uint64_t a, uint64_t b, int64_t& x, int64_t& y) {
  if (b == 0) { x = 1; y = 0; return; }
  exgcd(b, a % b, y, x);
  y -= (int64_t)(a / b) * x;
}
uint64_t mul_mod_inverse(int64_t a, int64_t b) {
  int64_t x, y;
  exgcd((uint64_t)a, (uint64_t)b, x, y);
  int64_t r = x - (x / b) * b;
  return (uint64_t)((r < 0) ? r + b : r);
}

// InverseRadicalInverse sampler.h:111-120
static uint64_t inverse_radical_inverse(int base, uint64_t inverse, int n_digits) {
  uint64_t index = 0;
  for (int i = 0; i < n_digits; ++i) {
    uint64_t digit = inverse % (uint64_t)base;
    inverse /= (uint64_t)base;
    index = index * (uint64_t)base + digit;
  }
  return index;
}

// HaltonSampler ctor sampler.h:145-163 (M_MAX_RESOLUTION = 128)
Halton::Halton(uint64_t spp_, int resx, int resy) : spp(spp_) {
  int interval[2] = {resx, resy};
  for (int i = 0; i < 2; ++i) {
    int base = i == 0 ? 2 : 3;
    int scale = 1, e = 0;
    int lim = interval[i] < 128 ? interval[i] : 128;
    while (scale < lim) { scale *= base; ++e; }
    base_scales[i] = scale;
    base_exp[i] = e;
  }
  stride = (uint64_t)base_scales[0] * (uint64_t)base_scales[1];
  mul_inverse[0] = (int64_t)mul_mod_inverse(base_scales[1], base_scales[0]);
  mul_inverse[1] = (int64_t)mul_mod_inverse(base_scales[0], base_scales[1]);
}

// getIndexForSample sampler.h:206-227 (the pixel offset; 0 for pixel (0,0),
// which the reference leaves uninitialised)
uint64_t Halton::offset_for_pixel(int x, int y) const {
  uint64_t off = 0;
  if (stride > 1) {
    int pm[2];
    int p[2] = {x, y};
    for (int i = 0; i < 2; ++i) {
      int r = p[i] - (p[i] / 128) * 128;
      pm[i] = r < 0 ? r + 128 : r;
    }
    for (int i = 0; i < 2; ++i) {
      uint64_t dim_offset = inverse_radical_inverse(i == 0 ? 2 : 3, (uint64_t)pm[i], base_exp[i]);
      off += dim_offset * (stride / (uint64_t)base_scales[i]) * (uint64_t)mul_inverse[i];
    }
    off %= stride;
  }
  return off;
}

void Halton::start(int x, int y, uint64_t sample) {
  px = x;
  py = y;
  offset = offset_for_pixel(x, y);
  index = offset + sample * stride;
  cnt = 0;
}

// setPixel: dimension 0 of the pixel's sample 0
void Halton::set_pixel(int x, int y) {
  start(x, y, 0);
  psi = 0;
}

// reset: dimension 0 of the next sample of the pixel
bool Halton::reset() {
  cnt = 0;
  index = offset + (psi + 1) * stride;
  return ++psi < spp;
}

// sampleDimension sampler.h:241-251
float Halton::sample_dimension(uint64_t idx, uint64_t dim) {
  const auto& pr = prime_list();
  if (dim >= pr.size()) {  // the reference would read a zero base and assert
    overflow = true;
    return 0.5f;
  }
  if (dim == 0) return radical_inverse_(pr[0], idx / (uint64_t)base_scales[0]);
  if (dim == 1) return radical_inverse_(pr[1], idx / (uint64_t)base_scales[1]);
  return radical_inverse_(pr[dim], idx);
}

// get1D sampler.h:171-174
float Halton::get1D() {
  float v = sample_dimension(index, cnt++);
  return smin(v, ONE_MINUS_EPSILON);
}

// Camera ctor camera.h:12-21: forward = Normalize(lookAt - pos),
// right = Normalize(up x forward); fov_ratio = tanf(PI * fov / 360) (C library,
// host-only constant).
Camera::Camera(V3 pos_, V3 look_at, V3 up_, float fov) : pos(pos_), up(up_) {
  forward = normalize(look_at - pos);
  right = normalize(cross(up, forward));
  up2 = normalize(cross(forward, right));
  fov_ratio = std::tan(PI * fov / 360.0f);
}

// Camera::generateRay camera.h:21-31: the "- 1.0" literal promotes the offset
// products to double before they are rounded back to float.
void Camera::generate(float x, float y, int w, int h, V3* o, V3* d) const {
  float aspect = (float)w / (float)h;
  float y_offset = (float)((double)fov_ratio * ((double)((y / (float)h) * 2.0f) - 1.0));
  float x_offset = (float)((double)(fov_ratio * aspect) * ((double)((x / (float)w) * 2.0f) - 1.0));
  V3 dir = (forward + right * x_offset) + up2 * y_offset;
  *o = pos;
  *d = normalize(dir);
}

}  // namespace orc
