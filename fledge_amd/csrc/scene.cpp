// Synthetic workload generators (host C++, libfledge_scene.so): PCG32
// streams, the random triangle soup, the fBm-displaced octahedral sphere,
// the fBm density grid and the batched mesh set.  Inputs only — nothing
// here is on the measured path.
//
// Fixed choices (the workload descriptions leave them open):
//  * PCG32 (O'Neill): state/stream seeded by pcg32_srandom(initstate,
//    initseq); u = (x >> 8) * 2^-24;
//  * value noise: lattice values hash(ix, iy, iz, seed) in [-1, 1), quintic
//    fade, trilinear; fBm = sum over 5 octaves of 0.5^o * noise(p * f0 * 2^o)
//    with base frequency f0;
//  * sphere: octahedron (6 vertices, 8 faces) subdivided k times by edge
//    midpoints projected to the unit sphere (8 * 4^k triangles,
//    4 * 4^k + 2 vertices), radius 1 + amplitude * fBm(direction).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fledge_scene.h"

namespace {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  void seed(uint64_t initstate, uint64_t initseq) {
    state = 0;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  uint32_t h = seed * 0x9E3779B1u;
  h ^= (uint32_t)x * 0x85EBCA77u;
  h = (h << 13) | (h >> 19);
  h ^= (uint32_t)y * 0xC2B2AE3Du;
  h = (h << 13) | (h >> 19);
  h ^= (uint32_t)z * 0x27D4EB2Fu;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return h;
}

inline float lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  return (float)(hash3(x, y, z, seed) >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

inline float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

float value_noise(float x, float y, float z, uint32_t seed) {
  float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
  float c[8];
  for (int k = 0; k < 8; ++k) c[k] = lattice(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2), seed);
  float x00 = c[0] + tx * (c[1] - c[0]), x10 = c[2] + tx * (c[3] - c[2]);
  float x01 = c[4] + tx * (c[5] - c[4]), x11 = c[6] + tx * (c[7] - c[6]);
  float y0 = x00 + ty * (x10 - x00), y1 = x01 + ty * (x11 - x01);
  return y0 + tz * (y1 - y0);
}

float fbm(float x, float y, float z, uint32_t seed, float f0, int octaves) {
  float sum = 0.0f, amp = 1.0f, f = f0;
  for (int o = 0; o < octaves; ++o) {
    sum += amp * value_noise(x * f, y * f, z * f, seed + (uint32_t)o * 1013u);
    amp *= 0.5f;
    f *= 2.0f;
  }
  return sum;
}

template <class F>
void parallel_rows(int64_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int64_t i = t; i < n; i += nt) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

void flg_scene_pcg32(uint64_t initstate, uint64_t initseq, int64_t n, uint32_t* out) {
  Pcg32 r;
  r.seed(initstate, initseq);
  for (int64_t i = 0; i < n; ++i) out[i] = r.next();
}

void flg_scene_soup(uint64_t initstate, uint64_t initseq, int32_t ntris, float* verts, int32_t* ind) {
  Pcg32 r;
  r.seed(initstate, initseq);
  for (int32_t t = 0; t < ntris; ++t) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = r.uniform();
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k)
        verts[(3 * (int64_t)t + v) * 3 + k] = c[k] + (r.uniform() * 0.02f - 0.01f);
    for (int v = 0; v < 3; ++v) ind[3 * (int64_t)t + v] = 3 * t + v;
  }
}

int32_t flg_scene_batch_sizes(uint64_t initstate, uint64_t initseq, int32_t nmeshes, int32_t lo,
                              int32_t hi, int32_t* sizes) {
  Pcg32 r;
  r.seed(initstate, initseq);
  double ratio = (double)hi / (double)lo;
  for (int32_t i = 0; i < nmeshes; ++i) {
    double u = (double)r.uniform();
    sizes[i] = (int32_t)std::llround((double)lo * std::pow(ratio, u));
  }
  return 0;
}

void flg_scene_fbm_grid(int32_t nx, int32_t ny, int32_t nz, uint32_t seed, float base_freq,
                        int32_t octaves, float bias, float scale, float* out) {
  parallel_rows((int64_t)nz * ny, [&](int64_t row) {
    int64_t k = row / ny, j = row % ny;
    float z = ((float)k + 0.5f) / (float)nz, y = ((float)j + 0.5f) / (float)ny;
    float* o = out + row * nx;
    for (int32_t i = 0; i < nx; ++i) {
      float x = ((float)i + 0.5f) / (float)nx;
      float v = bias + scale * fbm(x, y, z, seed, base_freq, octaves);
      o[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
  });
}

int64_t flg_scene_octasphere_counts(int32_t level, int64_t* nverts) {
  int64_t f = 8;
  for (int i = 0; i < level; ++i) f *= 4;
  if (nverts) *nverts = f / 2 + 2;
  return f;
}

int32_t flg_scene_octasphere(int32_t level, float amplitude, uint32_t seed, float base_freq,
                             float* verts, int32_t* ind) {
  std::vector<double> P = {1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
  std::vector<int32_t> F = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
  for (int lv = 0; lv < level; ++lv) {
    std::unordered_map<uint64_t, int32_t> mid;
    mid.reserve(F.size());
    std::vector<int32_t> G;
    G.reserve(F.size() * 4);
    auto midpoint = [&](int32_t a, int32_t b) {
      uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double x = P[3 * a] + P[3 * b], y = P[3 * a + 1] + P[3 * b + 1], z = P[3 * a + 2] + P[3 * b + 2];
      double n = std::sqrt(x * x + y * y + z * z);
      int32_t id = (int32_t)(P.size() / 3);
      P.push_back(x / n);
      P.push_back(y / n);
      P.push_back(z / n);
      mid.emplace(key, id);
      return id;
    };
    for (size_t t = 0; t < F.size(); t += 3) {
      int32_t a = F[t], b = F[t + 1], c = F[t + 2];
      int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int32_t q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
      G.insert(G.end(), q, q + 12);
    }
    F.swap(G);
  }
  const int64_t nv = (int64_t)P.size() / 3;
  parallel_rows(nv, [&](int64_t i) {
    double x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
    float r = 1.0f + amplitude * fbm((float)x, (float)y, (float)z, seed, base_freq, 5);
    verts[3 * i] = (float)x * r;
    verts[3 * i + 1] = (float)y * r;
    verts[3 * i + 2] = (float)z * r;
  });
  std::memcpy(ind, F.data(), F.size() * sizeof(int32_t));
  return (int32_t)nv;
}

// UV sphere of fledge's MakeMeshedSphere (src/plymesh.cpp:110-174): the
// fixture mesh of the reference's bvh_test and shape_tests.  Returns the
// triangle count; verts[3 * n_theta * n_phi], ind[3 * ntris].
int32_t flg_scene_uv_sphere(int32_t n_theta, int32_t n_phi, float radius, float* verts,
                            int32_t* ind) {
  const float PI = 3.14159265358979323846f;
  std::vector<float> v;
  v.reserve((size_t)n_theta * n_phi * 3);
  for (int t = 0; t < n_theta; ++t) {
    float theta = PI * (float)t / (float)(n_theta - 1);
    float cos_theta = std::cos(theta);
    float sin_theta = std::sin(theta);
    for (int p = 0; p < n_phi; ++p) {
      float phi = 2 * PI * (float)p / (float)(n_phi - 1);
      if (t == 0) {
        v.insert(v.end(), {0.0f, 0.0f, radius});
      } else if (t == n_theta - 1) {
        v.insert(v.end(), {0.0f, 0.0f, -radius});
      } else if (p == n_phi - 1) {
        size_t src = v.size() / 3 - (size_t)(n_phi - 1);
        float a = v[3 * src], b = v[3 * src + 1], c = v[3 * src + 2];
        v.insert(v.end(), {a, b, c});
      } else {
        float x = sin_theta * std::cos(phi), y = sin_theta * std::sin(phi), z = cos_theta;
        v.insert(v.end(), {0.0f + x * radius, 0.0f + y * radius, 0.0f + z * radius});
      }
    }
  }
  std::vector<int32_t> idx;
  auto off = [n_phi](int t, int p) { return t * n_phi + p; };
  for (int p = 0; p < n_phi - 1; ++p) idx.insert(idx.end(), {off(0, 0), off(1, p), off(1, p + 1)});
  for (int t = 1; t < n_theta - 2; ++t)
    for (int p = 0; p < n_phi - 1; ++p) {
      idx.insert(idx.end(), {off(t, p), off(t + 1, p), off(t + 1, p + 1)});
      idx.insert(idx.end(), {off(t, p), off(t + 1, p + 1), off(t, p + 1)});
    }
  for (int p = 0; p < n_phi - 1; ++p)
    idx.insert(idx.end(), {off(n_theta - 1, 0), off(n_theta - 2, p), off(n_theta - 2, p + 1)});
  if (verts) std::memcpy(verts, v.data(), v.size() * sizeof(float));
  if (ind) std::memcpy(ind, idx.data(), idx.size() * sizeof(int32_t));
  return (int32_t)(idx.size() / 3);
}

}  // extern "C"
