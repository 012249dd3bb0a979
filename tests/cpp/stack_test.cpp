// Host check of the traversal's stack policies (test infrastructure): the
// device kernels' closest-hit traversal (traverse.hpp) run on the CPU with the
// private stack and with the LDS ring stack at ring sizes 1, 2 and 8 (host
// memory standing in for LDS; small rings force the spill path), on a tree
// packed from the oracle's build, against the oracle's traversal.
// Also checks the device Halton digit loop (halton.hpp radical_inverse, run
// on the CPU) against integer division.
// Built with hipcc as host code; runs without a GPU.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../fledge_amd/csrc/traverse.hpp"
#include "../../fledge_amd/csrc/volume.hpp"
#include "fledge_scene.h"
#include "oracle_impl.hpp"

extern "C" void orc_bvh_test_rays(int64_t n, float* rays);

namespace {

int failures = 0;

void pack(const orc::Lbvh& b, const float* verts, const int32_t* ind, std::vector<flg::BNode>& nodes,
          std::vector<flg::BTri>& tris) {
  const int n = b.n;
  tris.resize(n);
  auto vert = [&](int v) { return flg::v3(verts[3 * v], verts[3 * v + 1], verts[3 * v + 2]); };
  std::vector<flg::V3> tlo(n), thi(n);
  for (int p = 0; p < n; ++p) {
    const int t = (int)b.order[p];
    flg::V3 a = vert(ind[3 * t]), c = vert(ind[3 * t + 1]), d = vert(ind[3 * t + 2]);
    tris[p].v0 = make_float4(a.x, a.y, a.z, flg::u2f((uint32_t)t));
    tris[p].v1 = make_float4(c.x, c.y, c.z, 0.0f);
    tris[p].v2 = make_float4(d.x, d.y, d.z, 0.0f);
    tlo[p] = flg::vmin(flg::vmin(a, c), d);
    thi[p] = flg::vmax(flg::vmax(a, c), d);
  }
  nodes.resize(n > 1 ? n - 1 : 0);
  for (int i = 0; i + 1 < n; ++i) {
    const int s = b.split[i];
    float box[12];
    int32_t ref[2];
    for (int side = 0; side < 2; ++side) {
      const int c = s + side;
      const bool leaf = side ? b.rleaf[i] : b.lleaf[i];
      flg::V3 lo = leaf ? tlo[c] : flg::v3(b.nlo[c].x, b.nlo[c].y, b.nlo[c].z);
      flg::V3 hi = leaf ? thi[c] : flg::v3(b.nhi[c].x, b.nhi[c].y, b.nhi[c].z);
      float* q = box + 6 * side;
      q[0] = lo.x; q[1] = lo.y; q[2] = lo.z; q[3] = hi.x; q[4] = hi.y; q[5] = hi.z;
      ref[side] = leaf ? (int32_t)(0x80000000u | (uint32_t)c) : c;
    }
    nodes[i].a = make_float4(box[0], box[1], box[2], box[3]);
    nodes[i].b = make_float4(box[4], box[5], box[6], box[7]);
    nodes[i].c = make_float4(box[8], box[9], box[10], box[11]);
    nodes[i].d = make_int4(ref[0], ref[1], 0, 0);
  }
}

template <int K>
int32_t trace_ring(const flg::TraceGeom& g, flg::V3 o, flg::V3 d, float& best, flg::Counters& c) {
  uint2 fast[K];
  uint2 spill[flg::kStack - K];
  flg::RingStack<K> st(fast, 1, spill);
  return flg::trace_one_s<false, true>(g, o, d, best, c, st);
}

void run_case(const char* name, const std::vector<float>& verts, const std::vector<int32_t>& ind,
              const std::vector<float>& rays) {
  orc::Lbvh b;
  const int n = (int)ind.size() / 3;
  orc::lbvh_build(b, verts.data(), ind.data(), n, (int)verts.size() / 3);
  std::vector<flg::BNode> nodes;
  std::vector<flg::BTri> tris;
  pack(b, verts.data(), ind.data(), nodes, tris);
  flg::TraceGeom g;
  g.nodes = nodes.data();
  g.tris = tris.data();
  g.n = n;
  const float root[6] = {b.glo.x, b.glo.y, b.glo.z, b.ghi.x, b.ghi.y, b.ghi.z};
  std::memcpy(g.root, root, sizeof root);
  const size_t nr = rays.size() / 8;
  size_t hits = 0;
  uint64_t visits = 0;
  for (size_t i = 0; i < nr; ++i) {
    const float* r = &rays[8 * i];
    const flg::V3 o = flg::v3(r[0], r[1], r[2]), d = flg::v3(r[4], r[5], r[6]);
    float t_ref = 0;
    orc::V3 ng;
    const int32_t id_ref = orc::lbvh_trace(b, orc::v3(r[0], r[1], r[2]), orc::v3(r[4], r[5], r[6]), r[3],
                                           false, &t_ref, &ng, nullptr);
    float t0 = r[3], t1 = r[3], t2 = r[3], t8 = r[3];
    flg::Counters c0{0, 0}, c1{0, 0}, c2{0, 0}, c8{0, 0};
    const int32_t p0 = flg::trace_one<false, true>(g, o, d, t0, c0);
    const int32_t p1 = trace_ring<1>(g, o, d, t1, c1);
    const int32_t p2 = trace_ring<2>(g, o, d, t2, c2);
    const int32_t p8 = trace_ring<8>(g, o, d, t8, c8);
    const int32_t id0 = p0 >= 0 ? (int32_t)flg::f2u(tris[p0].v0.w) : -1;
    bool ok = id0 == id_ref && p0 == p1 && p0 == p2 && p0 == p8;
    ok = ok && c0.nodes == c1.nodes && c0.nodes == c2.nodes && c0.nodes == c8.nodes && c0.tris == c8.tris;
    if (p0 >= 0) {
      uint32_t a, b1, b2, b8;
      std::memcpy(&a, &t0, 4);
      std::memcpy(&b1, &t1, 4);
      std::memcpy(&b2, &t2, 4);
      std::memcpy(&b8, &t8, 4);
      uint32_t r0;
      std::memcpy(&r0, &t_ref, 4);
      ok = ok && a == r0 && a == b1 && a == b2 && a == b8;
      ++hits;
    }
    visits += c0.nodes;
    if (!ok) {
      std::printf("FAIL %s ray %zu: ids %d/%d/%d/%d oracle %d, visits %u/%u/%u/%u\n", name, i, p0, p1, p2, p8,
                  id_ref, c0.nodes, c1.nodes, c2.nodes, c8.nodes);
      ++failures;
      return;
    }
  }
  std::printf("PASS %s: %zu rays, %zu hits, %.2f node visits per ray\n", name, nr, hits,
              (double)visits / (double)nr);
}

}  // namespace

// tile_order (the path start order of the mesh kernels, 16 x 16 tiles, and
// of the path tracer's first bounce, 8 x 64) is a bijection on
// [0, w h tasks): exhaustively on small ragged shapes; at widths of 2^27 and
// above (h = 1: the identity; h = 8: the 16 x 16 layout checked by hand; the
// 8 x 64 layout at 1920 x 1080 checked by hand).
template <uint32_t TW, uint32_t TH>
void check_tile_bijection() {
  const int shapes[][3] = {{1, 1, 1},   {17, 33, 3}, {16, 16, 2}, {40, 7, 5},  {3, 100, 1},
                           {1920, 24, 2}, {37, 131, 3}, {1, 70, 2}, {130, 1, 1}, {1920, 1080, 1}};
  for (const auto& s : shapes) {
    const uint32_t n = (uint32_t)s[0] * s[1] * s[2];
    std::vector<uint8_t> seen(n, 0);
    bool ok = true;
    for (uint32_t q = 0; q < n; ++q) {
      const uint32_t p = flg::tile_order<TW, TH>(q, s[2], s[0], s[1]);
      if (p >= n || seen[p]) ok = false;
      else seen[p] = 1;
    }
    if (!ok) {
      std::printf("FAIL tile_order<%u, %u> not a bijection at %d x %d x %d\n", TW, TH, s[0], s[1], s[2]);
      ++failures;
    }
  }
  std::mt19937 gen(5);
  for (int32_t w : {(1 << 28) + 3, (1 << 30) + 1, 2147483647}) {
    std::uniform_int_distribution<uint32_t> U(0, (uint32_t)w - 1);
    for (int k = 0; k < 100000; ++k) {
      const uint32_t q = U(gen);
      if (flg::tile_order<TW, TH>(q, 1, w, 1) != q) {
        std::printf("FAIL tile_order<%u, %u> at width %d: %u -> %u\n", TW, TH, w, q,
                    flg::tile_order<TW, TH>(q, 1, w, 1));
        ++failures;
        break;
      }
    }
  }
}

void check_tile_order() {
  check_tile_bijection<16, 16>();
  check_tile_bijection<8, 64>();
  {  // 1920 x 1080 in 8 x 64 tiles: a band holds 1920 * 64 pixels, a tile 512;
     // the last band is 1080 - 16 * 64 = 56 rows high, its tiles 448 pixels
    const int32_t w = 1920, h = 1080;
    const uint32_t cases[][3] = {{0, 0, 0},           {9, 1, 1},           {511, 7, 63},
                                 {512 + 8, 8, 1},      {1920 * 64, 0, 64},  {1920 * 64 * 16 + 447, 7, 1079},
                                 {1920 * 64 * 16 + 448 * 239 + 9, 239 * 8 + 1, 1025}};
    for (const auto& c : cases) {
      const uint32_t want = c[2] * (uint32_t)w + c[1];
      if (flg::tile_order<8, 64>(c[0], 1, w, h) != want) {
        std::printf("FAIL tile_order<8, 64>(%u) at %d x %d: %u, expected %u\n", c[0], w, h,
                    flg::tile_order<8, 64>(c[0], 1, w, h), want);
        ++failures;
      }
    }
  }
  {  // w = 2^27 + 5, h = 8, 1 task: one band of 8 rows; tile c holds pixels
     // (16 c + i % 16, i / 16) for i < 128, the last tile 5 wide
    const int32_t w = (1 << 27) + 5, h = 8;
    const uint32_t last = (uint32_t)w / 16;  // the narrow tile
    const uint32_t cases[][3] = {{0, 0, 0}, {17, 1, 1}, {127, 15, 7}, {128 * 5 + 33, 16 * 5 + 1, 2},
                                 {128 * last + 7, 16 * last + 2, 1}, {128 * last + 39, 16 * last + 4, 7}};
    for (const auto& c : cases) {
      const uint32_t want = c[2] * (uint32_t)w + c[1];
      if (flg::tile_order(c[0], 1, w, h) != want) {
        std::printf("FAIL tile_order(%u) at %d x %d: %u, expected %u\n", c[0], w, h,
                    flg::tile_order(c[0], 1, w, h), want);
        ++failures;
      }
    }
  }
  std::printf("tile_order: checked\n");
}

// radical_inverse divides in double with a biased refined reciprocal; its
// result must equal RadicalInverse_'s integer digit loop (sampler.h:59-72)
// for every prime below 2^18 and every 32-bit index.  Checked: the exact
// multiples of b, b^2 and b^3 and their neighbours at the top of the range
// and halfway (where the rounding argument is tightest), with remainders at
// the edges of the digit splits, prime powers, random indices below 2^32 and
// below 2^24 for every prime, every index below 2^24 for bases 2, 3, 5 and 7,
// and 200,000 random 32-bit indices for bases 2, 3 and 5, around 2^17 and at
// the top of the table.
float radical_inverse_int(uint32_t base, uint64_t a) {
  const float inv_base = 1.0f / (float)base;
  uint64_t reversed = 0;
  float inv_base_n = 1.0f;
  while (a) {
    const uint64_t next = a / base;
    reversed = reversed * base + (a - next * base);
    inv_base_n *= inv_base;
    a = next;
  }
  return (float)reversed * inv_base_n;
}

void check_radical_inverse() {
  std::vector<int32_t> primes;
  {
    std::vector<char> comp(1 << 18, 0);
    for (int i = 2; i < (1 << 18); ++i) {
      if (comp[i]) continue;
      primes.push_back(i);
      for (int64_t j = (int64_t)i * i; j < (1 << 18); j += i) comp[j] = 1;
    }
  }
  const uint64_t top = (uint64_t)1 << 32;
  std::mt19937_64 gen(11);
  std::uniform_int_distribution<uint64_t> U32(0, top - 1), U24(0, (1u << 24) - 1);
  int64_t checked = 0;
  auto one = [&](int32_t p, uint64_t a) {
    const float got = flg::radical_inverse(p, 1.0f / (float)p, (uint32_t)a),
                want = radical_inverse_int((uint32_t)p, a);
    ++checked;
    if (flg::f2u(got) != flg::f2u(want) && failures < 20) {
      std::printf("FAIL radical_inverse(%d, %llu) = %.9g, integer loop %.9g\n", p, (unsigned long long)a, got,
                  want);
      ++failures;
    }
  };
  for (int32_t p : primes) {
    const uint64_t b = (uint64_t)p, kmax = (top - 1) / b;
    for (uint64_t k : {kmax, kmax - 1, kmax / 2 + 1, (uint64_t)(((1ull << 24) - 1) / b), (uint64_t)1, (uint64_t)2})
      for (int d = -1; d <= 1; ++d) {
        const int64_t a = (int64_t)(k * b) + d;
        if (a >= 0 && (uint64_t)a < top) one(p, (uint64_t)a);
      }
    for (uint64_t pw = b; pw < top; pw *= b)
      for (int d = -1; d <= 1; ++d) one(p, pw + d);
    // multiples of b^2 (the two-digit quotient) at the top, halfway and low
    const uint64_t B = b * b;
    if (B < top) {
      const uint64_t kb = (top - 1) / B;
      for (uint64_t k : {kb, kb - 1, kb / 2 + 1, (uint64_t)1, (uint64_t)2})
        for (int64_t d : {(int64_t)-1, (int64_t)0, (int64_t)1, (int64_t)b - 1, (int64_t)b, (int64_t)B - 1}) {
          const int64_t a = (int64_t)(k * B) + d;
          if (a >= 0 && (uint64_t)a < top) one(p, (uint64_t)a);
        }
    }
    // multiples of b^3 (two two-digit steps), with remainders at the edges of
    // the two- and one-digit splits of the low three digits
    const uint64_t B3 = B * b;
    if (B3 < top) {
      const uint64_t kc = (top - 1) / B3;
      for (uint64_t k : {kc, kc - 1, kc / 2 + 1, (uint64_t)1, (uint64_t)2, (uint64_t)b - 1, (uint64_t)b})
        for (int64_t d : {(int64_t)-1, (int64_t)0, (int64_t)1, (int64_t)b - 1, (int64_t)b, (int64_t)B - 1,
                          (int64_t)B, (int64_t)B + 1, (int64_t)(B3 - 1), (int64_t)(B3 - b)}) {
          const int64_t a = (int64_t)(k * B3) + d;
          if (a >= 0 && (uint64_t)a < top) one(p, (uint64_t)a);
        }
    }
    for (int i = 0; i < 32; ++i) one(p, U32(gen));
    for (int i = 0; i < 32; ++i) one(p, U24(gen));
  }
  for (int32_t p : {2, 3, 5, 7})
    for (uint64_t a = 0; a < (1u << 24); ++a) one(p, a);
  for (int32_t p : {2, 3, 5, 131071, 131101, 262139})
    for (int i = 0; i < 200000; ++i) one(p, U32(gen));
  std::printf("radical_inverse: %lld cases checked\n", (long long)checked);
}

int main() {
  check_radical_inverse();
  check_tile_order();
  {  // bvh_test: MakeMeshedSphere(200, 200, 5), rays from inside
    const int nt = 200, np = 200;
    const int32_t ntris = flg_scene_uv_sphere(nt, np, 5.0f, nullptr, nullptr);
    std::vector<float> v(size_t(nt) * np * 3);
    std::vector<int32_t> ind(size_t(ntris) * 3);
    flg_scene_uv_sphere(nt, np, 5.0f, v.data(), ind.data());
    std::vector<float> rays(20000 * 8);
    orc_bvh_test_rays(20000, rays.data());
    run_case("uv_sphere", v, ind, rays);
  }
  {  // soup of 20000 triangles, rays from outside through the cube
    const int n = 20000;
    std::vector<float> v(size_t(n) * 9);
    std::vector<int32_t> ind(size_t(n) * 3);
    flg_scene_soup(42, 60, n, v.data(), ind.data());
    std::mt19937 gen(3);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    std::vector<float> rays;
    for (int i = 0; i < 20000; ++i) {
      orc::V3 o = orc::v3(U(gen) * 2 - 0.5f, U(gen) * 2 - 0.5f, -1.0f);
      orc::V3 tgt = orc::v3(U(gen), U(gen), U(gen));
      orc::V3 d = orc::normalize(tgt - o);
      rays.insert(rays.end(), {o.x, o.y, o.z, INFINITY, d.x, d.y, d.z, 0.0f});
    }
    run_case("soup", v, ind, rays);
  }
  {  // displaced octahedral sphere, diffuse-like rays from points near the surface
    const int level = 6;
    int64_t nv = 0;
    const int64_t ntris = flg_scene_octasphere_counts(level, &nv);
    std::vector<float> v(size_t(nv) * 3);
    std::vector<int32_t> ind(size_t(ntris) * 3);
    flg_scene_octasphere(level, 0.1f, 7, 2.0f, v.data(), ind.data());
    std::mt19937 gen(5);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::vector<float> rays;
    for (int i = 0; i < 20000; ++i) {
      orc::V3 o = orc::v3(U(gen) * 0.5f, U(gen) * 0.5f, U(gen) * 0.5f);
      orc::V3 d = orc::normalize(orc::v3(U(gen), U(gen), U(gen)));
      rays.insert(rays.end(), {o.x, o.y, o.z, INFINITY, d.x, d.y, d.z, 0.0f});
    }
    run_case("octasphere", v, ind, rays);
  }
  std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
  return failures ? 1 : 0;
}
