// Tests of the C++ drop-ins (fledge_gpu.hpp) against the CPU oracle, in the
// style of the reference's bvh_test harness (src/experimental/bvh_test.hpp:
// 107-168) and its shape/volume tests.
//   dropin_test --cpu   host-only checks (no GPU call): GridVolume::tr/sample
//                       against the oracle's OpenVDBVolume::tr/sample
//   dropin_test         everything, GPU included
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "fledge_gpu.hpp"
#include "fledge_scene.h"
#include "oracle_impl.hpp"

extern "C" void orc_bvh_test_rays(int64_t n, float* rays);

using fledge::Vector3f;

namespace {

int g_failures = 0;

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::printf(__VA_ARGS__);                            \
      std::printf("\n");                                   \
      std::fflush(stdout);                                 \
      ++g_failures;                                        \
      return;                                              \
    }                                                      \
  } while (0)

uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

bool same_bits(const Vector3f& a, orc::V3 b) {
  return bits(a[0]) == bits(b.x) && bits(a[1]) == bits(b.y) && bits(a[2]) == bits(b.z);
}

// A fledge::Sampler over the oracle's Halton stream of one pixel sample.
struct HaltonStream : fledge::Sampler {
  orc::Halton h;
  HaltonStream(int res, int x, int y, uint64_t s) : h(1, res, res) { h.start(x, y, s); }
  fledge::Float get1D() override { return h.get1D(); }
};

// ------------------------------------------------------------ host only
void test_grid_volume_host() {
  const int n = 24;
  std::vector<float> grid(size_t(n) * n * n);
  flg_scene_fbm_grid(n, n, n, 5, 4.0f, 5, 0.0f, 1.0f, grid.data());
  const float bounds[6] = {0, 0, 0, 1, 1, 1};
  const float medium[3] = {6.4f, 1.6f, 0.3f};
  fledge::GridVolume vol(grid.data(), n, n, n, fledge::AABB(Vector3f(0.0f), Vector3f(1.0f)),
                         Vector3f(medium[0]), Vector3f(medium[1]), medium[2]);
  orc::VolScene sc;
  orc::init_vol_grid(sc, grid.data(), n, n, n, bounds, medium);
  CHECK(bits(vol.maxDensity()) == bits(sc.med.max_density), "max density");
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  int ntr = 0, nsample_hits = 0;
  for (int i = 0; i < 4000; ++i) {
    orc::V3 o = orc::v3(U(gen) * 3 - 1, U(gen) * 3 - 1, -1.0f + U(gen));
    orc::V3 d = orc::normalize(orc::v3(0.5f + U(gen) * 0.2f, 0.5f + U(gen) * 0.2f, 1.0f) - o);
    const float tmax = (i % 3 == 0) ? 1.5f : INFINITY;
    fledge::Ray ray(Vector3f(o.x, o.y, o.z), Vector3f(d.x, d.y, d.z), tmax);
    HaltonStream a(128, i % 128, (i / 128) % 128, 1 + i / 16384), b = a;
    const Vector3f tr = vol.tr(ray, a);
    const float tr_ref = orc::vol_tr_pub(sc, b.h, o, d, tmax);
    CHECK(bits(tr[0]) == bits(tr_ref), "tr ray %d: %.9g vs %.9g", i, tr[0], tr_ref);
    CHECK(a.h.cnt == b.h.cnt, "tr ray %d: draw count", i);
    ntr += tr[0] < 1.0f;
    fledge::VInteraction vi;
    bool ok = false;
    const Vector3f f = vol.sample(ray, a, vi, ok);
    bool ok_ref = false;
    orc::V3 pos{0, 0, 0};
    const float f_ref = orc::vol_sample_pub(sc, b.h, o, d, tmax, &ok_ref, &pos);
    CHECK(ok == ok_ref && bits(f[0]) == bits(f_ref), "sample ray %d", i);
    if (ok) {
      CHECK(same_bits(vi.m_p, pos), "sample ray %d: position", i);
      CHECK(vi.m_g == medium[2] && bits(vi.m_wo[0]) == bits(-d.x), "sample ray %d: interaction", i);
      ++nsample_hits;
    }
  }
  CHECK(ntr > 100 && nsample_hits > 100, "too few interactions (%d, %d)", ntr, nsample_hits);
  std::printf("PASS grid_volume_host (%d attenuated, %d collisions)\n", ntr, nsample_hits);
}

// A sparse GridVolume (8^3 leaves with a brick pool) against the dense one
// holding the same values: bit-equal densities at random points, at voxel
// centres and on brick borders (inside, on the zero border, outside), equal
// tr/sample results and draw counts, equal maxima; grids whose sides are not
// multiples of 8; no leaves; every leaf; the origins it refuses.
void test_sparse_volume_host() {
  const int nx = 37, ny = 29, nz = 21;
  std::vector<float> grid(size_t(nx) * ny * nz);
  flg_scene_fbm_grid(nx, ny, nz, 3, 4.0f, 5, 0.0f, 1.0f, grid.data());
  const int bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 7) / 8;
  std::mt19937 gen(17);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  for (int mode = 0; mode < 3; ++mode) {  // 0: half the bricks, 1: every brick, 2: none
    std::vector<float> g(grid);
    std::vector<int32_t> origin;
    std::vector<float> values;
    for (int k = 0; k < bz; ++k)
      for (int j = 0; j < by; ++j)
        for (int i = 0; i < bx; ++i) {
          const bool keep = mode == 1 || (mode == 0 && U(gen) < 0.5f);
          for (int z = 8 * k; z < std::min(8 * k + 8, nz); ++z)
            for (int y = 8 * j; y < std::min(8 * j + 8, ny); ++y)
              for (int x = 8 * i; x < std::min(8 * i + 8, nx); ++x)
                if (!keep) g[(size_t(z) * ny + y) * nx + x] = 0.0f;
          if (!keep) continue;
          origin.insert(origin.end(), {8 * i, 8 * j, 8 * k});
          for (int q = 0; q < 512; ++q) {
            const int x = 8 * i + (q & 7), y = 8 * j + ((q >> 3) & 7), z = 8 * k + (q >> 6);
            // voxels past the grid's edge hold junk: they must be ignored
            values.push_back(x < nx && y < ny && z < nz ? g[(size_t(z) * ny + y) * nx + x] : 1e30f);
          }
        }
    const fledge::AABB box(Vector3f(-0.5f, 0.0f, 0.25f), Vector3f(1.0f, 1.25f, 1.0f));
    fledge::GridVolume dense(g.data(), nx, ny, nz, box, Vector3f(6.4f), Vector3f(1.6f), 0.3f);
    fledge::GridVolume sparse(int(origin.size() / 3), origin.data(), values.data(), nx, ny, nz, box,
                              Vector3f(6.4f), Vector3f(1.6f), 0.3f);
    CHECK(bits(dense.maxDensity()) == bits(sparse.maxDensity()), "mode %d: max density %.9g vs %.9g", mode,
          dense.maxDensity(), sparse.maxDensity());
    const float lo[3] = {-0.5f, 0.0f, 0.25f}, ext[3] = {1.5f, 1.25f, 0.75f};
    const int n[3] = {nx, ny, nz};
    int nonzero = 0;
    for (int t = 0; t < 200000; ++t) {
      Vector3f p;
      for (int a = 0; a < 3; ++a) {
        float u = U(gen) * 1.2f - 0.1f;
        if (t % 4 == 1) {  // on a brick border: index-space coordinate 8 m - 0.5 + small
          const int m = int(U(gen) * (n[a] / 8 + 2)) - 1;
          u = (8.0f * m + 0.5f * U(gen) - 0.25f + 0.5f) / n[a];
        } else if (t % 4 == 2) {  // a voxel centre, or halfway between two
          u = (std::floor(U(gen) * (n[a] + 2)) - 1.0f + 0.5f * (t % 8 == 2 ? 1.0f : 0.0f)) / n[a];
        }
        p[a] = lo[a] + u * ext[a];
      }
      const float a = dense.density(p), b = sparse.density(p);
      CHECK(bits(a) == bits(b), "mode %d: density at (%.9g, %.9g, %.9g): %.9g vs %.9g", mode, p[0], p[1], p[2], a,
            b);
      nonzero += a != 0.0f;
    }
    CHECK(mode == 2 ? nonzero == 0 : nonzero > 20000, "mode %d: %d non-zero lookups", mode, nonzero);
    int ncoll = 0;
    for (int i = 0; i < 3000; ++i) {
      const Vector3f o(U(gen) * 3 - 1.5f, U(gen) * 3 - 1, -1.0f + U(gen));
      const float tx = U(gen) * 0.6f - 0.2f, ty = 0.5f + U(gen) * 0.5f, tz = 0.6f + U(gen) * 0.2f;
      const float dx = tx - o[0], dy = ty - o[1], dz = tz - o[2];
      const float len = std::sqrt(dx * dx + dy * dy + dz * dz);
      const Vector3f d(dx / len, dy / len, dz / len);
      const fledge::Ray ray(o, d, (i % 3 == 0) ? 1.5f : INFINITY);
      HaltonStream a(128, i % 128, (i / 128) % 128, 1 + i / 16384), b = a;
      CHECK(bits(dense.tr(ray, a)[0]) == bits(sparse.tr(ray, b)[0]) && a.h.cnt == b.h.cnt, "mode %d: tr ray %d",
            mode, i);
      fledge::VInteraction va, vb;
      bool oa = false, ob = false;
      const Vector3f fa = dense.sample(ray, a, va, oa), fb = sparse.sample(ray, b, vb, ob);
      CHECK(oa == ob && bits(fa[0]) == bits(fb[0]) && a.h.cnt == b.h.cnt, "mode %d: sample ray %d", mode, i);
      if (oa) {
        CHECK(bits(va.m_p[0]) == bits(vb.m_p[0]) && bits(va.m_p[1]) == bits(vb.m_p[1]) &&
                  bits(va.m_p[2]) == bits(vb.m_p[2]),
              "mode %d: sample ray %d position", mode, i);
        ++ncoll;
      }
    }
    CHECK(mode == 2 ? ncoll == 0 : ncoll > 100, "mode %d: %d collisions", mode, ncoll);
    std::printf("PASS sparse_volume_host mode %d (%zu leaves, %d collisions)\n", mode, origin.size() / 3, ncoll);
  }
  // origins it refuses
  const std::vector<float> v(512 * 2, 1.0f);
  const int32_t bad[][6] = {{0, 0, 4, 8, 8, 8}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 40, 0, 0}, {-8, 0, 0, 0, 0, 0}};
  for (const auto& o : bad) {
    bool threw = false;
    try {
      fledge::GridVolume s(2, o, v.data(), nx, ny, nz, fledge::AABB(Vector3f(0.0f), Vector3f(1.0f)), Vector3f(1.0f),
                           Vector3f(0.0f), 0.0f);
    } catch (const fledge::gpu::Error& e) {
      threw = e.status() == FLG_ERR_INVALID;
    }
    CHECK(threw, "leaf origins (%d %d %d), (%d %d %d) must be refused", o[0], o[1], o[2], o[3], o[4], o[5]);
  }
  std::printf("PASS sparse_volume_host errors\n");
}

void test_host_errors() {
  float one = 1.0f;
  bool threw = false;
  try {
    fledge::GridVolume v(&one, 0, 1, 1, fledge::AABB(Vector3f(0.0f), Vector3f(1.0f)), Vector3f(1.0f),
                         Vector3f(0.0f), 0.0f);
  } catch (const fledge::gpu::Error& e) {
    threw = e.status() == FLG_ERR_INVALID;
  }
  CHECK(threw, "GridVolume with an empty grid must throw");
  threw = false;
  try {
    fledge::GridVolume v(&one, 1, 1, 1, fledge::AABB(Vector3f(0.0f), Vector3f(0.0f)), Vector3f(1.0f),
                         Vector3f(0.0f), 0.0f);
  } catch (const fledge::gpu::Error&) {
    threw = true;
  }
  CHECK(threw, "GridVolume with empty bounds must throw");
  std::printf("PASS host_errors\n");
}

// ------------------------------------------------------------------ GPU
struct Mesh {
  std::vector<float> verts;
  std::vector<int32_t> ind;
  std::vector<Vector3f> p;
  fledge::experimental::InternalTriangleMesh im{};
  void finish() {
    p.resize(verts.size() / 3);
    for (size_t i = 0; i < p.size(); ++i) p[i] = Vector3f(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    im.nInd = (int)ind.size();
    im.nVert = (int)p.size();
    im.ind = ind.data();
    im.p = p.data();
    im.n = nullptr;
    im.uv = nullptr;
  }
};

Mesh uv_sphere(int nt, int np, float r) {
  Mesh m;
  // MakeMeshedSphere layout: nt*np vertices; the triangle count is queried
  const int32_t ntris = flg_scene_uv_sphere(nt, np, r, nullptr, nullptr);
  m.verts.resize(size_t(nt) * np * 3);
  m.ind.resize(size_t(ntris) * 3);
  flg_scene_uv_sphere(nt, np, r, m.verts.data(), m.ind.data());
  m.finish();
  return m;
}

void test_lbvh_builder(fledge::gpu::Context& ctx) {
  // bvh_test: MakeMeshedSphere, rays from [0, 0.5)^3 with normalised
  // directions in the positive octant; every ray starts inside the sphere
  Mesh m = uv_sphere(200, 200, 5.0f);
  fledge::gpu_memory_resource mr;
  fledge::experimental::LBVHBuilder bvh(&m.im, &mr, &ctx);
  bvh.build();
  CHECK(bvh.getMemoryResource() == &mr, "memory resource");
  orc::Lbvh ref;
  orc::lbvh_build(ref, m.verts.data(), m.ind.data(), (int32_t)m.ind.size() / 3, (int32_t)m.p.size());
  auto bd = bvh.getBound();
  CHECK(same_bits(bd.lower, ref.glo) && same_bits(bd.upper, ref.ghi), "global bound");

  const int N = 20000;
  std::vector<float> rays(size_t(N) * 8);
  orc_bvh_test_rays(N, rays.data());
  std::vector<fledge::experimental::BVHBuilderBase::BVHRayHit> batch(N);
  for (int i = 0; i < N; ++i) {
    batch[i].ray_o = Vector3f(rays[8 * i], rays[8 * i + 1], rays[8 * i + 2]);
    batch[i].ray_d = Vector3f(rays[8 * i + 4], rays[8 * i + 5], rays[8 * i + 6]);
  }
  std::vector<int32_t> ids(N);
  bvh.intersectBatch(batch.data(), ids.data(), N);
  for (int i = 0; i < N; ++i) {
    fledge::experimental::BVHBuilderBase::BVHRayHit h;
    h.ray_o = batch[i].ray_o;
    h.ray_d = batch[i].ray_d;
    int32_t id = -2;
    const bool hit = bvh.intersect(h, &id);
    orc::V3 o = orc::v3(rays[8 * i], rays[8 * i + 1], rays[8 * i + 2]);
    orc::V3 d = orc::v3(rays[8 * i + 4], rays[8 * i + 5], rays[8 * i + 6]);
    float t_ref = 0;
    orc::V3 ng_ref{0, 0, 0};
    const int32_t id_ref = orc::lbvh_trace(ref, o, d, rays[8 * i + 3], false, &t_ref, &ng_ref, nullptr);
    CHECK(hit && id == id_ref, "ray %d: host mirror id %d vs oracle %d", i, id, id_ref);
    CHECK(bits(h.tfar) == bits(t_ref) && same_bits(h.hit_ng, ng_ref), "ray %d: host mirror t/ng", i);
    CHECK(batch[i].hit && ids[i] == id && bits(batch[i].tfar) == bits(h.tfar) &&
              same_bits(batch[i].hit_ng, ng_ref) && same_bits(batch[i].hit_ns, ng_ref),
          "ray %d: GPU batch differs from the host mirror", i);
  }
  // bvh_test correctness2: the fixed ray hits
  fledge::experimental::BVHBuilderBase::BVHRayHit fixed;
  fixed.ray_o = Vector3f(0.345564f, 0.422557f, 0.259793f);
  fixed.ray_d = Vector3f(0.524748325347900390625f, 0.714325249195098876953125f, 0.46301043033599853515625f);
  CHECK(bvh.intersect(fixed), "fixed ray must hit");
  // a ray leaving the sphere's bound misses
  fledge::experimental::BVHBuilderBase::BVHRayHit away;
  away.ray_o = Vector3f(20.0f, 0.0f, 0.0f);
  away.ray_d = Vector3f(1.0f, 0.0f, 0.0f);
  CHECK(!bvh.intersect(away) && !away.hit, "ray away from the mesh must miss");
  std::printf("PASS lbvh_builder (%d rays: host mirror, GPU batch and oracle agree bit for bit)\n", N);
}

void test_lbvh_accel(fledge::gpu::Context& ctx) {
  const int ntris = 4096;
  Mesh m;
  m.verts.resize(size_t(ntris) * 9);
  m.ind.resize(size_t(ntris) * 3);
  flg_scene_soup(42, 54, ntris, m.verts.data(), m.ind.data());
  m.finish();
  std::vector<int> ind(m.ind.begin(), m.ind.end());
  fledge::LBVHAccel accel(m.p, ind, ctx);
  CHECK(accel.getMaterial() == nullptr && accel.getAreaLight() == nullptr, "Accel has no material");
  orc::Lbvh ref;
  orc::lbvh_build(ref, m.verts.data(), m.ind.data(), ntris, ntris * 3);
  const fledge::AABB b = accel.getBound();
  CHECK(same_bits(b.m_min, ref.glo) && same_bits(b.m_max, ref.ghi), "bound");

  orc::Camera cam(orc::v3(0.5f, 0.5f, -1.5f), orc::v3(0.5f, 0.5f, 0.0f), orc::v3(0, 1, 0), 45.0f);
  const int W = 96, H = 96;
  std::vector<fledge::Ray> rays;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      orc::V3 o, d;
      cam.generate(x + 0.5f, y + 0.5f, W, H, &o, &d);
      rays.emplace_back(Vector3f(o.x, o.y, o.z), Vector3f(d.x, d.y, d.z));
    }
  const std::vector<flg_hit> hits = accel.intersectBatch(rays);
  int nhit = 0;
  for (size_t i = 0; i < rays.size(); ++i) {
    fledge::Ray r = rays[i];
    fledge::SInteraction isect;
    int32_t id = -2;
    const bool hit = accel.intersect(r, isect, &id);
    float t_ref = 0;
    orc::V3 ng_ref{0, 0, 0};
    orc::V3 o = orc::v3(r.m_o[0], r.m_o[1], r.m_o[2]), d = orc::v3(r.m_d[0], r.m_d[1], r.m_d[2]);
    const int32_t id_ref = orc::lbvh_trace(ref, o, d, INFINITY, false, &t_ref, &ng_ref, nullptr);
    CHECK(id == id_ref && hits[i].id == id_ref, "ray %zu: ids %d / %d vs oracle %d", i, id, hits[i].id, id_ref);
    if (!hit) {
      CHECK(std::isinf(r.m_tMax) && isect.m_primitive == nullptr, "ray %zu: a miss must not touch ray/isect", i);
      continue;
    }
    ++nhit;
    CHECK(bits(r.m_tMax) == bits(t_ref) && bits(hits[i].t) == bits(t_ref), "ray %zu: t", i);
    CHECK(same_bits(isect.m_ng, ng_ref) && same_bits(isect.m_ns, ng_ref), "ray %zu: normal", i);
    CHECK(isect.m_primitive == &accel, "ray %zu: primitive", i);
    const Vector3f p = rays[i](t_ref);
    CHECK(bits(isect.m_p[0]) == bits(p[0]) && bits(isect.m_wo[2]) == bits(-rays[i].m_d[2]), "ray %zu: p/wo", i);
    // the shortened ray finds the same hit (t <= tMax accepted) and nothing nearer
    fledge::SInteraction again;
    CHECK(accel.intersect(r, again, &id) && id == id_ref, "ray %zu: re-intersect", i);
  }
  CHECK(nhit > 100, "too few hits (%d)", nhit);  // sparse soup: about 5 % of the rays hit
  bool threw = false;
  try {
    std::vector<int> bad(ind.begin(), ind.begin() + 4);
    fledge::LBVHAccel broken(m.p, bad, ctx);
  } catch (const fledge::gpu::Error& e) {
    threw = e.status() == FLG_ERR_INVALID;
  }
  CHECK(threw, "index count not a multiple of 3 must throw");
  std::printf("PASS lbvh_accel (%d hits of %zu rays)\n", nhit, rays.size());
}

// ------------------------------------------------------------------ PLY
// An octahedral sphere written as a PLY file (ascii or binary little-endian)
// with an extra vertex property and a comment, for MakeTriangleMesh.
struct PlyCase {
  std::vector<float> v;
  std::vector<int32_t> ind;
};

PlyCase sphere_case(int level, float dx) {
  PlyCase c;
  int64_t nv = 0;
  const int64_t nt = flg_scene_octasphere_counts(level, &nv);
  c.v.resize(size_t(nv) * 3);
  c.ind.resize(size_t(nt) * 3);
  flg_scene_octasphere(level, 0.05f, 7, 2.0f, c.v.data(), c.ind.data());
  for (size_t i = 0; i < c.v.size(); i += 3) c.v[i] += dx;
  return c;
}

std::string write_ply(const std::string& name, bool binary, const PlyCase& c) {
  const std::string path = "/tmp/fledge_dropin_" + std::to_string((long)getpid()) + "_" + name + ".ply";
  FILE* f = std::fopen(path.c_str(), "wb");
  const size_t nv = c.v.size() / 3, nt = c.ind.size() / 3;
  std::fprintf(f, "ply\nformat %s 1.0\ncomment written by dropin_test\n", binary ? "binary_little_endian" : "ascii");
  std::fprintf(f, "element vertex %zu\nproperty float x\nproperty float y\nproperty float z\n", nv);
  std::fprintf(f, "property uchar flags\nelement face %zu\nproperty list uchar int vertex_indices\nend_header\n", nt);
  for (size_t i = 0; i < nv; ++i) {
    if (binary) {
      std::fwrite(&c.v[3 * i], 4, 3, f);
      const unsigned char fl = (unsigned char)(i & 0xff);
      std::fwrite(&fl, 1, 1, f);
    } else {
      std::fprintf(f, "%.9g %.9g %.9g %d\n", c.v[3 * i], c.v[3 * i + 1], c.v[3 * i + 2], (int)(i & 0xff));
    }
  }
  for (size_t t = 0; t < nt; ++t) {
    if (binary) {
      const unsigned char three = 3;
      std::fwrite(&three, 1, 1, f);
      std::fwrite(&c.ind[3 * t], 4, 3, f);
    } else {
      std::fprintf(f, "3 %d %d %d\n", c.ind[3 * t], c.ind[3 * t + 1], c.ind[3 * t + 2]);
    }
  }
  std::fclose(f);
  return path;
}

void test_ply_host() {
  const PlyCase c = sphere_case(3, 0.0f);
  for (bool binary : {false, true}) {
    const std::string path = write_ply(binary ? "bin" : "ascii", binary, c);
    auto m = fledge::MakeTriangleMesh(path);
    std::remove(path.c_str());
    CHECK(m->mesh.nVert == (int)(c.v.size() / 3) && m->mesh.nInd == (int)c.ind.size(), "ply counts");
    CHECK(m->mesh.n == nullptr && m->mesh.uv == nullptr, "ply: no normals or texture coordinates");
    for (int i = 0; i < m->mesh.nVert; ++i)
      for (int k = 0; k < 3; ++k)
        CHECK(bits(m->mesh.p[i][k]) == bits(c.v[3 * i + k]), "ply vertex %d", i);
    for (int i = 0; i < m->mesh.nInd; ++i) CHECK(m->mesh.ind[i] == c.ind[i], "ply index %d", i);
  }
  bool threw = false;
  try {
    (void)fledge::MakeTriangleMesh("/tmp/fledge_dropin_no_such_file.ply");
  } catch (const fledge::gpu::Error& e) {
    threw = e.status() == FLG_ERR_INVALID;
  }
  CHECK(threw, "a missing file must throw");
  std::printf("PASS ply_host (ascii and binary, %zu vertices)\n", c.v.size() / 3);
}

// Scene::parseXML's subset through ParseSceneXML on a file written here in
// the reference's scene format, then its PLY shapes through MakeTriangleMesh;
// and the files the reference refuses.
void test_scene_file_host() {
  const std::string dir = "/tmp/fledge_dropin_" + std::to_string((long)getpid()) + "_scene";
  ::mkdir(dir.c_str(), 0755);
  const PlyCase c = sphere_case(3, 0.0f);
  const std::string ply = write_ply("scene_mesh", false, c);
  (void)std::rename(ply.c_str(), (dir + "/mesh.ply").c_str());
  const char* text = R"(<?xml version='1.0' encoding='utf-8'?>
<!-- a scene in the reference's format -->
<scene version="0.6.0">
  <integrator type="volpath">
    <integer name="maxDepth" value="12" />
  </integrator>
  <sensor type="perspective">
    <float name="fov" value="35.5" />
    <string name="fovAxis" value="y" />
    <transform name="toWorld">
      <lookat up="0, 1, 0" origin="1.5, 2, -7.25" target="0,0.5,0" />
    </transform>
    <sampler type="independent"><integer name="sampleCount" value="24" /></sampler>
    <film type="hdrfilm">
      <integer name="width" value="320" />
      <integer name="height" value="200" />
      <rfilter type="box"><float name="size" value="0.5" /></rfilter>
    </film>
  </sensor>
  <medium id="m" type="homogeneous"><spectrum name="sigmaS" value="3.0" /></medium>
  <shape type="ply">
    <string name="filename" value="mesh.ply" />
    <bsdf type="roughconductor"><float name="alpha" value="0.1" /></bsdf>
    <ref name="interior" id="m" />
  </shape>
  <shape type="sphere">
    <point x="0.5" y="-1" z="2" />
    <float name="radius" value="0.75" />
    <bsdf type="dielectric"><float name="intIOR" value="1.5" /><float name="extIOR" value="1.25" /></bsdf>
  </shape>
  <emitter type="envmap"><string name="filename" value="/abs/sky.exr" /></emitter>
</scene>
)";
  const std::string xml = dir + "/scene.xml";
  {
    FILE* f = std::fopen(xml.c_str(), "wb");
    std::fputs(text, f);
    std::fclose(f);
  }
  fledge::Scene sc;
  const fledge::SceneFileDesc d = fledge::ParseSceneXML(xml, sc);
  CHECK(sc.m_integrator_type == fledge::EIntegratorType::EVolPathIntegrator && sc.m_maxDepth == 12, "integrator");
  CHECK(sc.m_resX == 320 && sc.m_resY == 200 && sc.m_SPP == 24 && sc.m_FoV == 35.5f, "sensor");
  CHECK(sc.m_origin[0] == 1.5f && sc.m_origin[2] == -7.25f && sc.m_target[1] == 0.5f && sc.m_up[1] == 1.0f,
        "lookat");
  CHECK(d.shapes.size() == 2 && d.envmaps.size() == 1 && d.envmaps[0] == "/abs/sky.exr", "entries");
  CHECK(d.shapes[0].type == "ply" && d.shapes[0].filename == dir + "/mesh.ply" &&
            d.shapes[0].bsdf == "roughconductor",
        "ply shape");
  CHECK(d.shapes[1].type == "sphere" && d.shapes[1].radius == 0.75f && d.shapes[1].center[1] == -1.0f &&
            d.shapes[1].int_ior == 1.5f && d.shapes[1].ext_ior == 1.25f,
        "sphere shape");
  auto m = fledge::MakeTriangleMesh(d.shapes[0].filename);
  CHECK(m->mesh.nInd == (int)c.ind.size(), "the scene's ply mesh");
  // what the reference refuses
  const char* bad[] = {"<scene><integrator type=\"path\"/></scene>",
                       "<scene version=\"0.6.0\"><integrator type=\"bdpt\"/></scene>",
                       "<scene version=\"0.6.0\"><integrator type=\"path\"/><sensor type=\"orthographic\"/></scene>",
                       "<scene version=\"0.6.0\"><integrator type=\"path\"/>",
                       "<scene version=\"0.6.0\"><integrator type=\"path\"/><sensor type=\"perspective\">"
                       "<transform><lookat up=\"0,1\" origin=\"0,0,0\" target=\"0,0,1\"/></transform></sensor></scene>"};
  std::string deep = "<scene version=\"0.6.0\">";  // nesting deep enough to exhaust a recursive parser's stack
  for (int k = 0; k < 200000; ++k) deep += "<a>";
  std::vector<std::string> files(std::begin(bad), std::end(bad));
  files.push_back(deep);
  int refused = 0;
  for (const std::string& b : files) {
    FILE* f = std::fopen(xml.c_str(), "wb");
    std::fputs(b.c_str(), f);
    std::fclose(f);
    try {
      fledge::Scene s2;
      (void)fledge::ParseSceneXML(xml, s2);
    } catch (const fledge::gpu::Error& e) {
      refused += e.status() == FLG_ERR_INVALID;
    }
  }
  std::remove(xml.c_str());
  std::remove((dir + "/mesh.ply").c_str());
  ::rmdir(dir.c_str());
  CHECK(refused == 6, "the malformed scene files must be refused (%d of 6)", refused);
  std::printf("PASS scene_file_host (2 shapes, 1 emitter, 6 refused files)\n");
}

// Scene::intersect through SceneAccel over three mesh primitives read from
// PLY files (the third repeats the first: they tie wherever it is hit), host
// mirror and GPU batch against the oracle's NaiveBVHAccel restatement.
void test_scene_accel(fledge::gpu::Context& ctx) {
  const PlyCase a = sphere_case(4, 0.0f), b = sphere_case(3, 0.7f);
  std::vector<std::unique_ptr<fledge::PlyMesh>> plys;
  for (const PlyCase* c : {&a, &b, &a}) {
    const std::string path = write_ply("scene" + std::to_string(plys.size()), true, *c);
    plys.push_back(fledge::MakeTriangleMesh(path));
    std::remove(path.c_str());
  }
  std::vector<std::unique_ptr<fledge::LBVHAccel>> accels;
  std::vector<const fledge::LBVHAccel*> ptrs;
  for (auto& m : plys) {
    accels.push_back(std::make_unique<fledge::LBVHAccel>(m->mesh, ctx));
    ptrs.push_back(accels.back().get());
  }
  fledge::SceneAccel scene(ptrs, ctx);
  std::vector<orc::Lbvh> refs(3);
  std::vector<const orc::Lbvh*> rp;
  for (int i = 0; i < 3; ++i) {
    const PlyCase& c = i == 1 ? b : a;
    orc::lbvh_build(refs[i], c.v.data(), c.ind.data(), (int)c.ind.size() / 3, (int)c.v.size() / 3);
    rp.push_back(&refs[i]);
  }
  orc::Camera cam(orc::v3(0.3f, 0.2f, -4.0f), orc::v3(0.3f, 0.0f, 0.0f), orc::v3(0, 1, 0), 40.0f);
  const int W = 96, H = 96;
  std::vector<fledge::Ray> rays;
  std::vector<float> rr;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      orc::V3 o, d;
      cam.generate(x + 0.5f, y + 0.5f, W, H, &o, &d);
      rays.emplace_back(Vector3f(o.x, o.y, o.z), Vector3f(d.x, d.y, d.z));
      rr.insert(rr.end(), {o.x, o.y, o.z, INFINITY, d.x, d.y, d.z, 0.0f});
    }
  std::vector<int32_t> mref(rays.size()), iref(rays.size());
  std::vector<float> tref(rays.size());
  orc::scene_trace(rp, false, rr.data(), (int64_t)rays.size(), mref.data(), iref.data(), tref.data(), 4);
  const std::vector<flg_scene_hit> hits = scene.intersectBatch(rays);
  int nhit = 0, n_second = 0;
  for (size_t i = 0; i < rays.size(); ++i) {
    fledge::Ray r = rays[i];
    fledge::SInteraction isect;
    int32_t mesh = -2, id = -2;
    const bool hit = scene.intersect(r, isect, &mesh, &id);
    CHECK(mesh == mref[i] && id == iref[i] && hits[i].mesh == mref[i] && hits[i].id == iref[i],
          "ray %zu: mesh/id %d/%d, gpu %d/%d, oracle %d/%d", i, mesh, id, hits[i].mesh, hits[i].id, mref[i], iref[i]);
    CHECK(hit == (mref[i] >= 0), "ray %zu: hit flag", i);
    if (!hit) continue;
    ++nhit;
    n_second += mesh == 1;
    CHECK(bits(r.m_tMax) == bits(tref[i]) && bits(hits[i].t) == bits(tref[i]), "ray %zu: t", i);
    CHECK(mesh != 2, "ray %zu: the repeated mesh must lose every tie", i);
  }
  CHECK(nhit > 1000 && n_second > 100, "too few hits (%d, %d on the second mesh)", nhit, n_second);
  std::printf("PASS scene_accel (%d hits of %zu rays)\n", nhit, rays.size());
}

void test_vol_integrator(fledge::gpu::Context& ctx) {
  const int n = 32, W = 40, H = 32, spp = 4, depth = 16;
  std::vector<float> grid(size_t(n) * n * n);
  flg_scene_fbm_grid(n, n, n, 9, 4.0f, 5, 0.0f, 1.0f, grid.data());
  fledge::GridVolume vol(grid.data(), n, n, n, fledge::AABB(Vector3f(0.0f), Vector3f(1.0f)),
                         Vector3f(6.4f), Vector3f(1.6f), 0.3f, &ctx);
  fledge::GPUVolLights lights;
  lights.has_point = true;
  lights.point_pos = Vector3f(2.0f, 2.0f, -2.0f);
  lights.point_intensity = Vector3f(50.0f, 30.0f, 12.5f);
  lights.has_env = true;
  lights.env_radiance = Vector3f(0.5f, 0.35f, 0.8f);
  fledge::GPUVolIntegrator integ(depth, lights);
  fledge::Film film(W, H);
  fledge::Scene scene;
  scene.m_resX = W;
  scene.m_resY = H;
  scene.m_SPP = spp;
  scene.m_maxDepth = depth;
  scene.m_FoV = 45.0f;
  scene.m_origin = Vector3f(0.5f, 0.5f, -1.5f);
  scene.m_target = Vector3f(0.5f, 0.5f, 0.0f);
  scene.m_up = Vector3f(0.0f, 1.0f, 0.0f);
  scene.m_volume = &vol;
  scene.m_film = &film;
  integ.render(scene);
  CHECK(integ.lastLookups() > 0, "lookup count");

  const float bounds[6] = {0, 0, 0, 1, 1, 1};
  const float medium[3] = {6.4f, 1.6f, 0.3f};
  orc::VolScene sc;
  orc::init_vol_grid(sc, grid.data(), n, n, n, bounds, medium);
  sc.point.present = true;
  sc.point.pos = orc::v3(2, 2, -2);
  sc.point.intensity = orc::v3(50.0f, 30.0f, 12.5f);
  sc.env.build(orc::v3(0.5f, 0.35f, 0.8f), sc.grid.bmin, sc.grid.bmax);
  orc::Camera cam(orc::v3(0.5f, 0.5f, -1.5f), orc::v3(0.5f, 0.5f, 0.0f), orc::v3(0, 1, 0), 45.0f);
  sc.cam = &cam;
  sc.width = W;
  sc.height = H;
  sc.spp = spp;
  sc.max_depth = depth;
  // render(const Scene&) is the reference's render loop and film
  std::vector<float> rgb(size_t(W) * H * 3);
  uint64_t lk = 0, paths = 0;
  orc::svol_render_film(sc, rgb.data(), nullptr, nullptr, &lk, &paths);
  double worst = 0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      for (int k = 0; k < 3; ++k) {
        const Vector3f& c = film.getBuffer(x, y, fledge::EFilmBufferType::EColor);
        const float r = rgb[(size_t(y) * W + x) * 3 + k];
        worst = std::max(worst, std::fabs(double(c[k]) - r) / std::max(std::fabs(double(r)), 1e-6));
      }
  CHECK(worst <= 1e-4, "film vs oracle: worst relative error %.3g", worst);
  CHECK(integ.lastLookups() == lk, "lookup counts %llu vs oracle %llu",
        (unsigned long long)integ.lastLookups(), (unsigned long long)lk);
  // update() and setBound() reach the device copy
  std::vector<float> half(grid);
  for (float& v : half) v *= 0.5f;
  vol.update(half.data());
  std::vector<float> out;
  integ.render(vol, scene.m_origin, scene.m_target, scene.m_up, 45.0f, W, H, spp, out);
  bool changed = false;
  for (int i = 0; i < W * H; ++i) changed |= out[3 * i] != film.m_color[i][0];
  CHECK(changed, "update() must change the image");
  std::printf("PASS vol_integrator (film vs oracle worst rel %.2g)\n", worst);
}

// HIPRender in the OptiX stub's place: a volume-only scene (the volume
// integrator's film) and a closed interface mesh holding the medium
// (VolPathIntegrator through the render loop and film, against the oracle's),
// saveImage, and the scenes it refuses.
void test_hip_render(fledge::gpu::Context& ctx) {
  const int n = 24, W = 40, H = 32, spp = 4, depth = 16;
  std::vector<float> grid(size_t(n) * n * n);
  flg_scene_fbm_grid(n, n, n, 5, 4.0f, 5, 0.0f, 1.0f, grid.data());
  fledge::GridVolume vol(grid.data(), n, n, n, fledge::AABB(Vector3f(0.0f), Vector3f(1.0f)), Vector3f(6.4f),
                         Vector3f(1.6f), 0.3f, &ctx);
  fledge::GPUVolLights lights;
  lights.has_point = true;
  lights.point_pos = Vector3f(2.0f, 2.0f, -2.0f);
  lights.point_intensity = Vector3f(50.0f, 30.0f, 12.5f);
  lights.has_env = true;
  lights.env_radiance = Vector3f(0.5f, 0.35f, 0.8f);
  fledge::Film film(W, H);
  fledge::Scene scene;
  scene.m_resX = W;
  scene.m_resY = H;
  scene.m_SPP = spp;
  scene.m_maxDepth = depth;
  scene.m_FoV = 45.0f;
  scene.m_origin = Vector3f(0.5f, 0.5f, -1.5f);
  scene.m_target = Vector3f(0.5f, 0.5f, 0.0f);
  scene.m_up = Vector3f(0.0f, 1.0f, 0.0f);
  scene.m_integrator_type = fledge::EIntegratorType::EVolPathIntegrator;
  scene.m_volume = &vol;
  scene.m_film = &film;
  {  // volume only: the volume integrator's render loop and film
    fledge::HIPRender r(&scene, lights);
    r.init();
    CHECK(r.getBackends() == fledge::kHIPBackend && r.preProcess() && r.render(), "volume-only render");
    fledge::GPUVolIntegrator integ(depth, lights);
    fledge::Film ref(W, H);
    fledge::Scene s2 = scene;
    s2.m_film = &ref;
    integ.render(s2);
    for (int i = 0; i < W * H; ++i)
      for (int k = 0; k < 3; ++k)
        CHECK(bits(film.m_color[i][k]) == bits(ref.m_color[i][k]), "volume-only pixel %d channel %d", i, k);
    CHECK(r.lastLookups() == integ.lastLookups(), "lookup counts");
  }
  // a closed interface mesh around the medium
  int64_t nv = 0;
  const int64_t nt = flg_scene_octasphere_counts(4, &nv);
  std::vector<float> v(size_t(nv) * 3);
  std::vector<int32_t> ind(size_t(nt) * 3);
  flg_scene_octasphere(4, 0.05f, 7, 2.0f, v.data(), ind.data());
  for (float& x : v) x = x * 0.5f + 0.5f;
  std::vector<Vector3f> p((size_t)nv);
  for (int64_t i = 0; i < nv; ++i) p[i] = Vector3f(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  fledge::LBVHAccel mesh(p, std::vector<int>(ind.begin(), ind.end()), ctx);
  scene.m_primitives = {&mesh};
  fledge::HIPRender r(&scene, lights);
  r.init();
  CHECK(r.render(), "mesh + medium render");
  orc::Lbvh b;
  orc::lbvh_build(b, v.data(), ind.data(), (int)nt, (int)nv);
  const float mb[6] = {b.glo.x, b.glo.y, b.glo.z, b.ghi.x, b.ghi.y, b.ghi.z};
  const float medium[3] = {6.4f, 1.6f, 0.3f};
  orc::VolScene sc;
  orc::init_vol_grid(sc, grid.data(), n, n, n, mb, medium);
  sc.point.present = true;
  sc.point.pos = orc::v3(2, 2, -2);
  sc.point.intensity = orc::v3(50.0f, 30.0f, 12.5f);
  sc.env.build(orc::v3(0.5f, 0.35f, 0.8f), sc.grid.bmin, sc.grid.bmax);
  orc::Camera cam(orc::v3(0.5f, 0.5f, -1.5f), orc::v3(0.5f, 0.5f, 0.0f), orc::v3(0, 1, 0), 45.0f);
  sc.cam = &cam;
  sc.width = W;
  sc.height = H;
  sc.spp = spp;
  sc.max_depth = depth;
  std::vector<float> rgb(size_t(W) * H * 3);
  uint64_t lk = 0;
  orc::volpath_mesh_render_film(b, false, sc, rgb.data(), &lk);
  double worst = 0;  // the film sums in another order than the oracle's splats
  for (int i = 0; i < 3 * W * H; ++i)
    worst = std::max(worst, std::fabs(double(film.m_color[i / 3][i % 3]) - rgb[i]) /
                                std::max(std::fabs(double(rgb[i])), 1e-6));
  CHECK(worst <= 1e-5, "mesh + medium film vs oracle: worst relative error %.3g", worst);
  CHECK(r.lastLookups() == lk, "mesh + medium lookup counts");
  const std::string out = "/tmp/fledge_dropin_" + std::to_string((long)getpid()) + "_render.exr";
  CHECK(r.saveImage(out) && !r.saveImage(out, /*denoise=*/true), "saveImage");
  FILE* f = std::fopen(out.c_str(), "rb");
  unsigned char magic[4] = {0, 0, 0, 0};
  CHECK(f && std::fread(magic, 1, 4, f) == 4, "the EXR file");
  std::fclose(f);
  std::remove(out.c_str());
  CHECK(magic[0] == 0x76 && magic[1] == 0x2f && magic[2] == 0x31 && magic[3] == 0x01, "EXR magic");
  // refused scenes
  int refused = 0;
  for (int k = 0; k < 2; ++k) {
    fledge::Scene bad = scene;
    if (k == 0) bad.m_integrator_type = fledge::EIntegratorType::EPathIntegrator;
    if (k == 1) bad.m_primitives = {&mesh, &mesh};
    try {
      fledge::HIPRender rb(&bad, lights);
      rb.init();
    } catch (const fledge::gpu::Error& e) {
      refused += e.status() == FLG_ERR_INVALID;
    }
  }
  CHECK(refused == 2, "HIPRender must refuse the path integrator and several primitives");
  std::printf("PASS hip_render (volume-only and mesh + medium, %llu lookups)\n", (unsigned long long)lk);
}

void test_memory_resources(fledge::gpu::Context& ctx) {
  fledge::gpu_memory_resource gmr;
  void* p = gmr.allocate(1 << 20, 256);
  std::vector<uint8_t> a(1 << 20), b(1 << 20);
  for (size_t i = 0; i < a.size(); ++i) a[i] = uint8_t(i * 131 + 7);
  ctx.check(flg_memcpy_htod(ctx.get(), p, a.data(), a.size()), "htod");
  ctx.check(flg_memcpy_dtoh(ctx.get(), b.data(), p, b.size()), "dtoh");
  gmr.deallocate(p, 1 << 20, 256);
  CHECK(a == b, "gpu_memory_resource round trip");
  CHECK(gmr.is_equal(gmr) && !gmr.is_equal(*std::pmr::new_delete_resource()), "is_equal");
  fledge::async_memory_resource amr(nullptr);
  void* q = amr.allocate(4096, 256);
  CHECK(q != nullptr, "async allocation");
  amr.deallocate(q, 4096, 256);
  ctx.synchronize();
  bool managed = true;
  try {
    void* m = fledge::GlobalManagedMemoryResource()->allocate(4096);
    fledge::GlobalManagedMemoryResource()->deallocate(m, 4096);
  } catch (const std::bad_alloc&) {
    managed = false;
  }
  std::printf("PASS memory_resources (managed memory %s)\n", managed ? "available" : "unavailable");
}

}  // namespace

int main(int argc, char** argv) {
  const bool cpu_only = argc > 1 && std::string(argv[1]) == "--cpu";
  test_grid_volume_host();
  if (!g_failures) test_sparse_volume_host();
  test_host_errors();
  if (!g_failures) test_ply_host();
  if (!g_failures) test_scene_file_host();
  if (!cpu_only && g_failures == 0) {
    try {
      fledge::gpu::Context ctx(0, nullptr);
      std::printf("context on device 0\n");
      std::fflush(stdout);
      if (!g_failures) test_memory_resources(ctx);
      if (!g_failures) test_lbvh_builder(ctx);
      if (!g_failures) test_lbvh_accel(ctx);
      if (!g_failures) test_scene_accel(ctx);
      if (!g_failures) test_vol_integrator(ctx);
      if (!g_failures) test_hip_render(ctx);
      ctx.synchronize();
    } catch (const std::exception& e) {
      std::printf("FAIL exception: %s\n", e.what());
      ++g_failures;
    }
  }
  std::printf("%s\n", g_failures ? "FAILED" : "ALL PASSED");
  return g_failures ? 1 : 0;
}
