// TEST INFRASTRUCTURE — CPU restatement of the segment sequence of fledge's
// PathIntegrator::Li (src/integrator.cpp:252-326) over a diffuse mesh with no
// lights, traced by the LBVH (radix_bvh.cpp:202-245): the incoming-ray
// workload of the incoherent closest-hit benchmark.
//
// Per segment: closest hit; stop on a miss or at max depth; else the diffuse
// sampleF in the shading frame (CoordinateTransition material_base.hpp:14-41,
// MaterialDispatcher::sampleF material_base.cpp:34-46,
// DiffuseMaterial::sampleF_impl builtin_materials.cpp:44-50,
// CosineSampleHemisphere / ConcentricSampleDisk math_utils.h:69-88) with two
// Halton dimensions; stop if pdf == 0; spawn with OffsetRayOrigin
// (ray.h:64-93).  With no lights UniformSampleOneLight consumes no dimension.
// The hit interaction of the LBVH accel: p = o + d t, ng = ns = the Pluecker
// normal.
#include <atomic>
#include <cmath>
#include <thread>

#include "oracle_impl.hpp"

namespace orc {

// OffsetRayOrigin ray.h:64-93
V3 offset_ray_origin(V3 p, V3 n, V3 dir) {
  const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
  V3 offset = n;
  if (dot(n, dir) <= 0) offset = -n;
  int oi[3] = {(int)(offset.x * int_scale), (int)(offset.y * int_scale), (int)(offset.z * int_scale)};
  float pc[3] = {p.x, p.y, p.z}, oc[3] = {offset.x, offset.y, offset.z}, r[3];
  for (int i = 0; i < 3; ++i) {
    int32_t bits = (int32_t)f2u(pc[i]);
    float pint = u2f((uint32_t)(bits + ((pc[i] < 0) ? -oi[i] : oi[i])));
    r[i] = std::fabs(pc[i]) < origin ? pc[i] + float_scale * oc[i] : pint;
  }
  return v3(r[0], r[1], r[2]);
}

struct Frame {  // CoordinateTransition material_base.hpp:14-41
  V3 n, t, b;
  explicit Frame(V3 normal) {
    n = normal / std::sqrt(sqnorm(normal));
    if (std::fabs(n.x) > std::fabs(n.z)) b = v3(-n.y, n.x, 0.0f);
    else b = v3(0.0f, -n.z, n.y);
    b = b / std::sqrt(sqnorm(b));
    t = cross(b, n);
  }
  V3 to_local(V3 p) const { return v3(dot(p, t), dot(p, b), dot(p, n)); }
  V3 to_world(V3 p) const { return (t * p.x + b * p.y) + n * p.z; }
};

// ConcentricSampleDisk / CosineSampleHemisphere math_utils.h:69-88
static V3 cosine_hemisphere(float u0, float u1) {
  float ux = (float)((double)u0 * 2.0 - 1.0), uy = (float)((double)u1 * 2.0 - 1.0);
  float dx, dy;
  if (ux == 0 && uy == 0) {
    dx = 0.0f;
    dy = 0.0f;
  } else {
    float theta, r;
    if (std::fabs(ux) > std::fabs(uy)) {
      r = ux;
      theta = PI_OVER4 * (uy / ux);
    } else {
      r = uy;
      theta = PI_OVER2 - (PI_OVER4 * (ux / uy));
    }
    float s, c;
    flg_sincosf(theta, &s, &c);
    dx = c * r;
    dy = s * r;
  }
  float z = std::sqrt(smax(0.0f, 1.0f - dx * dx - dy * dy));
  return v3(dx, dy, z);
}

// The diffuse bounce at a hit with geometric normal ng seen along -wo: the
// shading frame of ng, a cosine-weighted local direction from (u0, u1), the
// pdf (DiffuseMaterial::pdf_impl builtin_materials.cpp:38-41: zero unless wo
// and wi lie on the same side of ng) and wi back in world space.
void diffuse_sample(V3 ng, V3 wo, float u0, float u1, V3* wi, float* pdf) {
  Frame fr(ng);
  V3 wo_l = fr.to_local(wo);
  V3 wi_l = cosine_hemisphere(u0, u1);
  *pdf = wo_l.z * wi_l.z > 0 ? std::fabs(wi_l.z) * INV_PI : 0.0f;
  *wi = fr.to_world(wi_l);
}

// The shading frame's axes (tangent, binormal, normal) for a normal.
void shading_frame(V3 n, V3* t, V3* b, V3* nn) {
  Frame fr(n);
  *t = fr.t;
  *b = fr.b;
  *nn = fr.n;
}

// One path: returns the number of segments; writes (id, t) per segment.
int diffuse_path(const Lbvh& b, bool compat, Halton& hs, V3 o, V3 d, int max_depth, int32_t* ids,
                 float* ts, TraceCounters* ctr) {
  int nseg = 0;
  for (int bounces = 0;; ++bounces) {
    float t;
    V3 ng;
    int32_t id = lbvh_trace(b, o, d, INFINITY, compat, &t, &ng, ctr);
    ids[nseg] = id;
    ts[nseg] = t;
    ++nseg;
    if (id < 0 || bounces >= max_depth) break;
    V3 p = o + d * t;
    float u0 = hs.get1D();
    float u1 = hs.get1D();
    V3 wi;
    float pdf;
    diffuse_sample(ng, -d, u0, u1, &wi, &pdf);
    if (pdf == 0.0f) break;
    o = offset_ray_origin(p, ng, wi);
    d = normalize(wi);
  }
  return nseg;
}

// Paths [p0, p1) of a W x H x spp image, path index = pixel * spp + s.
void diffuse_render(const Lbvh& b, bool compat, const Camera& cam, int W, int H, int spp,
                    int max_depth, int64_t p0, int64_t p1, int32_t* nseg, int32_t* ids, float* ts,
                    int nthreads, uint64_t* rays_out) {
  if (nthreads < 1) nthreads = 1;
  std::atomic<int64_t> next{p0};
  std::vector<uint64_t> rays(nthreads, 0);
  std::vector<std::thread> th;
  const int S = max_depth + 1;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t] {
      Halton hs((uint64_t)spp, W, H);
      int32_t lid[64];
      float lt[64];
      while (true) {
        int64_t base = next.fetch_add(64);
        if (base >= p1) break;
        int64_t end = base + 64 < p1 ? base + 64 : p1;
        for (int64_t p = base; p < end; ++p) {
          int64_t pix = p / spp;
          int s = (int)(p % spp);
          int x = (int)(pix % W), y = (int)(pix / W);
          hs.start(x, y, (uint64_t)s);
          float ux = hs.get1D();
          float uy = hs.get1D();
          V3 o, d;
          cam.generate((float)x + ux, (float)y + uy, W, H, &o, &d);
          int32_t* pid = ids ? ids + (p - p0) * S : lid;
          float* pt = ts ? ts + (p - p0) * S : lt;
          int n = diffuse_path(b, compat, hs, o, d, max_depth, pid, pt, nullptr);
          if (nseg) nseg[p - p0] = n;
          rays[t] += (uint64_t)n;
        }
      }
    });
  }
  for (auto& x : th) x.join();
  uint64_t tot = 0;
  for (auto v : rays) tot += v;
  if (rays_out) *rays_out = tot;
}

}  // namespace orc
