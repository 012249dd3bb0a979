// TEST INFRASTRUCTURE — CPU restatement of VolPathIntegrator::Li
// (src/integrator.cpp:407-510) for a scene of one closed triangle mesh whose
// triangles are index-matched interfaces (no material) with the density volume
// inside, and of the pieces it calls: VolEstimateTr (integrator.cpp:40-64),
// EstimateDirect / UniformSampleOneLight for a volume vertex
// (integrator.cpp:66-120), SInteraction::SpawnRay / SpawnRayTo and
// VInteraction::SpawnRay (src/interaction.cpp:16-57, Entering from
// src/common/math_utils.h:149-151), OpenVDBVolume::sample / tr (the oracle's
// trackers, volume.cpp).  The scene's accel is the LBVH of the mesh (the
// oracle's build and DFS).
//
// Where the reference has no behaviour to follow: it dereferences the missing
// material of an interface (integrator.cpp:433,487,501).  Here the path
// crosses an interface as VolEstimateTr's shadow rays do (SpawnRay in the
// same direction, weight unchanged) and the crossing is not counted as a
// bounce, so a camera ray that leaves the medium unscattered still sees the
// environment at bounce 0.  At most kMaxCrossings crossings per path or
// shadow ray.  The medium fills the mesh's bounding box (ShapePrimitive's
// ctor, src/primitive.cpp:17-24: m_volume->setBound(getBound())).
#include <atomic>
#include <thread>
#include <vector>

#include "oracle_impl.hpp"

namespace orc {

namespace {

constexpr int kMaxCrossings = 256;

// VolEstimateTr: from the volume vertex p toward the light, up to tmax.
float shadow_tr(const Lbvh& b, bool compat, const VolScene& sc, Halton& hs, V3 o, V3 d, float tmax, bool inside,
                uint64_t* lookups) {
  float tr = 1.0f;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t = 0;
    V3 ng{};
    if (lbvh_trace(b, o, d, t_max, compat, &t, &ng, nullptr) < 0) return tr;  // no hit: done
    // an interface has no material: it does not occlude
    if (inside) tr *= vol_tr_lk(sc, hs, o, d, t, lookups);  // ray.m_volume->tr over [0, tHit]
    t_max -= t;
    if (t_max < 0) return tr;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);  // SInteraction::SpawnRay(ray.m_d)
    d = normalize(d);
    inside = dot(d, ng) < 0;  // Entering(d, ng): the mesh's medium, else none
  }
  return tr;
}

// UniformSampleOneLight + EstimateDirect at a volume vertex (Vector3f
// arithmetic component by component).
V3 one_light(const Lbvh& b, bool compat, const VolScene& sc, Halton& hs, V3 p, V3 wo, uint64_t* lookups) {
  int n_lights = (sc.point.present ? 1 : 0) + (sc.env.present ? 1 : 0);
  if (n_lights == 0) return v3(0, 0, 0);
  int light_num = (int)(hs.get1D() * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  float light_pdf = (float)(1.0 / (double)n_lights);
  bool use_point = sc.point.present && light_num == 0;
  float u0 = hs.get1D();
  float u1 = hs.get1D();
  V3 wi, p_light, Li;
  float pdf;
  if (use_point) {
    V3 diff = sc.point.pos - p;
    float d2 = sqnorm(diff);
    wi = normalize(diff);
    Li = sc.point.intensity / d2;
    pdf = 1.0f;
    p_light = sc.point.pos;
  } else {
    Li = env_sample_li(sc.env, p, u0, u1, &wi, &pdf, &p_light);
  }
  V3 L = v3(0, 0, 0);
  if (pdf == 0.0f || (Li.x == 0.0f && Li.y == 0.0f && Li.z == 0.0f)) return L / light_pdf;
  float f = hg_p(wi, wo, sc.med.g);
  if (f != 0.0f) {
    // VInteraction::SpawnRayTo: no normal, so no offset; tMax = norm - SHADOW_EPS
    V3 diff = p_light - p;
    float nrm = norm(diff);
    V3 d = diff / nrm;
    float tr = shadow_tr(b, compat, sc, hs, p, d, nrm - SHADOW_EPS, true, lookups);
    Li = Li * tr;
    L = L + (Li * f) / pdf;
  }
  return L / light_pdf;
}

// VolPathIntegrator::Li for one camera ray.
V3 volpath_Li(const Lbvh& b, bool compat, const VolScene& sc, Halton& hs, V3 o, V3 d, uint64_t* lookups) {
  V3 L = v3(0, 0, 0);
  float beta = 1.0f;
  bool inside = false;  // the camera is outside the mesh
  int bounces = 0, crossings = 0;
  while (true) {
    float t = 0;
    V3 ng{};
    const bool hit = lbvh_trace(b, o, d, INFINITY, compat, &t, &ng, nullptr) >= 0;
    if (bounces == 0 && !hit && sc.env.present) L = L + sc.env.radiance * beta;  // Le of the environment
    if (!hit || bounces >= sc.max_depth) break;
    bool success = false;
    V3 pos{};
    if (inside) beta *= vol_sample_lk(sc, hs, o, d, t, &success, &pos, lookups);
    if (success) {
      const V3 wo = -d;
      L = L + one_light(b, compat, sc, hs, pos, wo, lookups) * beta;
      const float u = hs.get1D();
      const float v = hs.get1D();
      const V3 wi = hg_sample(wo, u, v, sc.med.g);
      if (bounces >= sc.max_depth || beta * beta + beta * beta + beta * beta == 0.0f) break;
      o = pos;  // VInteraction::SpawnRay: no normal, no offset
      d = normalize(wi);
      ++bounces;
      continue;
    }
    if (++crossings > kMaxCrossings) break;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);
    d = normalize(d);
    inside = dot(d, ng) < 0;
  }
  return L;
}

}  // namespace

// Per pixel the in-order mean of spp samples on the per-sample schedule;
// pixels 0, pix_stride, 2 pix_stride, ... only (the others keep rgb's values).
void volpath_mesh_render(const Lbvh& b, bool compat, const VolScene& sc, float* rgb, int nthreads,
                         uint64_t* lookups) {
  const int W = sc.width, H = sc.height;
  const int64_t npix = (int64_t)W * H;
  if (nthreads < 1) nthreads = 1;
  std::vector<uint64_t> lk(nthreads, 0);
  std::vector<std::thread> th;
  std::atomic<int64_t> next{0};
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      Halton hs((uint64_t)sc.spp, W, H);
      while (true) {
        const int64_t pix = next.fetch_add(1) * sc.pix_stride;
        if (pix >= npix) break;
        const int x = (int)(pix % W), y = (int)(pix / W);
        V3 color = v3(0, 0, 0);
        for (int s = 0; s < sc.spp; ++s) {
          hs.start(x, y, (uint64_t)s);
          const float ux = hs.get1D();
          const float uy = hs.get1D();
          V3 o, d;
          sc.cam->generate((float)x + ux, (float)y + uy, W, H, &o, &d);
          color = color + volpath_Li(b, compat, sc, hs, o, d, &lk[t]);
        }
        color = color / (float)sc.spp;
        rgb[3 * pix] = color.x;
        rgb[3 * pix + 1] = color.y;
        rgb[3 * pix + 2] = color.z;
      }
    });
  for (auto& x : th) x.join();
  if (lookups) {
    uint64_t tot = 0;
    for (auto v : lk) tot += v;
    *lookups = tot;
  }
}

void volpath_mesh_render_film(const Lbvh& b, bool compat, const VolScene& sc, float* rgb, uint64_t* lookups) {
  film_render(
      sc, [&](Halton& hs, V3 o, V3 d, uint64_t* lk) { return volpath_Li(b, compat, sc, hs, o, d, lk); }, rgb,
      nullptr, nullptr, lookups, nullptr);
}

}  // namespace orc
