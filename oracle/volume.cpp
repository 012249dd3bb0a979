// TEST INFRASTRUCTURE — CPU restatement of fledge's volume-only path
// integrator: SVolIntegrator::Li (src/integrator.cpp:328-405) with delta
// tracking and ratio tracking (src/volume.cpp:75-148), NEE
// (src/integrator.cpp:41-120), Henyey-Greenstein phase
// (src/common/math_utils.h:103-128), the constant-texture environment light
// (src/light.cpp:93-159, src/distribution.hpp) and a point light.
//
// Declared deviations ("fixed" semantics, the reference path is unreachable
// as shipped):
//  * the density lookup is a dense grid with trilinear interpolation defined
//    below (the reference uses OpenVDB's BoxSampler, whose arithmetic is not
//    in the tree);
//  * the scattered VInteraction carries its ray and medium, so shadow rays
//    are attenuated by Volume::tr (the reference loses m_ray);
//  * one of n lights is picked with probability 1/n (the reference asserts
//    n == 1); lights are ordered [point, environment];
//  * PointLight (absent in the reference): Li = I / |p - x|^2, pdf 1;
//  * per-sample schedule: sample s of pixel p uses Halton index
//    offset(p) + s * stride, dimensions 0,1 for the pixel position and 2.. for
//    Li; the pixel value is the float sum over s in order, divided by spp.
#include <atomic>
#include <cmath>
#include <thread>

#include "oracle_impl.hpp"

namespace orc {

// Free-flight walks.  Halton index 0 (pixel offset 0, sample 0) gives 0 in
// every dimension >= 2, so t never advances and the reference's loops
// (volume.cpp:86, :123) never end.  A walk ends at a zero draw (only index 0
// gives one) and in any case after kMaxTrackSteps steps: ratio tracking
// returns the current transmittance, delta tracking reports no collision.
static constexpr int kMaxTrackSteps = 4096;

// Trilinear lookup on the dense grid: index space g = (p - bmin) * scale - 0.5,
// zero outside the grid, lerps along x, then y, then z.
Lookup grid_lookup(const Grid& g, V3 p) {
  float gx = (p.x - g.bmin.x) * g.scale.x - 0.5f;
  float gy = (p.y - g.bmin.y) * g.scale.y - 0.5f;
  float gz = (p.z - g.bmin.z) * g.scale.z - 0.5f;
  float flx = std::floor(gx), fly = std::floor(gy), flz = std::floor(gz);
  Lookup L;
  L.fx = gx - flx;
  L.fy = gy - fly;
  L.fz = gz - flz;
  // clamp far-outside positions; anything outside [-1, n] reads zero anyway
  float cx = smin(smax(flx, -2.0f), (float)g.nx + 1.0f);
  float cy = smin(smax(fly, -2.0f), (float)g.ny + 1.0f);
  float cz = smin(smax(flz, -2.0f), (float)g.nz + 1.0f);
  L.i = (int)cx;
  L.j = (int)cy;
  L.k = (int)cz;
  float wx0 = 1.0f - L.fx, wx1 = L.fx;
  float wy0 = 1.0f - L.fy, wy1 = L.fy;
  float wz0 = 1.0f - L.fz, wz1 = L.fz;
  int i = L.i, j = L.j, k = L.k;
  float c00 = g.voxel(i, j, k) * wx0 + g.voxel(i + 1, j, k) * wx1;
  float c10 = g.voxel(i, j + 1, k) * wx0 + g.voxel(i + 1, j + 1, k) * wx1;
  float c01 = g.voxel(i, j, k + 1) * wx0 + g.voxel(i + 1, j, k + 1) * wx1;
  float c11 = g.voxel(i, j + 1, k + 1) * wx0 + g.voxel(i + 1, j + 1, k + 1) * wx1;
  float c0 = c00 * wy0 + c10 * wy1;
  float c1 = c01 * wy0 + c11 * wy1;
  L.rho = c0 * wz0 + c1 * wz1;
  return L;
}

// AABB::intersect_pbrt aabb.h:38-59
bool box_intersect_pbrt(V3 lo, V3 hi, V3 o, V3 d, float tmax, float* t_min, float* t_max) {
  float t0 = 0.0f, t1 = tmax;
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / comp(d, i);
    float tn = (comp(lo, i) - comp(o, i)) * inv;
    float tf = (comp(hi, i) - comp(o, i)) * inv;
    if (tn > tf) { float t = tn; tn = tf; tf = t; }
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  *t_min = t0;
  *t_max = t1;
  return true;
}

// AABB::intersect aabb.h:20-36: inverse via the double literal 1.0
// (cwiseInverse, vector.h:248-250 and operator/(T_, Vector) :291-294).
bool box_intersect(V3 lo, V3 hi, V3 o, V3 d, float* t_min, float* t_max) {
  V3 inv = v3((float)(1.0 / (double)d.x), (float)(1.0 / (double)d.y), (float)(1.0 / (double)d.z));
  V3 t1 = (lo - o) * inv;
  V3 t2 = (hi - o) * inv;
  V3 vmn = vmin(t1, t2), vmx = vmax(t1, t2);
  float lmin = smax(smax(vmn.x, vmn.y), vmn.z);  // maxCoeff (reduce with max)
  float lmax = smin(smin(vmx.x, vmx.y), vmx.z);  // minCoeff
  if (lmin > lmax) { float t = lmin; lmin = lmax; lmax = t; }
  if (lmin < lmax && lmax >= 0) {
    *t_min = lmin;
    *t_max = lmax;
    return true;
  }
  return false;
}

// OpenVDBVolume::tr volume.cpp:75-107 (ratio tracking with Russian roulette)
static float vol_tr(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, uint64_t* lookups) {
  float t_min, t_max;
  if (!box_intersect_pbrt(sc.grid.bmin, sc.grid.bmax, o, d, ray_tmax, &t_min, &t_max)) return 1.0f;
  float tr = 1.0f, t = smax(t_min, 0.0f);
  t_max = smin(t_max, ray_tmax);
  float max_mu = sc.med.max_mu;
  if (!(max_mu > 0.0f)) return 1.0f;  // empty medium (the reference divides by 0)
  for (int step = 0; step < kMaxTrackSteps; ++step) {
    const float u = hs.get1D();
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / max_mu;
    if (t >= t_max) break;
    V3 pos = o + d * t;
    float density = grid_lookup(sc.grid, pos).rho;
    if (lookups) ++*lookups;
    tr *= 1.0f - smax(0.0f, density * sc.med.inv_max_density);
    const float rr = 0.1f;
    if (tr < rr) {
      float q = smax(0.05f, 1.0f - tr);
      if (hs.get1D() < q) return 0.0f;
      tr /= 1.0f - q;
    }
  }
  return tr;
}

// OpenVDBVolume::sample volume.cpp:109-148 (delta tracking)
static float vol_sample(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, bool* success,
                        V3* pos_out, uint64_t* lookups) {
  float t_min, t_max;
  if (!box_intersect_pbrt(sc.grid.bmin, sc.grid.bmax, o, d, ray_tmax, &t_min, &t_max)) {
    *success = false;
    return 1.0f;
  }
  float t = smax(t_min, 0.0f);
  float max_mu = sc.med.max_mu;
  t_max = smin(t_max, ray_tmax);
  if (!(max_mu > 0.0f)) {
    *success = false;
    return 1.0f;
  }
  for (int step = 0; step < kMaxTrackSteps; ++step) {
    const float u = hs.get1D();
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / max_mu;
    if (t >= t_max) break;
    V3 pos = o + d * t;
    float density = grid_lookup(sc.grid, pos).rho;
    if (lookups) ++*lookups;
    float p_s = sc.med.sigma_s * density / max_mu;
    float p_a = sc.med.sigma_a * density / max_mu;
    float p_n = 1.0f - p_s - p_a;
    if (hs.get1D() > p_n) {
      *success = true;
      *pos_out = pos;
      return sc.med.sigma_s / sc.med.sigma_t;
    }
  }
  *success = false;
  return 1.0f;
}

float vol_tr_lk(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, uint64_t* lookups) {
  return vol_tr(sc, hs, o, d, ray_tmax, lookups);
}

float vol_sample_lk(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, bool* success, V3* pos,
                    uint64_t* lookups) {
  return vol_sample(sc, hs, o, d, ray_tmax, success, pos, lookups);
}

float vol_tr_pub(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax) {
  uint64_t lk = 0;
  return vol_tr(sc, hs, o, d, ray_tmax, &lk);
}

float vol_sample_pub(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, bool* success,
                     V3* pos) {
  uint64_t lk = 0;
  return vol_sample(sc, hs, o, d, ray_tmax, success, pos, &lk);
}

// HGP math_utils.h:103-107
float hg_p(V3 wi, V3 wo, float g) {
  float cos_theta = dot(wi, wo);
  float denom = 1.0f + g * g + 2.0f * g * cos_theta;
  return (1.0f - g * g) / (denom * std::sqrt(denom)) * INV_4PI;
}

// CoordinateSystem math_utils.h:57-66
void coordinate_system(V3 v1, V3* v2, V3* v3o) {
  if (std::fabs(v1.x) > std::fabs(v1.y))
    *v2 = v3(-v1.z, 0.0f, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
  else
    *v2 = v3(0.0f, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
  *v3o = cross(v1, *v2);
}

// HGSampleP math_utils.h:112-128 (u = first draw, v = second: clang's
// left-to-right argument evaluation at integrator.cpp:378)
V3 hg_sample(V3 wo, float u, float v, float g) {
  float cos_theta;
  if ((double)std::fabs(g) < 1e-3) {
    cos_theta = 1.0f - 2.0f * u;
  } else {
    float tmp = (1.0f - g * g) / (1.0f + g - 2.0f * g * u);
    cos_theta = -(1.0f + g * g - tmp * tmp) / (2.0f * g);
  }
  float phi = 2.0f * PI * v;
  // sqrt(fmax(0, x)) promotes to double; the float rounding of the double
  // square root of a float equals sqrtf.
  float x = 1.0f - cos_theta * cos_theta;
  float sin_theta = (float)std::sqrt(std::fmax(0.0, (double)x));
  V3 v1, v2;
  coordinate_system(wo, &v1, &v2);
  float sp, cp;
  flg_sincosf(phi, &sp, &cp);
  return (v1 * (sin_theta * cp) + v2 * (sin_theta * sp)) + wo * cos_theta;
}

V3 hg_sample_pub(V3 wo, float u, float v, float g) { return hg_sample(wo, u, v, g); }
float hg_p_pub(V3 wi, V3 wo, float g) { return hg_p(wi, wo, g); }

// Dist1D distribution.hpp:14-22, 26-50
void Dist1D::init(const float* fv, int n) {
  f.assign(fv, fv + n);
  cdf.assign(n + 1, 0.0f);
  cdf[0] = 0.0f;
  for (int i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / (float)n;
  integral = cdf[n];
  if (integral == 0.0f)
    for (int i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
  else
    for (int i = 1; i < n + 1; ++i) cdf[i] /= integral;
}
int Dist1D::inverse_cdf(float u) const {  // note r = mid - 1 (reference quirk)
  int l = 0, r = (int)f.size();
  while (r - l > 1) {
    int mid = (l + r) / 2;
    if (cdf[mid] <= u) l = mid;
    else r = mid - 1;
  }
  return l;
}
float Dist1D::sample_c(float u, float* pdf, int* off) const {
  int idx = inverse_cdf(u);
  *off = idx;
  float du = u - cdf[idx];
  if ((cdf[idx + 1] - cdf[idx]) > 0) du /= (cdf[idx + 1] - cdf[idx]);
  *pdf = f[idx] / integral;
  return ((float)idx + du) / (float)f.size();
}

// InfiniteAreaLight(texture) light.cpp:93-108 with ConstTexture (texture.cpp:52-55)
// and preprocess light.cpp:110-117 (scene bound = the volume box).
void EnvLight::build(V3 rad, V3 lo, V3 hi) {
  present = true;
  radiance = rad;
  const int NU = 512, NV = 512;
  std::vector<float> fv(NU * NV, 0.0f);
  float nrm = norm(rad);  // m_tex->eval(u, v).norm()
  for (int v = 0; v < NV; ++v) {
    float sin_theta = std::sin(PI * (float)(v + 0.5f) / (float)NV);
    for (int u = 0; u < NU; ++u) {
      fv[v * NU + u] = nrm;
      fv[v * NU + u] *= sin_theta;
    }
  }
  rows.resize(NV);
  std::vector<float> marg_f(NV);
  for (int v = 0; v < NV; ++v) {
    rows[v].init(&fv[v * NU], NU);
    marg_f[v] = rows[v].integral;
  }
  marg.init(marg_f.data(), NV);
  center = (lo + hi) / 2.0f;  // AABB::boundSphere aabb.h:67-70
  radius = norm(hi - center);
  radius *= 1000.0f;
  radius = smin(FLT_MAX / 2, radius);
}

// Light::sampleLi for the environment, light.cpp:119-145
V3 env_sample_li(const EnvLight& env, V3 ref, float u0, float u1, V3* wi, float* pdf,
                           V3* p_light) {
  float pdfs[2];
  int offs[2];
  float d1 = env.marg.sample_c(u1, &pdfs[1], &offs[1]);
  float d0 = env.rows[offs[1]].sample_c(u0, &pdfs[0], &offs[0]);
  *pdf = pdfs[0] * pdfs[1];
  float phi = d0 * 2.0f * PI;
  float theta = d1 * PI;
  float st, ct, sp, cp;
  flg_sincosf(theta, &st, &ct);
  flg_sincosf(phi, &sp, &cp);
  V3 dir = v3(st * cp, st * sp, ct);  // SphericalDirection math_utils.h:15-18
  *pdf /= (2.0f * PI * PI * smax(st, 1e-4f));
  dir = v3(-dir.y, dir.z, -dir.x);  // LightToWorld light.cpp:33-35
  V3 p = env.center + (dir * 2.0f) * env.radius;
  *p_light = p;
  *wi = normalize(p - ref);
  return env.radiance;  // Le(ray) of a constant texture
}

// EstimateDirect integrator.cpp:69-107 for a volume vertex, VolEstimateTr
// reduced to the medium's transmittance (no surfaces in the scene);
// UniformSampleOneLight integrator.cpp:109-120.  Vector3f arithmetic
// (vector.inc): products and quotients component by component.
static V3 sample_one_light(const VolScene& sc, Halton& hs, V3 p, V3 wo, uint64_t* lookups) {
  int n_lights = (sc.point.present ? 1 : 0) + (sc.env.present ? 1 : 0);
  if (n_lights == 0) return v3(0, 0, 0);
  int light_num = (int)(hs.get1D() * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  float light_pdf = (float)(1.0 / (double)n_lights);
  bool use_point = sc.point.present && light_num == 0;
  // EstimateDirect: u_light drawn before sampleLi
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
    // Interaction::SpawnRayTo interaction.hpp:33-39 (ng = 0: no offset)
    V3 diff = p_light - p;
    float nrm = norm(diff);
    V3 d = diff / nrm;
    float tr = vol_tr(sc, hs, p, d, nrm - SHADOW_EPS, lookups);
    Li = Li * tr;
    L = L + (Li * f) / pdf;
  }
  return L / light_pdf;
}

// SVolIntegrator::Li integrator.cpp:328-405 (m_rrThreshold = 0.01)
V3 svol_Li(const VolScene& sc, Halton& hs, V3 o, V3 d, uint64_t* lookups) {
  V3 L = v3(0, 0, 0);
  float beta = 1.0f;
  bool in_volume = false;
  const float rr = 0.01f;
  for (int bounces = 0;; ++bounces) {
    if (bounces == 0) {
      float t_min, t_max;
      if (box_intersect(sc.grid.bmin, sc.grid.bmax, o, d, &t_min, &t_max)) {
        float tr = vol_tr(sc, hs, o, d, INFINITY, lookups);
        if (sc.env.present) L = L + sc.env.radiance * tr;
        in_volume = true;
        o = o + d * t_min;
        continue;
      } else {
        if (sc.env.present) L = L + sc.env.radiance * beta;
        break;
      }
    }
    if (in_volume) {
      bool success = false;
      V3 pos;
      float f = vol_sample(sc, hs, o, d, INFINITY, &success, &pos, lookups);
      if (success) {
        beta = beta * f;
        V3 wo = -d;
        float u = hs.get1D();
        float v = hs.get1D();
        V3 wi = hg_sample(wo, u, v, sc.med.g);
        L = L + sample_one_light(sc, hs, pos, wo, lookups) * beta;
        // beta.isZero(): SquaredNorm == 0 (vector.h:268-270)
        if (bounces >= sc.max_depth || beta * beta + beta * beta + beta * beta == 0.0f) break;
        o = pos;  // VInteraction::SpawnRay: ng = 0, origin unchanged
        d = normalize(wi);
      } else {
        break;
      }
    }
    if (bounces > 10) {
      if (hs.get1D() < rr) break;
      beta /= (1.0f - rr);
    }
  }
  return L;
}

void svol_render(const VolScene& sc, float* rgb, int nthreads, uint64_t* lookups_out,
                 uint64_t* paths_out) {
  const int W = sc.width, H = sc.height;
  const int64_t npix = (int64_t)W * H;
  if (nthreads < 1) nthreads = 1;
  std::vector<uint64_t> lk(nthreads, 0);
  std::vector<std::thread> th;
  std::atomic<int64_t> next{0};
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t] {
      Halton hs((uint64_t)sc.spp, W, H);
      while (true) {
        int64_t pix = next.fetch_add(1) * sc.pix_stride;
        if (pix >= npix) break;
        int x = (int)(pix % W), y = (int)(pix / W);
        V3 color = v3(0, 0, 0);
        for (int s = 0; s < sc.spp; ++s) {
          hs.start(x, y, (uint64_t)s);
          float ux = hs.get1D();
          float uy = hs.get1D();
          V3 o, d;
          sc.cam->generate((float)x + ux, (float)y + uy, W, H, &o, &d);
          color = color + svol_Li(sc, hs, o, d, &lk[t]);
        }
        color = color / (float)sc.spp;
        rgb[3 * pix] = color.x;
        rgb[3 * pix + 1] = color.y;
        rgb[3 * pix + 2] = color.z;
      }
    });
  }
  for (auto& x : th) x.join();
  uint64_t tot = 0;
  for (auto v : lk) tot += v;
  if (lookups_out) *lookups_out = tot;
  if (paths_out) *paths_out = (uint64_t)((npix + sc.pix_stride - 1) / sc.pix_stride) * (uint64_t)sc.spp;
}

}  // namespace orc
