// TEST INFRASTRUCTURE — FP64 twin of the differentiable volume estimator and
// its analytic gradient (no counterpart in the reference; the closest
// reference code is the ratio tracking of OpenVDBVolume::tr,
// src/volume.cpp:75-107, and the delta tracking of OpenVDBVolume::sample,
// src/volume.cpp:109-148, inside SVolIntegrator::Li,
// src/integrator.cpp:328-405).
//
// Estimator ("differential delta tracking", Nimier-David et al. 2022, with
// path replay, Vicini et al. 2021): the same path structure as
// SVolIntegrator::Li, the main path drawing the Halton dimensions in its
// order, with
//  * a fixed majorant mu (caller-given);
//  * collision decisions taken on a proposal grid theta_p: real iff
//    u > 1 - s rho_p / mu (float, as the GPU kernel), weight
//    alpha * rho / rho_p on a real collision and
//    (1 - s rho / mu) / (1 - s rho_p / mu) on a null one, so the estimator is
//    a smooth function of the evaluated grid theta (no branch depends on it);
//  * ratio tracking without Russian roulette on the camera and shadow rays,
//    factor 1 - s rho / mu per tentative collision (no clamp at 0), each walk
//    drawing from its own counter-based stream (walk_draw; walk 0 is the
//    camera ray, walk b the shadow ray of bounce b), walked to its end here;
//  * the path-level roulette after 10 bounces kept (it depends on u only).
// Positions, free-flight distances and decisions are computed in float
// exactly like the GPU kernel; densities of theta, weights, radiance and
// gradients are computed in double.
//
// The same estimator over an interface mesh (diff_path_mesh) follows
// VolPathIntegrator::Li (src/integrator.cpp:407-510) as the oracle's
// volpath_mesh.cpp restates it: closest hit on the mesh, the environment at
// bounce 0 on escape, differential delta tracking inside the medium up to the
// hit, the light sample before the phase sample, no roulette; each shadow
// ray's ratio track runs through the interfaces (VolEstimateTr,
// integrator.cpp:40-64) on walk stream b (the shadow ray of bounce b >= 1),
// its draws continuing from one inside segment to the next.
#include <atomic>
#include <cmath>
#include <thread>

#include "oracle_impl.hpp"

namespace orc {

namespace {

constexpr int kMaxTrackSteps = 4096;

// Steps a walk may take over [t, t_max]: 8 times the expected count
// (mu * length) plus 64, at most kMaxTrackSteps (the estimator's rule for
// the minute free flights of the first Halton indices; it bounds every walk).
int walk_cap(float mu, float t, float t_max) {
  const float e = 8.0f * mu * (t_max - t) + 64.0f;
  return e < (float)kMaxTrackSteps ? (int)e : kMaxTrackSteps;
}

struct DiffScene {
  Grid gp;                 // proposal grid (float), also gives lookup positions/weights
  const double* theta;     // evaluated grid
  float s_f, alpha_f, g_f, mu_f;
  double s, alpha, mu;
  EnvLight env;
  PointLight point;
  int ch = 0;  // the colour channel whose radiance this run computes
  Camera* cam;
  int W, H, spp, max_depth;
  int64_t pix_stride = 1;  // pixels 0, stride, 2 stride, ... (1: all)
};

struct Tri {  // lookup record: base voxel, weights, densities
  int i, j, k;
  double w[8];
  float rho_p;
  double rho;
};

Tri lookup2(const DiffScene& sc, V3 pos) {
  Lookup L = grid_lookup(sc.gp, pos);
  Tri t;
  t.i = L.i;
  t.j = L.j;
  t.k = L.k;
  t.rho_p = L.rho;
  double fx = L.fx, fy = L.fy, fz = L.fz;
  double r = 0;
  for (int c = 0; c < 8; ++c) {
    int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    double w = (dx ? fx : 1 - fx) * (dy ? fy : 1 - fy) * (dz ? fz : 1 - fz);
    t.w[c] = w;
    int ii = L.i + dx, jj = L.j + dy, kk = L.k + dz;
    if (ii < 0 || jj < 0 || kk < 0 || ii >= sc.gp.nx || jj >= sc.gp.ny || kk >= sc.gp.nz) continue;
    r += w * sc.theta[((int64_t)kk * sc.gp.ny + jj) * sc.gp.nx + ii];
  }
  t.rho = r;
  return t;
}

// Ratio tracking without roulette on the stream of walk `key`; records the
// lookups and returns T.
// *n: the stream's next draw (advanced; 0 when not given).
double ratio_track(const DiffScene& sc, uint32_t key, V3 o, V3 d, float tmax, std::vector<Tri>* rec,
                   uint32_t* n = nullptr) {
  uint32_t n0 = 0;
  uint32_t& draw = n ? *n : n0;
  float t_min, t_max;
  if (!box_intersect_pbrt(sc.gp.bmin, sc.gp.bmax, o, d, tmax, &t_min, &t_max)) return 1.0;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  if (!(sc.mu_f > 0.0f)) return 1.0;
  double T = 1.0;
  const int cap = walk_cap(sc.mu_f, t, t_max);
  for (int step = 0; step < cap; ++step) {
    t += -flg_logf(walk_draw(key, draw++)) / sc.mu_f;
    if (t >= t_max) break;
    V3 pos = o + d * t;
    Tri q = lookup2(sc, pos);
    T *= 1.0 - sc.s * q.rho / sc.mu;
    if (rec) rec->push_back(q);
  }
  return T;
}

struct Event {  // main-path tentative collision
  Tri q;
  bool real;
};

// Differential delta tracking along (o, d); records events; returns true on a
// real collision (position in *pos).
bool delta_track(const DiffScene& sc, Halton& hs, V3 o, V3 d, float tmax, V3* pos, std::vector<Event>* ev) {
  float t_min, t_max;
  if (!box_intersect_pbrt(sc.gp.bmin, sc.gp.bmax, o, d, tmax, &t_min, &t_max)) return false;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  if (!(sc.mu_f > 0.0f)) return false;
  const int cap = walk_cap(sc.mu_f, t, t_max);
  for (int step = 0; step < cap; ++step) {
    const float u = hs.get1D();
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / sc.mu_f;
    if (t >= t_max) break;
    V3 p = o + d * t;
    Tri q = lookup2(sc, p);
    float pr = sc.s_f * q.rho_p / sc.mu_f;
    float p_n = 1.0f - pr;
    bool real = hs.get1D() > p_n;
    ev->push_back(Event{q, real});
    if (real) {
      *pos = p;
      return true;
    }
  }
  return false;
}

double w_real(const DiffScene& sc, const Tri& q) { return sc.alpha * (q.rho / (double)q.rho_p); }
double w_null(const DiffScene& sc, const Tri& q) {
  // a null collision is only chosen with probability 1 - s rho_p / mu > 0,
  // except for u == 0 exactly; that measure-zero case keeps weight 1
  double den = 1.0 - sc.s * (double)q.rho_p / sc.mu;
  if (den == 0.0) return 1.0;
  return (1.0 - sc.s * q.rho / sc.mu) / den;
}

struct NeeSample {
  double K;  // f * Li / pdf / light_pdf (no transmittance)
  bool valid;
  V3 p_light;
};

// Light choice and sample of UniformSampleOneLight / EstimateDirect
// (integrator.cpp:69-120) without the shadow transmittance.
NeeSample nee_sample(const DiffScene& sc, Halton& hs, V3 p, V3 wo) {
  NeeSample r{0.0, false, v3(0, 0, 0)};
  int n_lights = (sc.point.present ? 1 : 0) + (sc.env.present ? 1 : 0);
  if (n_lights == 0) return r;
  int light_num = (int)(hs.get1D() * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  double light_pdf = (double)(float)(1.0 / (double)n_lights);
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
  // a light sample counts when Li is non-zero in some channel
  if (pdf == 0.0f || (Li.x == 0.0f && Li.y == 0.0f && Li.z == 0.0f)) return r;
  float f = hg_p(wi, wo, sc.g_f);
  if (f == 0.0f) return r;
  r.K = (double)f * (double)channel(Li, sc.ch) / (double)pdf / light_pdf;
  r.valid = true;
  r.p_light = p_light;
  return r;
}

// grad += g * trilinear weights of lookup q (dense grid, x fastest)
void deposit(const DiffScene& sc, double* grad, const Tri& q, double g) {
  if (!grad || g == 0.0) return;
  for (int c = 0; c < 8; ++c) {
    int ii = q.i + (c & 1), jj = q.j + ((c >> 1) & 1), kk = q.k + (c >> 2);
    if (ii < 0 || jj < 0 || kk < 0 || ii >= sc.gp.nx || jj >= sc.gp.ny || kk >= sc.gp.nz) continue;
    grad[((int64_t)kk * sc.gp.ny + jj) * sc.gp.nx + ii] += g * q.w[c];
  }
}

// scale * d(prod f)/d rho_i for every recorded ratio-tracking lookup
void deposit_tr(const DiffScene& sc, double* grad, const std::vector<Tri>& r, double scale) {
  if (!grad) return;
  size_t n = r.size();
  std::vector<double> f(n), pre(n + 1, 1.0), suf(n + 1, 1.0);
  for (size_t i = 0; i < n; ++i) f[i] = 1.0 - sc.s * r[i].rho / sc.mu;
  for (size_t i = 0; i < n; ++i) pre[i + 1] = pre[i] * f[i];
  for (size_t i = n; i-- > 0;) suf[i] = suf[i + 1] * f[i];
  for (size_t i = 0; i < n; ++i) deposit(sc, grad, r[i], scale * (-sc.s / sc.mu) * pre[i] * suf[i + 1]);
}

// The weights of a delta-tracking walk's events into beta; with grad, each
// event's derivative weighted by the radiance gathered after it.
void apply_events(const DiffScene& sc, const std::vector<Event>& ev, double L, double L_total, double delta,
                  double& beta, double* grad, double* galpha) {
  for (const Event& e : ev) {
    if (e.real) {
      if (grad) {
        double S = L_total - L;
        if (e.q.rho != 0.0) deposit(sc, grad, e.q, delta * S / e.q.rho);
        *galpha += delta * S / sc.alpha;
      }
      beta *= w_real(sc, e.q);
    } else {
      double fnull = 1.0 - sc.s * e.q.rho / sc.mu;
      if (grad && fnull != 0.0) deposit(sc, grad, e.q, delta * (L_total - L) * (-sc.s / sc.mu) / fnull);
      beta *= w_null(sc, e.q);
    }
  }
}

// Replays one path.  grad == nullptr: forward only (returns L).  Otherwise
// accumulates delta * dL/dtheta into grad and delta * dL/dalpha into *galpha,
// given the path's total L_total.
double diff_path(const DiffScene& sc, Halton& hs, uint32_t path, V3 o, V3 d, double delta,
                 double L_total, double* grad, double* galpha, uint64_t* lookups) {
  double L = 0.0;
  const double Le = sc.env.present ? (double)channel(sc.env.radiance, sc.ch) : 0.0;
  float t_min, t_max;
  if (!box_intersect(sc.gp.bmin, sc.gp.bmax, o, d, &t_min, &t_max)) return Le;
  std::vector<Tri> rec;
  std::vector<Event> ev;
  double T = ratio_track(sc, walk_key(path, 0), o, d, INFINITY, grad ? &rec : nullptr);
  if (lookups) *lookups += rec.size();
  deposit_tr(sc, grad, rec, delta * Le);
  L += T * Le;
  o = o + d * t_min;
  double beta = 1.0;
  const float rr = 0.01f;
  for (int bounces = 1;; ++bounces) {
    ev.clear();
    V3 pos;
    bool hit = delta_track(sc, hs, o, d, INFINITY, &pos, &ev);
    if (lookups) *lookups += ev.size();
    apply_events(sc, ev, L, L_total, delta, beta, grad, galpha);
    if (!hit) break;
    V3 wo = -d;
    float u = hs.get1D();
    float v = hs.get1D();
    V3 wi = hg_sample(wo, u, v, sc.g_f);
    NeeSample ns = nee_sample(sc, hs, pos, wo);
    if (ns.valid) {
      V3 diff = ns.p_light - pos;
      float nrm = norm(diff);
      V3 dd = diff / nrm;
      rec.clear();
      double Tsh = ratio_track(sc, walk_key(path, (uint32_t)bounces), pos, dd, nrm - SHADOW_EPS,
                               grad ? &rec : nullptr);
      if (lookups) *lookups += rec.size();
      deposit_tr(sc, grad, rec, delta * beta * ns.K);
      L += beta * ns.K * Tsh;
    }
    if (bounces >= sc.max_depth || beta == 0.0) break;
    o = pos;
    d = normalize(wi);
    if (bounces > 10) {
      if (hs.get1D() < rr) break;
      beta /= (double)(1.0f - rr);
    }
  }
  return L;
}

constexpr int kMaxCrossings = 256;

// Ratio track of walk `key` from a medium vertex through the interfaces
// (VolEstimateTr): the product over the inside segments up to the last
// interface before tmax; lookups recorded.
double mesh_ratio_track(const Lbvh& b, bool compat, const DiffScene& sc, uint32_t key, V3 o, V3 d, float tmax,
                        std::vector<Tri>* rec) {
  double T = 1.0;
  uint32_t n = 0;
  bool inside = true;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t = 0;
    V3 ng{};
    if (lbvh_trace(b, o, d, t_max, compat, &t, &ng, nullptr) < 0) return T;
    if (inside) T *= ratio_track(sc, key, o, d, t, rec, &n);
    t_max -= t;
    if (t_max < 0) return T;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);
    d = normalize(d);
    inside = dot(d, ng) < 0;
  }
  return T;
}

// One path of the estimator over the interface mesh (VolPathIntegrator's
// loop); grad as diff_path.
double diff_path_mesh(const Lbvh& b, bool compat, const DiffScene& sc, Halton& hs, uint32_t path, V3 o, V3 d,
                      double delta, double L_total, double* grad, double* galpha, uint64_t* lookups) {
  double L = 0.0, beta = 1.0;
  const double Le = sc.env.present ? (double)channel(sc.env.radiance, sc.ch) : 0.0;
  bool inside = false;  // the camera is outside the mesh
  int bounces = 0, crossings = 0;
  std::vector<Tri> rec;
  std::vector<Event> ev;
  while (true) {
    float t = 0;
    V3 ng{};
    const bool hit = lbvh_trace(b, o, d, INFINITY, compat, &t, &ng, nullptr) >= 0;
    if (bounces == 0 && !hit && sc.env.present) L += beta * Le;
    if (!hit || bounces >= sc.max_depth) break;
    bool scattered = false;
    V3 pos{};
    if (inside) {
      ev.clear();
      scattered = delta_track(sc, hs, o, d, t, &pos, &ev);
      if (lookups) *lookups += ev.size();
      apply_events(sc, ev, L, L_total, delta, beta, grad, galpha);
    }
    if (scattered) {
      const V3 wo = -d;
      NeeSample ns = nee_sample(sc, hs, pos, wo);
      if (ns.valid) {
        V3 diff = ns.p_light - pos;
        float nrm = norm(diff);
        V3 dd = diff / nrm;
        rec.clear();
        double Tsh = mesh_ratio_track(b, compat, sc, walk_key(path, (uint32_t)bounces + 1), pos, dd,
                                      nrm - SHADOW_EPS, grad ? &rec : nullptr);
        if (lookups) *lookups += rec.size();
        deposit_tr(sc, grad, rec, delta * beta * ns.K);
        L += beta * ns.K * Tsh;
      }
      const float u = hs.get1D();
      const float v = hs.get1D();
      const V3 wi = hg_sample(wo, u, v, sc.g_f);
      if (beta == 0.0) break;
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

void setup_scene(DiffScene& sc, Camera& cam, const double* theta, const float* theta_p, int32_t nx, int32_t ny,
                 int32_t nz, const float* bounds6, const float* params4, const float* lights11, int32_t W, int32_t H,
                 int32_t spp, int32_t max_depth, int32_t ch) {
  sc.gp.nx = nx; sc.gp.ny = ny; sc.gp.nz = nz;
  sc.gp.data = theta_p;
  sc.gp.bmin = v3(bounds6[0], bounds6[1], bounds6[2]);
  sc.gp.bmax = v3(bounds6[3], bounds6[4], bounds6[5]);
  sc.gp.scale = v3((float)nx / (sc.gp.bmax.x - sc.gp.bmin.x), (float)ny / (sc.gp.bmax.y - sc.gp.bmin.y),
                   (float)nz / (sc.gp.bmax.z - sc.gp.bmin.z));
  sc.theta = theta;
  sc.s_f = params4[0];
  sc.alpha_f = params4[1];
  sc.g_f = params4[2];
  sc.mu_f = params4[3];
  sc.s = sc.s_f;
  sc.alpha = sc.alpha_f;
  sc.mu = sc.mu_f;
  sc.point.present = lights11[0] != 0.0f;
  sc.point.pos = v3(lights11[1], lights11[2], lights11[3]);
  sc.point.intensity = v3(lights11[4], lights11[5], lights11[6]);
  if (lights11[7] != 0.0f) sc.env.build(v3(lights11[8], lights11[9], lights11[10]), sc.gp.bmin, sc.gp.bmax);
  sc.ch = ch < 0 ? 0 : ch > 2 ? 2 : ch;
  sc.cam = &cam;
  sc.W = W; sc.H = H; sc.spp = spp; sc.max_depth = max_depth;
}

// Every path of the image on the per-sample schedule, over nthreads threads:
// the forward (lpath, img) or, with dldpix, the adjoint (grad, galpha).
template <class PathFn>
void run_paths(const DiffScene& sc, PathFn path_fn, double* lpath, double* img, const double* dldpix,
               double* grad, double* galpha, int nthreads, uint64_t* lookups) {
  const int W = sc.W, H = sc.H, spp = sc.spp;
  const bool adj = dldpix != nullptr;
  const int64_t nvox = (int64_t)sc.gp.nx * sc.gp.ny * sc.gp.nz;
  if (nthreads < 1) nthreads = 1;
  std::vector<std::vector<double>> gpart(adj ? nthreads : 0);
  std::vector<double> apart(nthreads, 0.0);
  std::vector<uint64_t> lk(nthreads, 0);
  std::atomic<int64_t> next{0};
  const int64_t npix = (int64_t)W * H;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      if (adj) gpart[t].assign(nvox, 0.0);
      Halton hs((uint64_t)spp, W, H);
      while (true) {
        int64_t pix = next.fetch_add(1) * sc.pix_stride;
        if (pix >= npix) break;
        int x = (int)(pix % W), y = (int)(pix / W);
        double acc = 0;
        for (int s = 0; s < spp; ++s) {
          hs.start(x, y, (uint64_t)s);
          float ux = hs.get1D();
          float uy = hs.get1D();
          V3 o, d;
          sc.cam->generate((float)x + ux, (float)y + uy, W, H, &o, &d);
          int64_t p = pix * spp + s;
          if (adj) {
            path_fn(hs, (uint32_t)p, o, d, dldpix[pix] / spp, lpath[p], gpart[t].data(), &apart[t], &lk[t]);
          } else {
            double L = path_fn(hs, (uint32_t)p, o, d, 0.0, 0.0, nullptr, nullptr, &lk[t]);
            if (lpath) lpath[p] = L;
            acc += L;
          }
        }
        if (!adj && img) img[pix] = acc / spp;
      }
    });
  for (auto& x : th) x.join();
  if (adj) {
    for (int64_t v = 0; v < nvox; ++v) {
      double sum = 0;
      for (int t = 0; t < nthreads; ++t) sum += gpart[t][v];
      grad[v] = sum;
    }
    double a = 0;
    for (double x : apart) a += x;
    *galpha = a;
  }
  if (lookups) {
    uint64_t tot = 0;
    for (auto v : lk) tot += v;
    *lookups = tot;
  }
}

}  // namespace

// The estimator over the interface mesh b (the medium on b's bounding box).
void diff_mesh_render(const Lbvh& b, bool compat, const double* theta, const float* theta_p, int32_t nx,
                      int32_t ny, int32_t nz, const float* params4, const float* lights11, const float* cam10,
                      int32_t W, int32_t H, int32_t spp, int32_t max_depth, double* lpath, double* img,
                      const double* dldpix, double* grad, double* galpha, int nthreads, uint64_t* lookups,
                      int32_t channel) {
  const float bounds6[6] = {b.glo.x, b.glo.y, b.glo.z, b.ghi.x, b.ghi.y, b.ghi.z};
  DiffScene sc;
  Camera cam(v3(cam10[0], cam10[1], cam10[2]), v3(cam10[3], cam10[4], cam10[5]),
             v3(cam10[6], cam10[7], cam10[8]), cam10[9]);
  setup_scene(sc, cam, theta, theta_p, nx, ny, nz, bounds6, params4, lights11, W, H, spp, max_depth, channel);
  run_paths(
      sc,
      [&](Halton& hs, uint32_t p, V3 o, V3 d, double delta, double Lt, double* g, double* ga, uint64_t* lk) {
        return diff_path_mesh(b, compat, sc, hs, p, o, d, delta, Lt, g, ga, lk);
      },
      lpath, img, dldpix, grad, galpha, nthreads, lookups);
}

}  // namespace orc

using namespace orc;

extern "C" {

// theta: evaluated grid (double, nx*ny*nz, x fastest); theta_p: proposal
// grid (float).  params: s, alpha, g, mu.  lights11 as orc_volpath_render
// (RGB); channel: the colour channel (0, 1, 2) whose radiance the call
// computes (the estimator is linear in the lights' radiance and no decision
// depends on it, so the three channels are three such calls).
// Outputs: lpath[W*H*spp] (double, may be NULL), img[W*H] (double mean over
// samples, may be NULL).  With dldpix != NULL the adjoint runs instead:
// grad (nx*ny*nz doubles, overwritten) and galpha receive
// sum_pixels dldpix[pixel] * dI[pixel]/dtheta (dI = mean over samples),
// using lpath (from a previous forward call) as the paths' totals.
void orc_diff_render(const double* theta, const float* theta_p, int32_t nx, int32_t ny, int32_t nz,
                     const float* bounds6, const float* params4, const float* lights11,
                     const float* cam10, int32_t W, int32_t H, int32_t spp, int32_t max_depth,
                     double* lpath, double* img, const double* dldpix, double* grad,
                     double* galpha, int nthreads, uint64_t* lookups, int32_t channel) {
  DiffScene sc;
  Camera cam(v3(cam10[0], cam10[1], cam10[2]), v3(cam10[3], cam10[4], cam10[5]),
             v3(cam10[6], cam10[7], cam10[8]), cam10[9]);
  setup_scene(sc, cam, theta, theta_p, nx, ny, nz, bounds6, params4, lights11, W, H, spp, max_depth, channel);
  run_paths(
      sc,
      [&](Halton& hs, uint32_t p, V3 o, V3 d, double delta, double Lt, double* g, double* ga, uint64_t* lk) {
        return diff_path(sc, hs, p, o, d, delta, Lt, g, ga, lk);
      },
      lpath, img, dldpix, grad, galpha, nthreads, lookups);
}

// orc_diff_render over the pixels 0, stride, 2 stride, ... only (the CPU
// baseline's bounded sample): lpath and img entries of other pixels are left
// untouched; the adjoint's gradient covers the sampled pixels.
void orc_diff_render_strided(const double* theta, const float* theta_p, int32_t nx, int32_t ny, int32_t nz,
                             const float* bounds6, const float* params4, const float* lights11,
                             const float* cam10, int32_t W, int32_t H, int32_t spp, int32_t max_depth,
                             int64_t stride, double* lpath, double* img, const double* dldpix, double* grad,
                             double* galpha, int nthreads, uint64_t* lookups, int32_t channel) {
  DiffScene sc;
  Camera cam(v3(cam10[0], cam10[1], cam10[2]), v3(cam10[3], cam10[4], cam10[5]),
             v3(cam10[6], cam10[7], cam10[8]), cam10[9]);
  setup_scene(sc, cam, theta, theta_p, nx, ny, nz, bounds6, params4, lights11, W, H, spp, max_depth, channel);
  sc.pix_stride = stride < 1 ? 1 : stride;
  run_paths(
      sc,
      [&](Halton& hs, uint32_t p, V3 o, V3 d, double delta, double Lt, double* g, double* ga, uint64_t* lk) {
        return diff_path(sc, hs, p, o, d, delta, Lt, g, ga, lk);
      },
      lpath, img, dldpix, grad, galpha, nthreads, lookups);
}

}  // extern "C"
