// Differentiable volume rendering: forward and path-replay adjoint.
//
// No counterpart in the reference (the closest code is OpenVDBVolume::tr and
// ::sample, src/volume.cpp:75-148, inside SVolIntegrator::Li,
// src/integrator.cpp:328-405).  The estimator keeps SVolIntegrator's path
// structure and makes every branch independent of the evaluated density grid
// theta (differential delta tracking, Nimier-David et al. 2022; path replay
// backpropagation, Vicini et al. 2021):
//   * fixed majorant mu; collision decisions on the proposal grid theta_p
//     (real iff u > 1 - s rho_p / mu) with weight alpha rho / rho_p on a real
//     collision and (1 - s rho / mu) / (1 - s rho_p / mu) on a null one;
//   * ratio tracking without roulette on camera and shadow rays, factor
//     1 - s rho / mu per tentative collision;
//   * the path roulette after 10 bounces is kept (it depends on u only).
// The main path (delta tracking, phase, light choice, roulette) draws the
// path's Halton dimensions in SVolIntegrator's order; every ratio-tracking
// walk draws from its own counter-based stream (walk_draw: walk 0 is the
// camera ray, walk b the shadow ray of bounce b).  A walk therefore stops as
// soon as its product and every derivative are fixed (two zero factors, or a
// product that reached 0) without shifting any later draw.
// The adjoint replays each path with the same draws.  For a main-path event j
// the contributions that carry its weight sum to L_total - (radiance emitted
// before j), so the gradient needs no division by the throughput.  A
// ratio-tracking product is measured (product of the non-zero factors, count
// of zero factors), then walked again to deposit d(product)/d(rho) per
// lookup, which stays exact when a factor is 0.  The forward can record every
// walk's measurement on a tape, which spares the adjoint the first walk.
#include <algorithm>
#include <cmath>

#include "volume.hpp"

namespace flg {

namespace {

constexpr int kMaxSteps = 4096;

// Steps a walk of this estimator may take over [t, t_max]: 8 times the
// expected count (mu * length) plus 64, at most kMaxSteps.  A walk of ordinary
// draws never comes near it; on the main path it trims the delta-tracking
// walks of the first Halton indices, whose high dimensions are tiny
// (index / prime) and give minute free flights, and it bounds every loop.
FLG_D int walk_cap(float mu, float t, float t_max) {
  const float e = 8.0f * mu * (t_max - t) + 64.0f;
  return e < (float)kMaxSteps ? (int)e : kMaxSteps;
}

struct DiffDev {
  VolDev v;   // evaluated grid theta (densities rho)
  VolDev vp;  // proposal grid theta_p (decisions)
  bool same;  // theta_p is theta
  float s, alpha, g, mu;
  // walk tape (NULL: none).  Walk w of path p: tape[w * npaths + p] = product
  // of its non-zero factors (0 once two factors are 0); bit w of
  // mask_hi[p]:mask_lo[p] set iff exactly one factor is 0.
  float* tape;
  uint32_t* mask_lo;
  uint32_t* mask_hi;
  uint32_t npaths;
};

struct Look2 {
  Tri8 t;     // lookup on theta (weights, base, rho)
  float rho_p;
};

FLG_D Look2 look2(const DiffDev& D, V3 p) {
  Look2 r;
  r.t = lookup(D.v, p);
  r.rho_p = D.same ? r.t.rho : lookup(D.vp, p).rho;
  return r;
}

// gradient scatter: g * trilinear weight into the caller's grid (x fastest)
FLG_D void deposit(const DiffDev& D, float* grad, const Tri8& t, float g) {
  if (t.base < 0 || g == 0.0f) return;
  const int px = D.v.px, py = D.v.ny + 2;
  const int b = t.base;
  const int i = b % px - 1, j = (b / px) % py - 1, k = b / D.v.pxy - 1;
  const float wx[2] = {1.0f - t.fx, t.fx}, wy[2] = {1.0f - t.fy, t.fy}, wz[2] = {1.0f - t.fz, t.fz};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    const int ii = i + dx, jj = j + dy, kk = k + dz;
    if (ii < 0 || jj < 0 || kk < 0 || ii >= D.v.nx || jj >= D.v.ny || kk >= D.v.nz) continue;
    const float w = wx[dx] * wy[dy] * wz[dz];
    unsafeAtomicAdd(grad + ((int64_t)kk * D.v.ny + jj) * D.v.nx + ii, g * w);
  }
}

// A ratio-tracking product: nonzero = product of the non-zero factors,
// zeros = number of factors equal to 0.
struct Track {
  float nonzero;
  int zeros;
};

// First walk of a ratio track without roulette (the camera and shadow rays):
// factor 1 - s rho / mu per tentative collision.  Stops once the product and
// every derivative are fixed: two zero factors (every derivative is 0), or a
// product of the non-zero factors that reached 0.
FLG_D Track track_measure(const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax) {
  Track r{1.0f, 0};
  float t_min, t_max;
  if (!box_pbrt(D.v.bmin, D.v.bmax, o, d, tmax, &t_min, &t_max)) return r;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  const int cap = walk_cap(D.mu, t, t_max);
  for (int step = 0; step < cap; ++step) {
    t += -flg_logf(walk_draw(key, (uint32_t)step)) / D.mu;
    if (t >= t_max) break;
    const float f = 1.0f - D.s * lookup(D.v, o + d * t).rho / D.mu;
    if (f == 0.0f) {
      if (++r.zeros >= 2) break;
    } else {
      r.nonzero *= f;
      if (r.nonzero == 0.0f) break;
    }
  }
  return r;
}

// The product's value.
FLG_D float track_value(const Track& r) { return r.zeros > 0 ? 0.0f : r.nonzero; }

// Whether a measured track has a non-zero derivative.
FLG_D bool track_live(const Track& r) { return r.zeros < 2 && r.nonzero != 0.0f; }

// Second walk (same draws): deposits scale * d(product)/d(rho_i) for every
// lookup i, d/d(rho_i) = k prod_{j != i} f_j with k = -s / mu.
FLG_D void track_deposit(const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax, const Track& r,
                         float scale, float* grad) {
  float t_min, t_max;
  if (!box_pbrt(D.v.bmin, D.v.bmax, o, d, tmax, &t_min, &t_max)) return;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  const float k = -D.s / D.mu;
  const int cap = walk_cap(D.mu, t, t_max);
  for (int step = 0; step < cap; ++step) {
    t += -flg_logf(walk_draw(key, (uint32_t)step)) / D.mu;
    if (t >= t_max) break;
    const Tri8 q = lookup(D.v, o + d * t);
    const float f = 1.0f - D.s * q.rho / D.mu;
    if (r.zeros == 0) {
      deposit(D, grad, q, scale * k * (r.nonzero / f));
    } else if (f == 0.0f) {  // the one zero factor: the others' product
      deposit(D, grad, q, scale * k * r.nonzero);
      break;
    }
  }
}

// Walk w of path p: measured (and recorded on the tape by a forward that has
// one), or read back from the tape by the adjoint.
template <bool ADJ>
FLG_D Track track(const DiffDev& D, uint32_t key, uint32_t p, uint32_t w, V3 o, V3 d, float tmax,
                  uint64_t& mask) {
  if (ADJ && D.tape) {
    return Track{D.tape[(int64_t)w * D.npaths + p], (int)((mask >> w) & 1u)};
  }
  const Track r = track_measure(D, key, o, d, tmax);
  if (!ADJ && D.tape) {
    D.tape[(int64_t)w * D.npaths + p] = r.zeros >= 2 ? 0.0f : r.nonzero;
    if (r.zeros == 1) mask |= (uint64_t)1 << w;
  }
  return r;
}

// Light choice and sample (UniformSampleOneLight / EstimateDirect,
// integrator.cpp:69-120) without the shadow transmittance: returns
// K = f Li / pdf / light_pdf, 0 if no contribution.
FLG_D float nee_sample(const DiffDev& D, const LightsDev& L, const HaltonTables& tab,
                       const HaltonParams& hp, PathSampler& ps, V3 p, V3 wo, V3* p_light) {
  const int n_lights = (L.has_point ? 1 : 0) + (L.env.present ? 1 : 0);
  if (n_lights == 0) return 0.0f;
  int light_num = (int)(ps.next(tab, hp) * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  const float light_pdf = (float)(1.0 / (double)n_lights);
  const bool use_point = L.has_point && light_num == 0;
  float u0 = ps.next(tab, hp);
  float u1 = ps.next(tab, hp);
  V3 wi;
  float pdf, Li;
  if (use_point) {
    V3 diff = L.point_pos - p;
    float d2 = sqnorm(diff);
    wi = normalize(diff);
    Li = L.point_intensity / d2;
    pdf = 1.0f;
    *p_light = L.point_pos;
  } else {
    Li = env_sample_li(L.env, p, u0, u1, &wi, &pdf, p_light);
  }
  if (pdf == 0.0f || Li == 0.0f) return 0.0f;
  float f = hg_p(wi, wo, D.g);
  if (f == 0.0f) return 0.0f;
  return f * Li / pdf / light_pdf;
}

// One path, as a start (camera ray, box entry, primary transmittance) and one
// step per scattering event, so that persistent lanes can interleave paths.
// ADJ = false: accumulates the radiance.  ADJ = true: replays with the path's
// total L_total and adjoint weight delta, depositing gradients.
struct DiffPath {
  PathSampler ps;
  V3 o, d;
  float beta, acc, delta, L_total;
  int bounces;
  uint32_t path;
  uint64_t mask;  // tape bits of the path's walks with exactly one zero factor
};

// Returns true when the path is complete (or, in the adjoint, has no weight).
template <bool ADJ>
FLG_D bool diff_begin(const DiffDev& D, const LightsDev& L, const HaltonTables& tab,
                      const HaltonParams& hp, const CamDevV& cam, int32_t spp, uint32_t p,
                      const float* lpath, const float* dl_drgb, float* grad, DiffPath& s) {
  const uint32_t pix = p / (uint32_t)spp, smp = p - pix * (uint32_t)spp;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.acc = 0.0f;
  s.path = p;
  s.mask = 0;
  if (ADJ) {
    const float* g = dl_drgb + 3 * (int64_t)pix;
    s.delta = (g[0] + g[1] + g[2]) / (float)spp;
    if (s.delta == 0.0f) return true;
    s.L_total = lpath[p];
    if (D.tape) s.mask = ((uint64_t)D.mask_hi[p] << 32) | D.mask_lo[p];
  }
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  V3 o, d;
  gen_ray_v(cam, (float)x + ux, (float)y + uy, &o, &d);
  const float Le = L.env.present ? L.env.radiance : 0.0f;
  float t_min, t_max;
  if (!box_enter(D.v.bmin, D.v.bmax, o, d, &t_min, &t_max)) {
    s.acc = Le;
    return true;
  }
  const uint32_t key = walk_key(p, 0);
  const Track tr = track<ADJ>(D, key, p, 0, o, d, __builtin_inff(), s.mask);
  if (ADJ && track_live(tr) && s.delta * Le != 0.0f)
    track_deposit(D, key, o, d, __builtin_inff(), tr, s.delta * Le, grad);
  s.acc += track_value(tr) * Le;
  s.o = o + d * t_min;
  s.d = d;
  s.beta = 1.0f;
  s.bounces = 1;
  return false;
}

// One scattering event; returns true when the path is complete.
template <bool ADJ>
FLG_D bool diff_step(const DiffDev& D, const LightsDev& L, const HaltonTables& tab,
                     const HaltonParams& hp, int max_depth, float* grad, float* galpha, DiffPath& s) {
  const float rr = 0.01f;
  const float k = -D.s / D.mu;
  // differential delta tracking
  float tb0, tb1;
  bool hit = false;
  V3 pos = s.o;
  if (box_pbrt(D.v.bmin, D.v.bmax, s.o, s.d, __builtin_inff(), &tb0, &tb1) && D.mu > 0.0f) {
    float t = smax(tb0, 0.0f);
    tb1 = smin(tb1, __builtin_inff());
    const int cap = walk_cap(D.mu, t, tb1);
    for (int step = 0; step < cap; ++step) {
      const float u = s.ps.next(tab, hp);
      if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
      t += -flg_logf(1.0f - u) / D.mu;
      if (t >= tb1) break;
      const V3 p = s.o + s.d * t;
      const Look2 q = look2(D, p);
      const float p_n = 1.0f - D.s * q.rho_p / D.mu;
      const bool real = s.ps.next(tab, hp) > p_n;
      if (real) {
        if (ADJ) {
          const float S = s.L_total - s.acc;
          if (q.t.rho != 0.0f) deposit(D, grad, q.t, s.delta * S / q.t.rho);
          *galpha += s.delta * S / D.alpha;
        }
        s.beta *= D.alpha * (q.t.rho / q.rho_p);
        pos = p;
        hit = true;
        break;
      }
      const float fnull = 1.0f - D.s * q.t.rho / D.mu;
      if (ADJ && fnull != 0.0f) deposit(D, grad, q.t, s.delta * (s.L_total - s.acc) * k / fnull);
      // a null collision has probability p_n > 0 except for u == 0
      // exactly; that measure-zero case keeps weight 1
      if (p_n != 0.0f) s.beta *= fnull / p_n;
    }
  }
  if (!hit) return true;
  const V3 wo = -s.d;
  const float u = s.ps.next(tab, hp);
  const float vv = s.ps.next(tab, hp);
  const V3 wi = hg_sample(wo, u, vv, D.g);
  V3 p_light;
  const float K = nee_sample(D, L, tab, hp, s.ps, pos, wo, &p_light);
  if (K != 0.0f) {
    const V3 diff = p_light - pos;
    const float nrm = sqrtf(sqnorm(diff));
    const V3 dd = diff / nrm;
    const uint32_t w = (uint32_t)s.bounces;
    const uint32_t key = walk_key(s.path, w);
    const Track tr = track<ADJ>(D, key, s.path, w, pos, dd, nrm - SHADOW_EPS, s.mask);
    if (ADJ && track_live(tr) && s.delta * s.beta * K != 0.0f)
      track_deposit(D, key, pos, dd, nrm - SHADOW_EPS, tr, s.delta * s.beta * K, grad);
    s.acc += s.beta * K * track_value(tr);
  }
  if (s.bounces >= max_depth || s.beta == 0.0f) return true;
  s.o = pos;
  s.d = normalize(wi);
  if (s.bounces > 10) {
    if (s.ps.next(tab, hp) < rr) return true;
    s.beta /= (1.0f - rr);
  }
  ++s.bounces;
  return false;
}

// Forward: the path's radiance and its tape bits.
FLG_D void diff_finish(const DiffDev& D, uint32_t p, const DiffPath& s, float* lpath) {
  lpath[p] = s.acc;
  if (D.tape) {
    D.mask_lo[p] = (uint32_t)s.mask;
    D.mask_hi[p] = (uint32_t)(s.mask >> 32);
  }
}

template <bool ADJ>
FLG_D void reduce_galpha(float ga, float* galpha) {
  if (ADJ) {
    for (int off = 32; off >= 1; off >>= 1) ga += __shfl_xor(ga, off, 64);
    if ((threadIdx.x & 63) == 0 && ga != 0.0f) unsafeAtomicAdd(galpha, ga);
  }
}

// One lane per path sample.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_diff(DiffDev D, LightsDev L, HaltonTables tab_g,
                                              HaltonParams hp, CamDevV cam, int32_t spp,
                                              int32_t max_depth, float* lpath,
                                              const float* dl_drgb, float* grad, float* galpha) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npaths = (int64_t)cam.w * cam.h * spp;
  float ga = 0.0f;
  if (p < npaths) {
    DiffPath s;
    bool done = diff_begin<ADJ>(D, L, tab, hp, cam, spp, (uint32_t)p, lpath, dl_drgb, grad, s);
    while (!done) done = diff_step<ADJ>(D, L, tab, hp, max_depth, grad, &ga, s);
    if (!ADJ) diff_finish(D, (uint32_t)p, s, lpath);
  }
  reduce_galpha<ADJ>(ga, galpha);
}

// Persistent lanes with path regeneration (regen_loop); same per-path code.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_diff_regen(DiffDev D, LightsDev L, HaltonTables tab_g,
                                                    HaltonParams hp, CamDevV cam, int32_t spp,
                                                    int32_t max_depth, float* lpath,
                                                    const float* dl_drgb, float* grad, float* galpha,
                                                    unsigned int* next_path) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t npaths = (uint32_t)((int64_t)cam.w * cam.h * spp);
  float ga = 0.0f;
  regen_loop<DiffPath>(
      npaths, next_path,
      [&](uint32_t p, DiffPath& s) {
        return diff_begin<ADJ>(D, L, tab, hp, cam, spp, p, lpath, dl_drgb, grad, s);
      },
      [&](DiffPath& s) { return diff_step<ADJ>(D, L, tab, hp, max_depth, grad, &ga, s); },
      [&](uint32_t p, const DiffPath& s) {
        if (!ADJ) diff_finish(D, p, s, lpath);
      });
  reduce_galpha<ADJ>(ga, galpha);
}

__global__ void k_zero_counter(unsigned int* c) { *c = 0u; }

// Launches one pass (forward or adjoint) with the schedule of cfg.
template <bool ADJ>
int launch_diff(flg_ctx* ctx, const flg_render_cfg* cfg, const DiffDev& D, const LightsDev& L,
                const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, float* lpath,
                const float* dl, float* grad, float* galpha) {
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  const int sched = cfg->schedule ? cfg->schedule : 2;
  if (sched == 1) {
    k_diff<ADJ><<<div_up(npaths, 256), 256, 0, ctx->stream>>>(D, L, tab, hp, cam, cfg->spp, cfg->max_depth,
                                                               lpath, dl, grad, galpha);
    return FLG_OK;
  }
  int st = FLG_OK;
  unsigned int* counter = (unsigned int*)scratch(ctx, ADJ ? "diff_counter_adj" : "diff_counter_fwd", 256, &st);
  if (!counter) return st;
  k_zero_counter<<<1, 1, 0, ctx->stream>>>(counter);
  const int grid = (int)std::min<int64_t>(div_up(npaths, 256), (int64_t)ctx->num_cus * 8);
  k_diff_regen<ADJ><<<grid, 256, 0, ctx->stream>>>(D, L, tab, hp, cam, cfg->spp, cfg->max_depth, lpath, dl,
                                                   grad, galpha, counter);
  return FLG_OK;
}

__global__ __launch_bounds__(256) void k_mean(const float* lpath, int64_t npix, int32_t spp,
                                              float* rgb) {
  int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= npix) return;
  float c = 0.0f;
  for (int s = 0; s < spp; ++s) c += lpath[pix * spp + s];
  c = c / (float)spp;
  rgb[3 * pix] = c;
  rgb[3 * pix + 1] = c;
  rgb[3 * pix + 2] = c;
}

// Walk tape: one float per walk (camera walk and one shadow walk per bounce)
// per path, walk-major, then the low and high words of each path's mask.
constexpr int kMaxTapeDepth = 63;

int64_t tape_bytes(const flg_render_cfg* cfg) {
  if (!cfg || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 || cfg->max_depth < 1 ||
      cfg->max_depth > kMaxTapeDepth)
    return -1;
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  if (npaths >= ((int64_t)1 << 31)) return -1;
  return npaths * 4 * ((int64_t)cfg->max_depth + 3);
}

int prepare(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal, const flg_vol_scene* scene,
            const flg_render_cfg* cfg, const flg_diff_params* prm, const void* tape, DiffDev* D,
            LightsDev* L, HaltonTables* tab, HaltonParams* hp, CamDevV* cam) {
  if (!ctx || !vol || !scene || !cfg || !prm || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
      cfg->max_depth < 1)
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: bad arguments");
  if (cfg->schedule < 0 || cfg->schedule > 2)
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: schedule must be 0, 1 or 2");
  if ((int64_t)cfg->width * cfg->height * cfg->spp >= ((int64_t)1 << 31))
    return set_error(ctx, FLG_ERR_RANGE, "differentiable render: too many paths");
  if (proposal && (proposal->nx != vol->nx || proposal->ny != vol->ny || proposal->nz != vol->nz))
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: proposal grid dimensions differ");
  if (!(prm->majorant > 0.0f) || !(prm->albedo > 0.0f))
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: majorant and albedo must be > 0");
  *hp = make_halton_params(cfg->width, cfg->height);
  if ((uint64_t)hp->stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
    return set_error(ctx, FLG_ERR_RANGE, "differentiable render: spp * Halton stride must stay below 2^24");
  int st = get_halton_tables(ctx, tab);
  if (st) return st;
  if ((st = make_lights(ctx, vol, scene, L))) return st;
  D->v = make_voldev(vol);
  D->vp = proposal ? make_voldev(proposal) : D->v;
  D->same = proposal == nullptr || proposal == vol;
  D->s = prm->sigma_t_scale;
  D->alpha = prm->albedo;
  D->g = prm->g;
  D->mu = prm->majorant;
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  D->npaths = (uint32_t)npaths;
  D->tape = nullptr;
  D->mask_lo = D->mask_hi = nullptr;
  if (tape) {
    if (cfg->max_depth > kMaxTapeDepth)
      return set_error(ctx, FLG_ERR_RANGE, "differentiable render: a walk tape needs max_depth <= 63");
    D->tape = (float*)tape;
    D->mask_lo = (uint32_t*)(D->tape + (int64_t)(cfg->max_depth + 1) * npaths);
    D->mask_hi = D->mask_lo + npaths;
  }
  *cam = make_cam_v(&scene->camera, cfg->width, cfg->height);
  return FLG_OK;
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int64_t flg_volpath_diff_tape_bytes(const flg_render_cfg* cfg) { return tape_bytes(cfg); }

int flg_volpath_diff_forward(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                             const flg_vol_scene* scene, const flg_render_cfg* cfg,
                             const flg_diff_params* params, float* d_rgb, float* d_lpath, void* d_tape) {
  DiffDev D;
  LightsDev L;
  HaltonTables tab;
  HaltonParams hp;
  CamDevV cam;
  int st = prepare(ctx, vol, proposal, scene, cfg, params, d_tape, &D, &L, &tab, &hp, &cam);
  if (st) return st;
  if (!d_rgb && !d_lpath) return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_diff_forward: no output");
  const int64_t npix = (int64_t)cfg->width * cfg->height, npaths = npix * cfg->spp;
  float* lpath = d_lpath;
  if (!lpath) {
    lpath = (float*)scratch(ctx, "diff_lpath", (size_t)npaths * 4, &st);
    if (!lpath) return st;
  }
  {
    ScopedTimer tm(ctx, "diff_forward");
    if ((st = launch_diff<false>(ctx, cfg, D, L, tab, hp, cam, lpath, nullptr, nullptr, nullptr))) return st;
    if (d_rgb) k_mean<<<div_up(npix, 256), 256, 0, ctx->stream>>>(lpath, npix, cfg->spp, d_rgb);
  }
  FLG_CHECK_LAUNCH(ctx);
  return FLG_OK;
}

int flg_volpath_adjoint(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                        const flg_vol_scene* scene, const flg_render_cfg* cfg,
                        const flg_diff_params* params, const float* d_lpath, const void* d_tape,
                        const float* d_dl_drgb, float* d_grad, float* d_grad_albedo) {
  DiffDev D;
  LightsDev L;
  HaltonTables tab;
  HaltonParams hp;
  CamDevV cam;
  int st = prepare(ctx, vol, proposal, scene, cfg, params, d_tape, &D, &L, &tab, &hp, &cam);
  if (st) return st;
  if (!d_lpath || !d_dl_drgb || !d_grad || !d_grad_albedo)
    return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_adjoint: null buffer");
  {
    ScopedTimer tm(ctx, "diff_adjoint");
    if ((st = launch_diff<true>(ctx, cfg, D, L, tab, hp, cam, (float*)d_lpath, d_dl_drgb, d_grad,
                                d_grad_albedo)))
      return st;
  }
  FLG_CHECK_LAUNCH(ctx);
  return FLG_OK;
}

}  // extern "C"
