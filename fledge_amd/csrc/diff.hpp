// Pieces of the differentiable volume estimator shared by the volume-only
// integrator (adjoint.hip) and the one over an interface mesh (volpath.hip):
// the evaluated and proposal grids, the gradient layout, ratio tracks on
// counter-based walk streams, differential delta tracking, the light sample.
//
// No counterpart in the reference; the closest code is OpenVDBVolume::tr and
// ::sample (src/volume.cpp:75-148).  Every branch is independent of the
// evaluated grid theta:
//   * fixed majorant mu; collision decisions on the proposal grid theta_p
//     (real iff u > 1 - s rho_p / mu) with weight alpha rho / rho_p on a real
//     collision and (1 - s rho / mu) / (1 - s rho_p / mu) on a null one;
//   * ratio tracking without roulette, factor 1 - s rho / mu per tentative
//     collision, on the walk's own stream walk_draw(key, n).
#pragma once
#include "volume.hpp"

namespace flg {

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
  // gradient layout: NULL = dense (nx*ny*nz, x fastest); otherwise the slot
  // of each brick of 2^brick_shift voxels per edge (bricks x fastest, nbx*nby
  // per layer), -1 for a brick without a gradient, and the gradient brick
  // after brick in slot order, each x fastest
  const int32_t* brick_slot;
  int brick_shift, nbx, nby;
  // NULL, or a counter of the grid lookups (fetches of 8 voxels) made
  unsigned long long* nlook;
};

FLG_D void count_lookups(const DiffDev& D, unsigned n) {
  if (D.nlook) atomicAdd(D.nlook, (unsigned long long)n);
}

struct Look2 {
  Tri8 t;     // lookup on theta (weights, base, rho)
  float rho_p;
};

FLG_D Look2 look2(const DiffDev& D, V3 p) {
  Look2 r;
  r.t = lookup(D.v, p);
  r.rho_p = D.same ? r.t.rho : lookup(D.vp, p).rho;
  count_lookups(D, D.same ? 1u : 2u);
  return r;
}

// gradient scatter: g * trilinear weight into the caller's gradient
FLG_D void deposit(const DiffDev& D, float* grad, const Tri8& t, float g) {
  if (!t.inside || g == 0.0f) return;
  const int i = t.i, j = t.j, k = t.k;
  const float wx[2] = {1.0f - t.fx, t.fx}, wy[2] = {1.0f - t.fy, t.fy}, wz[2] = {1.0f - t.fz, t.fz};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    const int ii = i + dx, jj = j + dy, kk = k + dz;
    if (ii < 0 || jj < 0 || kk < 0 || ii >= D.v.nx || jj >= D.v.ny || kk >= D.v.nz) continue;
    int64_t at;
    if (D.brick_slot) {
      const int sh = D.brick_shift, m = (1 << sh) - 1;
      const int slot = D.brick_slot[((int64_t)(kk >> sh) * D.nby + (jj >> sh)) * D.nbx + (ii >> sh)];
      if (slot < 0) continue;
      at = ((int64_t)slot << (3 * sh)) + (((kk & m) << (2 * sh)) | ((jj & m) << sh) | (ii & m));
    } else {
      at = ((int64_t)kk * D.v.ny + jj) * D.v.nx + ii;
    }
    const float w = wx[dx] * wy[dy] * wz[dz];
    unsafeAtomicAdd(grad + at, g * w);
  }
}

// A ratio-tracking product: nonzero = product of the non-zero factors,
// zeros = number of factors equal to 0.
struct Track {
  float nonzero;
  int zeros;
};

// The product's value.
FLG_D float track_value(const Track& r) { return r.zeros > 0 ? 0.0f : r.nonzero; }

// Whether a measured track has a non-zero derivative.
FLG_D bool track_live(const Track& r) { return r.zeros < 2 && r.nonzero != 0.0f; }

// First walk of a ratio track over one segment: the medium's part of (o, d)
// up to tmax, draws n, n + 1, ... of the walk's stream (n advances).  Returns
// false once the product and every derivative are fixed (two zero factors,
// or a product of the non-zero factors that reached 0).
FLG_D bool seg_measure(const DiffDev& D, uint32_t key, uint32_t& n, V3 o, V3 d, float tmax, Track& r) {
  float t_min, t_max;
  if (!box_pbrt(D.v.bmin, D.v.bmax, o, d, tmax, &t_min, &t_max)) return true;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  const int cap = walk_cap(D.mu, t, t_max);
  for (int step = 0; step < cap; ++step) {
    t += -flg_logf(walk_draw(key, n++)) / D.mu;
    if (t >= t_max) break;
    const float f = 1.0f - D.s * lookup(D.v, o + d * t).rho / D.mu;
    count_lookups(D, 1u);
    if (f == 0.0f) {
      if (++r.zeros >= 2) return false;
    } else {
      r.nonzero *= f;
      if (r.nonzero == 0.0f) return false;
    }
  }
  return true;
}

// Second walk of a segment (same draws): deposits scale * d(product)/d(rho_i)
// for every lookup i, d/d(rho_i) = k prod_{j != i} f_j with k = -s / mu.
// Returns false once the track's derivative is complete (its one zero factor
// has been deposited).
FLG_D bool seg_deposit(const DiffDev& D, uint32_t key, uint32_t& n, V3 o, V3 d, float tmax, const Track& r,
                       float scale, float* grad) {
  float t_min, t_max;
  if (!box_pbrt(D.v.bmin, D.v.bmax, o, d, tmax, &t_min, &t_max)) return true;
  float t = smax(t_min, 0.0f);
  t_max = smin(t_max, tmax);
  const float k = -D.s / D.mu;
  const int cap = walk_cap(D.mu, t, t_max);
  for (int step = 0; step < cap; ++step) {
    t += -flg_logf(walk_draw(key, n++)) / D.mu;
    if (t >= t_max) break;
    const Tri8 q = lookup(D.v, o + d * t);
    count_lookups(D, 1u);
    const float f = 1.0f - D.s * q.rho / D.mu;
    if (r.zeros == 0) {
      deposit(D, grad, q, scale * k * (r.nonzero / f));
    } else if (f == 0.0f) {  // the one zero factor: the others' product
      deposit(D, grad, q, scale * k * r.nonzero);
      return false;
    }
  }
  return true;
}

// Walk w of path p: measured by `measure` (and recorded on the tape by a
// forward that has one), or read back from the tape by the adjoint.
template <bool ADJ, class Measure>
FLG_D Track track(const DiffDev& D, uint32_t p, uint32_t w, uint64_t& mask, Measure measure) {
  if (ADJ && D.tape) {
    return Track{D.tape[(int64_t)w * D.npaths + p], (int)((mask >> w) & 1u)};
  }
  const Track r = measure();
  if (!ADJ && D.tape) {
    D.tape[(int64_t)w * D.npaths + p] = r.zeros >= 2 ? 0.0f : r.nonzero;
    if (r.zeros == 1) mask |= (uint64_t)1 << w;
  }
  return r;
}

// Light choice and sample (UniformSampleOneLight / EstimateDirect,
// integrator.cpp:69-120) without the shadow transmittance.  valid: pdf != 0,
// Li != 0 in some channel and f != 0; then the estimate of one channel is
// nee_k(n, Li) times the shadow transmittance, Li = light_li(L, n.d2) per
// channel.  The adjoint prices the estimate in the loss's units instead:
// nee_k with Li = sum_c dLoss/dL_c Li_c (nee_li_w), the linear combination of
// the channels the gradient needs.
struct NeeSample {
  V3 p_light;
  float f, d2, pdf, light_pdf;  // d2 < 0: the environment light
  bool valid;
};

FLG_D NeeSample nee_sample(const DiffDev& D, const LightsDev& L, const HaltonTables& tab,
                           const HaltonParams& hp, PathSampler& ps, V3 p, V3 wo) {
  NeeSample n;
  n.valid = false;
  const int n_lights = (L.has_point ? 1 : 0) + (L.env.present ? 1 : 0);
  if (n_lights == 0) return n;
  int light_num = (int)(ps.next(tab, hp) * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  n.light_pdf = (float)(1.0 / (double)n_lights);
  const bool use_point = L.has_point && light_num == 0;
  float u0 = ps.next(tab, hp);
  float u1 = ps.next(tab, hp);
  V3 wi;
  if (use_point) {
    V3 diff = L.point_pos - p;
    n.d2 = sqnorm(diff);
    wi = normalize(diff);
    n.pdf = 1.0f;
    n.p_light = L.point_pos;
  } else {
    env_sample_dir(L.env, p, u0, u1, &wi, &n.pdf, &n.p_light);
    n.d2 = -1.0f;
  }
  const V3 Li = light_li(L, n.d2);
  if (n.pdf == 0.0f || (Li.x == 0.0f && Li.y == 0.0f && Li.z == 0.0f)) return n;
  n.f = hg_p(wi, wo, D.g);
  n.valid = n.f != 0.0f;
  return n;
}

// f Li / pdf / light_pdf for one channel's Li.
FLG_D float nee_k(const NeeSample& n, float Li) { return n.f * Li / n.pdf / n.light_pdf; }

// The sampled light's Li in the loss's units: wp = sum_c g_c I_c for the
// point light (over d2), we = sum_c g_c E_c for the environment.
FLG_D float nee_li_w(const NeeSample& n, float wp, float we) { return n.d2 < 0.0f ? we : wp / n.d2; }

// The adjoint's per-pixel weights of the two lights: g = dLoss/dL of the
// pixel (RGB).
FLG_D void light_weights(const LightsDev& L, const float* g, float* wp, float* we) {
  const V3 I = L.point_intensity, E = L.env.radiance;
  *wp = L.has_point ? g[0] * I.x + g[1] * I.y + g[2] * I.z : 0.0f;
  *we = L.env.present ? g[0] * E.x + g[1] * E.y + g[2] * E.z : 0.0f;
}

// acc += beta * K * tr per channel, K the channel's nee_k.
FLG_D void add_nee(V3& acc, const LightsDev& L, const NeeSample& n, float beta, float tr) {
  const V3 Li = light_li(L, n.d2);
  acc.x += beta * nee_k(n, Li.x) * tr;
  acc.y += beta * nee_k(n, Li.y) * tr;
  acc.z += beta * nee_k(n, Li.z) * tr;
}

// Differential delta tracking along (o, d) within the medium and [0,
// ray_tmax], on the path's Halton draws.  Multiplies the collision weights
// into beta; the adjoint (ADJ) deposits each event's derivative, weighted by
// the radiance the path gathers after it (L_total - acc).  Returns true on a
// real collision, at *pos.
template <bool ADJ>
FLG_D bool diff_delta_track(const DiffDev& D, const HaltonTables& tab, const HaltonParams& hp, PathSampler& ps,
                            V3 o, V3 d, float ray_tmax, float& beta, float acc, float delta, float L_total,
                            float* grad, float* galpha, V3* pos) {
  float tb0, tb1;
  if (!box_pbrt(D.v.bmin, D.v.bmax, o, d, ray_tmax, &tb0, &tb1) || !(D.mu > 0.0f)) return false;
  const float k = -D.s / D.mu;
  float t = smax(tb0, 0.0f);
  tb1 = smin(tb1, ray_tmax);
  const int cap = walk_cap(D.mu, t, tb1);
  for (int step = 0; step < cap; ++step) {
    const float u = ps.next(tab, hp);
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / D.mu;
    if (t >= tb1) break;
    const V3 p = o + d * t;
    const Look2 q = look2(D, p);
    const float p_n = 1.0f - D.s * q.rho_p / D.mu;
    const bool real = ps.next(tab, hp) > p_n;
    if (real) {
      if (ADJ) {
        const float S = L_total - acc;
        if (q.t.rho != 0.0f) deposit(D, grad, q.t, delta * S / q.t.rho);
        *galpha += delta * S / D.alpha;
      }
      beta *= D.alpha * (q.t.rho / q.rho_p);
      *pos = p;
      return true;
    }
    const float fnull = 1.0f - D.s * q.t.rho / D.mu;
    if (ADJ && fnull != 0.0f) deposit(D, grad, q.t, delta * (L_total - acc) * k / fnull);
    // a null collision has probability p_n > 0 except for u == 0 exactly;
    // that measure-zero case keeps weight 1
    if (p_n != 0.0f) beta *= fnull / p_n;
  }
  return false;
}

template <bool ADJ>
FLG_D void reduce_galpha(float ga, float* galpha) {
  if (ADJ) {
    for (int off = 32; off >= 1; off >>= 1) ga += __shfl_xor(ga, off, 64);
    if ((threadIdx.x & 63) == 0 && ga != 0.0f) unsafeAtomicAdd(galpha, ga);
  }
}

// Walk tape: one float per walk (walk 0 and one walk per bounce) per path,
// walk-major, then the low and high words of each path's mask.
constexpr int kMaxTapeDepth = 63;

// Checks the arguments of a differentiable render and sets up its device
// description: medium on `bounds` (NULL: the volume's box), lights, Halton
// tables, camera, tape.  Dense gradient layout.
int diff_prepare(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal, const flg_vol_scene* scene,
                 const flg_render_cfg* cfg, const flg_diff_params* prm, const void* tape, const float* bounds,
                 DiffDev* D, LightsDev* L, HaltonTables* tab, HaltonParams* hp, CamDevV* cam);

// Gradient brick layout (NULL: keeps the dense layout).
int diff_set_bricks(flg_ctx* ctx, const flg_brick_layout* bl, DiffDev* D);

}  // namespace flg
