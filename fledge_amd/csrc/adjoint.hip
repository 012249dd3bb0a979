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

#include "diff.hpp"

namespace flg {

namespace {

// The ratio track of walk `key` along (o, d) up to tmax: measured, and its
// derivative deposited.
FLG_D Track box_track_measure(const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax) {
  Track r{1.0f, 0};
  uint32_t n = 0;
  seg_measure(D, key, n, o, d, tmax, r);
  return r;
}

FLG_D void box_track_deposit(const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax, const Track& r, float scale,
                             float* grad) {
  uint32_t n = 0;
  seg_deposit(D, key, n, o, d, tmax, r, scale, grad);
}

// One path, as a start (camera ray, box entry, primary transmittance) and one
// step per scattering event, so that persistent lanes can interleave paths.
// ADJ = false: accumulates the RGB radiance (rgb).  ADJ = true: replays the
// path in the loss's units: the lights' weights wp and we (light_weights),
// acc the weighted radiance emitted so far, L_total the path's weighted
// total, delta = 1 / spp, depositing gradients.
struct DiffPath {
  PathSampler ps;
  V3 o, d, rgb;
  float beta, acc, delta, L_total, wp, we;
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
  s.rgb = v3(0.0f, 0.0f, 0.0f);
  s.path = p;
  s.mask = 0;
  s.delta = s.L_total = s.wp = s.we = 0.0f;
  if (ADJ) {
    const float* g = dl_drgb + 3 * (int64_t)pix;
    light_weights(L, g, &s.wp, &s.we);
    if (s.wp == 0.0f && s.we == 0.0f) return true;
    s.delta = 1.0f / (float)spp;
    const float* l = lpath + 3 * (int64_t)p;
    s.L_total = g[0] * l[0] + g[1] * l[1] + g[2] * l[2];
    if (D.tape) s.mask = ((uint64_t)D.mask_hi[p] << 32) | D.mask_lo[p];
  }
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  V3 o, d;
  gen_ray_v(cam, (float)x + ux, (float)y + uy, &o, &d);
  const V3 Le = L.env.present ? L.env.radiance : v3(0.0f, 0.0f, 0.0f);
  float t_min, t_max;
  if (!box_enter(D.v.bmin, D.v.bmax, o, d, &t_min, &t_max)) {
    s.rgb = Le;
    return true;
  }
  const uint32_t key = walk_key(p, 0);
  const Track tr = track<ADJ>(D, p, 0, s.mask, [&] { return box_track_measure(D, key, o, d, kInf); });
  if (ADJ && track_live(tr) && s.delta * s.we != 0.0f)
    box_track_deposit(D, key, o, d, kInf, tr, s.delta * s.we, grad);
  if (ADJ) s.acc += track_value(tr) * s.we;
  else add_scaled(s.rgb, track_value(tr), Le);
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
  V3 pos;
  const bool hit =
      diff_delta_track<ADJ>(D, tab, hp, s.ps, s.o, s.d, kInf, s.beta, s.acc, s.delta, s.L_total, grad, galpha, &pos);
  if (!hit) return true;
  const V3 wo = -s.d;
  const float u = s.ps.next(tab, hp);
  const float vv = s.ps.next(tab, hp);
  const V3 wi = hg_sample(wo, u, vv, D.g);
  const NeeSample ns = nee_sample(D, L, tab, hp, s.ps, pos, wo);
  if (ns.valid) {
    const V3 diff = ns.p_light - pos;
    const float nrm = sqrtf(sqnorm(diff));
    const V3 dd = diff / nrm;
    const uint32_t w = (uint32_t)s.bounces;
    const uint32_t key = walk_key(s.path, w);
    const Track tr =
        track<ADJ>(D, s.path, w, s.mask, [&] { return box_track_measure(D, key, pos, dd, nrm - SHADOW_EPS); });
    if (ADJ) {
      const float K = nee_k(ns, nee_li_w(ns, s.wp, s.we));
      if (track_live(tr) && s.delta * s.beta * K != 0.0f)
        box_track_deposit(D, key, pos, dd, nrm - SHADOW_EPS, tr, s.delta * s.beta * K, grad);
      s.acc += s.beta * K * track_value(tr);
    } else {
      add_nee(s.rgb, L, ns, s.beta, track_value(tr));
    }
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

// Forward: the path's RGB radiance and its tape bits.
FLG_D void diff_finish(const DiffDev& D, uint32_t p, const DiffPath& s, float* lpath) {
  float* o = lpath + 3 * (int64_t)p;
  o[0] = s.rgb.x;
  o[1] = s.rgb.y;
  o[2] = s.rgb.z;
  if (D.tape) {
    D.mask_lo[p] = (uint32_t)s.mask;
    D.mask_hi[p] = (uint32_t)(s.mask >> 32);
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
  regen_loop<DiffPath, kVolChunk>(
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

int64_t tape_bytes(const flg_render_cfg* cfg) {
  if (!cfg || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 || cfg->max_depth < 1 ||
      cfg->max_depth > kMaxTapeDepth)
    return -1;
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  if (npaths >= ((int64_t)1 << 31)) return -1;
  return npaths * 4 * ((int64_t)cfg->max_depth + 3);
}

// Active bricks: one block per brick, active iff one of its voxels is
// non-zero (flag written into the slot array).
__global__ __launch_bounds__(256) void k_brick_flags(const float* rho, int32_t nx, int32_t ny, int32_t nz,
                                                     int32_t b, int32_t nbx, int32_t nby, int32_t* slot) {
  const int64_t br = blockIdx.x;
  const int bx = (int)(br % nbx), by = (int)((br / nbx) % nby), bz = (int)(br / ((int64_t)nbx * nby));
  const int nvox = b * b * b;
  int any = 0;
  for (int v = threadIdx.x; v < nvox; v += blockDim.x) {
    const int ix = bx * b + v % b, iy = by * b + (v / b) % b, iz = bz * b + v / (b * b);
    if (ix < nx && iy < ny && iz < nz && rho[((int64_t)iz * ny + iy) * nx + ix] != 0.0f) any = 1;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) slot[br] = any;
}

// Flags to slots in brick order (-1 inactive), in one 1024-lane block; the
// number of active bricks into *count.
__global__ __launch_bounds__(1024) void k_brick_slots(int32_t* slot, int64_t nb, int32_t* count) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t base;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (int64_t c = 0; c < nb; c += 1024) {
    const int64_t i = c + threadIdx.x;
    const int f = i < nb ? slot[i] : 0;
    const uint64_t m = __ballot(f != 0);
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wid] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wid; ++w) off += wsum[w];
    if (i < nb) slot[i] = f ? off + rank : -1;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < 16; ++w) tot += wsum[w];
      base += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = base;
}

}  // namespace

int diff_prepare(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal, const flg_vol_scene* scene,
                 const flg_render_cfg* cfg, const flg_diff_params* prm, const void* tape, const float* bounds,
                 DiffDev* D, LightsDev* L, HaltonTables* tab, HaltonParams* hp, CamDevV* cam) {
  if (!ctx || !vol || !scene || !cfg || !prm || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
      cfg->max_depth < 1)
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: bad arguments");
  if (cfg->schedule < 0 || cfg->schedule > 2)
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: schedule must be 0, 1 or 2");
  if (cfg->film != 0) return set_error(ctx, FLG_ERR_INVALID, "differentiable render: film must be 0");
  if ((int64_t)cfg->width * cfg->height * cfg->spp >= ((int64_t)1 << 31))
    return set_error(ctx, FLG_ERR_RANGE, "differentiable render: too many paths");
  if (proposal && (proposal->nx != vol->nx || proposal->ny != vol->ny || proposal->nz != vol->nz))
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: proposal grid dimensions differ");
  if (!(prm->majorant > 0.0f) || !(prm->albedo > 0.0f))
    return set_error(ctx, FLG_ERR_INVALID, "differentiable render: majorant and albedo must be > 0");
  *hp = make_halton_params(cfg->width, cfg->height);
  if ((uint64_t)hp->stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
    return set_error(ctx, FLG_ERR_RANGE, "differentiable render: spp * Halton stride must stay at or below 2^32");
  int st = get_halton_tables(ctx, tab);
  if (st) return st;
  if ((st = make_lights(ctx, vol, scene, L, bounds))) return st;
  const float* b = bounds ? bounds : vol->bounds;
  D->v = voldev_of(vol->grid, vol->bmap, vol->nx, vol->ny, vol->nz, b, vol->medium, vol->max_density);
  D->vp = proposal ? voldev_of(proposal->grid, proposal->bmap, proposal->nx, proposal->ny, proposal->nz, b,
                               proposal->medium, proposal->max_density)
                   : D->v;
  D->same = proposal == nullptr || proposal == vol;
  D->s = prm->sigma_t_scale;
  D->alpha = prm->albedo;
  D->g = prm->g;
  D->mu = prm->majorant;
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  D->npaths = (uint32_t)npaths;
  D->tape = nullptr;
  D->mask_lo = D->mask_hi = nullptr;
  D->brick_slot = nullptr;
  D->brick_shift = D->nbx = D->nby = 0;
  D->nlook = prm->d_lookups;
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


int diff_set_bricks(flg_ctx* ctx, const flg_brick_layout* bl, DiffDev* D) {
  if (!bl) return FLG_OK;
  int sh = 0;
  while (sh <= 5 && (1 << sh) != bl->brick) ++sh;
  if (sh > 5 || !bl->d_slot || bl->n_active < 0)
    return set_error(ctx, FLG_ERR_INVALID, "gradient bricks: brick must be 1, 2, 4, 8, 16 or 32 voxels, with a slot map");
  if (bl->nx != D->v.nx || bl->ny != D->v.ny || bl->nz != D->v.nz)
    return set_error(ctx, FLG_ERR_INVALID, "gradient bricks: the layout was built for a grid of other dimensions");
  D->brick_slot = bl->d_slot;
  D->brick_shift = sh;
  D->nbx = (D->v.nx + bl->brick - 1) >> sh;
  D->nby = (D->v.ny + bl->brick - 1) >> sh;
  return FLG_OK;
}

}  // namespace flg

using namespace flg;

extern "C" {

int64_t flg_volpath_diff_tape_bytes(const flg_render_cfg* cfg) { return tape_bytes(cfg); }

int flg_bricks_build(flg_ctx* ctx, const float* d_density, int32_t nx, int32_t ny, int32_t nz, int32_t brick,
                     int32_t* d_slot, int32_t* h_n_active) {
  return host_guard(ctx, "flg_bricks_build", [&]() -> int {
    if (!ctx || !d_density || !d_slot || !h_n_active || nx <= 0 || ny <= 0 || nz <= 0)
      return set_error(ctx, FLG_ERR_INVALID, "flg_bricks_build: bad arguments");
    if (brick != 1 && brick != 2 && brick != 4 && brick != 8 && brick != 16 && brick != 32)
      return set_error(ctx, FLG_ERR_INVALID, "flg_bricks_build: brick must be 1, 2, 4, 8, 16 or 32");
    const int32_t nbx = (nx + brick - 1) / brick, nby = (ny + brick - 1) / brick, nbz = (nz + brick - 1) / brick;
    const int64_t nb = (int64_t)nbx * nby * nbz;
    if (nb >= ((int64_t)1 << 31)) return set_error(ctx, FLG_ERR_RANGE, "flg_bricks_build: too many bricks");
    int st = FLG_OK;
    int32_t* count = (int32_t*)scratch(ctx, "brick_count", 256, &st);
    if (!count) return st;
    k_brick_flags<<<(unsigned)nb, 256, 0, ctx->stream>>>(d_density, nx, ny, nz, brick, nbx, nby, d_slot);
    k_brick_slots<<<1, 1024, 0, ctx->stream>>>(d_slot, nb, count);
    FLG_CHECK_LAUNCH(ctx);
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(h_n_active, count, 4, hipMemcpyDeviceToHost, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FLG_OK;
  });
}

int flg_volpath_diff_forward(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                             const flg_vol_scene* scene, const flg_render_cfg* cfg,
                             const flg_diff_params* params, float* d_rgb, float* d_lpath, void* d_tape) {
  return host_guard(ctx, "flg_volpath_diff_forward", [&]() -> int {
    DiffDev D;
    LightsDev L;
    HaltonTables tab;
    HaltonParams hp;
    CamDevV cam;
    int st = diff_prepare(ctx, vol, proposal, scene, cfg, params, d_tape, nullptr, &D, &L, &tab, &hp, &cam);
    if (st) return st;
    if (!d_rgb && !d_lpath) return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_diff_forward: no output");
    const int64_t npix = (int64_t)cfg->width * cfg->height, npaths = npix * cfg->spp;
    float* lpath = d_lpath;
    if (!lpath) {
      lpath = (float*)scratch(ctx, "diff_lpath", (size_t)npaths * 12, &st);
      if (!lpath) return st;
    }
    {
      ScopedTimer tm(ctx, "diff_forward");
      if ((st = launch_diff<false>(ctx, cfg, D, L, tab, hp, cam, lpath, nullptr, nullptr, nullptr))) return st;
      if (d_rgb) pixel_mean(ctx->stream, lpath, npix, cfg->spp, d_rgb);
    }
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

int flg_volpath_adjoint(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                        const flg_vol_scene* scene, const flg_render_cfg* cfg,
                        const flg_diff_params* params, const float* d_lpath, const void* d_tape,
                        const float* d_dl_drgb, float* d_grad, float* d_grad_albedo) {
  return host_guard(ctx, "flg_volpath_adjoint", [&]() -> int {
    DiffDev D;
    LightsDev L;
    HaltonTables tab;
    HaltonParams hp;
    CamDevV cam;
    int st = diff_prepare(ctx, vol, proposal, scene, cfg, params, d_tape, nullptr, &D, &L, &tab, &hp, &cam);
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
  });
}

}  // extern "C"
