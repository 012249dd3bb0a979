// Surface + medium coupling: VolPathIntegrator::Li (src/integrator.cpp:
// 407-510) for a scene of one closed triangle mesh whose surfaces are
// index-matched interfaces (no material) and whose inside holds the dense
// density volume, lit by the point and environment lights.
//
// Per path, as in the reference: closest hit of the ray on the mesh; the
// environment added at bounce 0 when the ray escapes; inside the medium, delta
// tracking (OpenVDBVolume::sample) up to the hit; on a medium event, next-event
// estimation with the shadow transmittance of VolEstimateTr
// (integrator.cpp:40-64: ratio tracking per segment inside the medium,
// interfaces crossed, nothing applied after the last crossing), then HG
// sampling.  The reference dereferences the missing material at an interface
// (integrator.cpp:433,487,501); here an interface is crossed as
// VolEstimateTr crosses it: SInteraction::SpawnRay in the same direction
// (interaction.cpp:16-26: origin offset along the geometric normal, the medium
// entered when the direction faces against the normal), weight unchanged and
// not counted as a bounce.  The medium fills the mesh's bounding box
// (ShapePrimitive's ctor, primitive.cpp:17-24: m_volume->setBound(getBound())).
#include <algorithm>
#include <cmath>

#include "traverse.hpp"
#include "volume.hpp"

namespace flg {

namespace {

// Interface crossings of one path (or one shadow ray) at most; a closed mesh
// is crossed twice per passage, so this only guards a ray that keeps
// re-hitting its own surface.
constexpr int kMaxCrossings = 256;

// Closest hit on the mesh within tmax: sorted position (-1 on a miss), t, and
// the geometric normal of the hit as the Pluecker test computes it.
template <bool COMPAT>
FLG_D int32_t mesh_hit(const TraceGeom& g, V3 o, V3 d, float tmax, float* t, V3* ng) {
  float best = tmax;
  Counters c{0, 0};
  const int32_t p = trace_one<COMPAT, false>(g, o, d, best, c);
  *t = best;
  if (p >= 0) {
    const BTri tri = g.tris[p];
    const V3 v0 = v3(tri.v0.x, tri.v0.y, tri.v0.z) - o, v1 = v3(tri.v1.x, tri.v1.y, tri.v1.z) - o,
             v2 = v3(tri.v2.x, tri.v2.y, tri.v2.z) - o;
    *ng = stable_normalize(stable_triangle_normal(v2 - v0, v0 - v1, v1 - v2), COMPAT);
  }
  return p;
}

// VolEstimateTr (integrator.cpp:40-64) along (o, d) up to tmax, starting
// inside the medium or not.
template <bool COMPAT, bool COUNT>
FLG_D float shadow_tr(const TraceGeom& g, const VolDev& v, HaltonDraw& rng, V3 o, V3 d, float tmax,
                      bool inside, uint32_t& nlook) {
  float tr = 1.0f;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t;
    V3 ng;
    if (mesh_hit<COMPAT>(g, o, d, t_max, &t, &ng) < 0) return tr;
    if (inside) tr *= vol_tr<COUNT>(v, rng, o, d, t, nlook);
    t_max -= t;
    if (t_max < 0.0f) return tr;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);
    d = normalize(d);
    inside = dot(d, ng) < 0.0f;
  }
  return tr;
}

struct VPath {
  PathSampler ps;
  V3 o, d;
  float L, beta;
  int bounces, crossings;
  bool inside;
};

// Camera ray of path p (pixel p / spp, sample p % spp) on the per-sample
// Halton schedule; the camera is outside the mesh.
FLG_D void vp_begin(const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, int32_t spp,
                    uint32_t p, VPath& s) {
  const uint32_t pix = p / (uint32_t)spp, smp = p - pix * (uint32_t)spp;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  gen_ray_v(cam, (float)x + ux, (float)y + uy, &s.o, &s.d);
  s.L = 0.0f;
  s.beta = 1.0f;
  s.bounces = 0;
  s.crossings = 0;
  s.inside = false;
}

// One trip of the reference's loop; returns true when the path is complete.
template <bool COMPAT, bool COUNT>
FLG_D bool vp_step(const TraceGeom& g, const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                   const HaltonParams& hp, int max_depth, VPath& s, uint32_t& nlook) {
  float t;
  V3 ng;
  const bool hit = mesh_hit<COMPAT>(g, s.o, s.d, kInf, &t, &ng) >= 0;
  if (s.bounces == 0 && !hit && L.env.present) s.L += s.beta * L.env.radiance;
  if (!hit || s.bounces >= max_depth) return true;
  bool success = false;
  V3 pos;
  if (s.inside) {
    HaltonDraw rng{s.ps, tab, hp};
    s.beta *= vol_sample<COUNT>(v, rng, s.o, s.d, t, &success, &pos, nlook);
  }
  if (success) {
    const V3 wo = -s.d;
    s.L += s.beta * sample_one_light(L, v.g, tab, hp, s.ps, pos, wo, [&](V3 p, V3 d, float tmax) {
      HaltonDraw rng{s.ps, tab, hp};
      return shadow_tr<COMPAT, COUNT>(g, v, rng, p, d, tmax, true, nlook);
    });
    const float u = s.ps.next(tab, hp);
    const float vv = s.ps.next(tab, hp);
    const V3 wi = hg_sample(wo, u, vv, v.g);
    if (s.bounces >= max_depth || s.beta * s.beta + s.beta * s.beta + s.beta * s.beta == 0.0f) return true;
    s.o = pos;  // VInteraction::SpawnRay: no normal, no offset
    s.d = normalize(wi);
    ++s.bounces;
    return false;
  }
  // an interface: cross it
  if (++s.crossings > kMaxCrossings) return true;
  const V3 ph = s.o + s.d * t;
  s.o = offset_ray_origin(ph, ng, s.d);
  s.d = normalize(s.d);
  s.inside = dot(s.d, ng) < 0.0f;
  return false;
}

template <bool COUNT>
FLG_D void vp_count(uint32_t nlook, unsigned long long* nlookups) {
  if (COUNT) {
    unsigned long long c = nlook;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(nlookups, c);
  }
}

// One lane per path sample.
template <bool COMPAT, bool COUNT>
__global__ __launch_bounds__(256) void k_volpath(TraceGeom g, VolDev v, LightsDev L, HaltonTables tab_g,
                                                 HaltonParams hp, CamDevV cam, int32_t spp, int32_t max_depth,
                                                 float* lpath, unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npaths = (int64_t)cam.w * cam.h * spp;
  uint32_t nlook = 0;
  if (p < npaths) {
    VPath s;
    vp_begin(tab, hp, cam, spp, (uint32_t)p, s);
    while (!vp_step<COMPAT, COUNT>(g, v, L, tab, hp, max_depth, s, nlook)) {
    }
    lpath[p] = s.L;
  }
  vp_count<COUNT>(nlook, nlookups);
}

// Persistent lanes with path regeneration (regen_loop); same per-path code.
template <bool COMPAT, bool COUNT>
__global__ __launch_bounds__(256) void k_volpath_regen(TraceGeom g, VolDev v, LightsDev L, HaltonTables tab_g,
                                                       HaltonParams hp, CamDevV cam, int32_t spp,
                                                       int32_t max_depth, float* lpath, unsigned int* next_path,
                                                       unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t npaths = (uint32_t)((int64_t)cam.w * cam.h * spp);
  uint32_t nlook = 0;
  regen_loop<VPath>(
      npaths, next_path,
      [&](uint32_t p, VPath& s) {
        vp_begin(tab, hp, cam, spp, p, s);
        return false;
      },
      [&](VPath& s) { return vp_step<COMPAT, COUNT>(g, v, L, tab, hp, max_depth, s, nlook); },
      [&](uint32_t p, const VPath& s) { lpath[p] = s.L; });
  vp_count<COUNT>(nlook, nlookups);
}

__global__ void k_vp_zero(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) void k_vp_mean(const float* lpath, int64_t npix, int32_t spp, float* rgb) {
  const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= npix) return;
  float c = 0.0f;
  const float* l = lpath + pix * spp;
  for (int s = 0; s < spp; ++s) c += l[s];
  c = c / (float)spp;
  rgb[3 * pix] = c;
  rgb[3 * pix + 1] = c;
  rgb[3 * pix + 2] = c;
}

template <bool COMPAT, bool COUNT>
void launch_volpath(int sched, int blocks1, int blocks2, hipStream_t st, const TraceGeom& g, const VolDev& v,
                    const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam,
                    int32_t spp, int32_t D, float* lpath, unsigned int* next_path, unsigned long long* nl) {
  if (sched == 1)
    k_volpath<COMPAT, COUNT><<<blocks1, 256, 0, st>>>(g, v, L, tab, hp, cam, spp, D, lpath, nl);
  else
    k_volpath_regen<COMPAT, COUNT><<<blocks2, 256, 0, st>>>(g, v, L, tab, hp, cam, spp, D, lpath, next_path, nl);
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_volpath_mesh_render(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol, const flg_vol_scene* scene,
                            const flg_render_cfg* cfg, float* d_rgb, uint64_t* h_stats) {
  if (!ctx || !bvh || !vol || !scene || !cfg || !d_rgb || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
      cfg->max_depth < 1)
    return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: bad arguments");
  if (cfg->film != 0) return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: film must be 0");
  if ((int64_t)cfg->width * cfg->height * cfg->spp >= ((int64_t)1 << 31))
    return set_error(ctx, FLG_ERR_RANGE, "flg_volpath_mesh_render: too many paths");
  const int sched = cfg->schedule ? cfg->schedule : 2;
  if (sched < 1 || sched > 2)
    return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: schedule must be 0, 1 or 2");
  if (bvh->n <= 0) return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: empty mesh");
  const float* mb = bvh->bound;
  if (!(mb[3] > mb[0] && mb[4] > mb[1] && mb[5] > mb[2]))
    return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: the mesh's bounding box is flat");
  HaltonParams hp = make_halton_params(cfg->width, cfg->height);
  if ((uint64_t)hp.stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
    return set_error(ctx, FLG_ERR_RANGE, "flg_volpath_mesh_render: spp * Halton stride must stay below 2^24");
  const int64_t npix = (int64_t)cfg->width * cfg->height, npaths = npix * cfg->spp;
  int st = FLG_OK;
  char* sc = (char*)scratch(ctx, "volpath_mesh", (size_t)npaths * 4 + 256, &st);
  if (!sc) return st;
  unsigned long long* nl = (unsigned long long*)sc;
  unsigned int* next_path = (unsigned int*)(nl + 1);  // cleared with the lookup count
  float* lpath = (float*)(sc + 256);
  HaltonTables tab;
  if ((st = get_halton_tables(ctx, &tab))) return st;
  LightsDev L;
  if ((st = make_lights(ctx, vol, scene, &L, mb))) return st;
  // the medium on the mesh's bounding box (setBound)
  const VolDev v = voldev_of(vol->grid, vol->nx, vol->ny, vol->nz, mb, vol->medium, vol->max_density);
  const CamDevV cam = make_cam_v(&scene->camera, cfg->width, cfg->height);
  const TraceGeom g = geom_of(bvh);
  const bool compat = bvh->flags & FLG_NORMAL_COMPAT;
  const bool count = h_stats != nullptr;
  const int blocks1 = (int)div_up(npaths, 256);
  const int blocks2 = (int)std::min<int64_t>(div_up(npaths, 256), (int64_t)ctx->num_cus * 8);
  {
    ScopedTimer tm(ctx, "volpath_mesh_paths");
    k_vp_zero<<<1, 64, 0, ctx->stream>>>(nl);
    if (compat && count)
      launch_volpath<true, true>(sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, cfg->spp,
                                 cfg->max_depth, lpath, next_path, nl);
    else if (compat)
      launch_volpath<true, false>(sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, cfg->spp,
                                  cfg->max_depth, lpath, next_path, nl);
    else if (count)
      launch_volpath<false, true>(sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, cfg->spp,
                                  cfg->max_depth, lpath, next_path, nl);
    else
      launch_volpath<false, false>(sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, cfg->spp,
                                   cfg->max_depth, lpath, next_path, nl);
  }
  k_vp_mean<<<div_up(npix, 256), 256, 0, ctx->stream>>>(lpath, npix, cfg->spp, d_rgb);
  FLG_CHECK_LAUNCH(ctx);
  if (h_stats) {
    unsigned long long c = 0;
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(&c, nl, 8, hipMemcpyDeviceToHost, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    h_stats[0] = (uint64_t)npaths;
    h_stats[1] = c;
  }
  return FLG_OK;
}

}  // extern "C"
