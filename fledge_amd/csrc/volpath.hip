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

#include "diff.hpp"
#include "traverse.hpp"

namespace flg {

namespace {

// Interface crossings of one path (or one shadow ray) at most; a closed mesh
// is crossed twice per passage, so this only guards a ray that keeps
// re-hitting its own surface.
constexpr int kMaxCrossings = 256;

// Every kernel over the mesh runs 256-lane blocks and keeps the 8 most recent
// traversal stack entries of each lane in LDS (the ring stack of the path
// tracer's traversal), the older ones in private memory.
constexpr int kMeshBlock = 256;
constexpr int kMeshRing = 8;
// The kernels ask for 6 waves per SIMD: their traversals are latency-bound,
// and the register cap (the few values beyond it spill to scratch) measured
// faster than the 3-4 waves of the unconstrained allocation.
constexpr int kMeshWaves = 6;

struct MeshGeom {
  TraceGeom g;
  uint2* ring;  // the block's LDS ring, kMeshBlock * kMeshRing entries
};

// Counts of a measurement pass: density lookups, rays traced on the mesh,
// and their node visits and triangle tests.
struct VpCount {
  uint32_t look = 0, rays = 0, nodes = 0, tris = 0;
};

// Closest hit on the mesh within tmax: sorted position (-1 on a miss), t, and
// the geometric normal of the hit as the Pluecker test computes it.
template <bool COMPAT, bool COUNT>
FLG_D int32_t mesh_hit(const MeshGeom& m, V3 o, V3 d, float tmax, float* t, V3* ng, VpCount& cnt) {
  float best = tmax;
  Counters c{0, 0};
  const int32_t p = trace_lane<COMPAT, COUNT, kMeshBlock, kMeshRing, true>(m.g, o, d, best, c, m.ring);
  if (COUNT) {
    ++cnt.rays;
    cnt.nodes += c.nodes;
    cnt.tris += c.tris;
  }
  *t = best;
  if (p >= 0) {
    const BTri tri = m.g.tris[p];
    const V3 v0 = v3(tri.v0.x, tri.v0.y, tri.v0.z) - o, v1 = v3(tri.v1.x, tri.v1.y, tri.v1.z) - o,
             v2 = v3(tri.v2.x, tri.v2.y, tri.v2.z) - o;
    *ng = stable_normalize(stable_triangle_normal(v2 - v0, v0 - v1, v1 - v2), COMPAT);
  }
  return p;
}

template <bool COMPAT>
FLG_D int32_t mesh_hit(const MeshGeom& m, V3 o, V3 d, float tmax, float* t, V3* ng) {
  VpCount none;
  return mesh_hit<COMPAT, false>(m, o, d, tmax, t, ng, none);
}

// VolEstimateTr (integrator.cpp:40-64) along (o, d) up to tmax, starting
// inside the medium or not.
template <bool COMPAT, bool COUNT>
FLG_D float shadow_tr(const MeshGeom& g, const VolDev& v, HaltonDraw& rng, V3 o, V3 d, float tmax,
                      bool inside, VpCount& cnt) {
  float tr = 1.0f;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t;
    V3 ng;
    if (mesh_hit<COMPAT, COUNT>(g, o, d, t_max, &t, &ng, cnt) < 0) return tr;
    if (inside) tr *= vol_tr<COUNT>(v, rng, o, d, t, cnt.look);
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
  V3 L;  // RGB
  float beta;
  int bounces, crossings;
  bool inside;
  // film mode: the pixel, the film position of the splatted sample, and
  // whether the path in flight is the pixel-centre path (not splatted)
  uint32_t pix;
  float fx, fy;
  bool centre;
};

// The camera ray at film position (fx, fy); the camera is outside the mesh.
FLG_D void vp_start(const CamDevV& cam, float fx, float fy, VPath& s) {
  gen_ray_v(cam, fx, fy, &s.o, &s.d);
  s.L = v3(0.0f, 0.0f, 0.0f);
  s.beta = 1.0f;
  s.bounces = 0;
  s.crossings = 0;
  s.inside = false;
}

// Camera ray of path p (pixel p / spp, sample p % spp) on the per-sample
// Halton schedule.
FLG_D void vp_begin(const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, int32_t spp,
                    uint32_t p, VPath& s) {
  const uint32_t pix = p / (uint32_t)spp, smp = p - pix * (uint32_t)spp;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  vp_start(cam, (float)x + ux, (float)y + uy, s);
}

// Film task p of the reference's render loop (tasks = spp - 1 per pixel, the
// schedule of the volume-only integrator's film mode): task 0 runs the
// pixel-centre path, then sample 1 on the same Halton index from the
// dimension the centre path reached; task t >= 1 is sample t + 1 on index
// offset + t * stride from dimension 0.
FLG_D void vp_film_begin(const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, int32_t tasks,
                         uint32_t p, VPath& s) {
  const uint32_t pix = p / (uint32_t)tasks, t = p - pix * (uint32_t)tasks;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.pix = pix;
  s.ps.index = halton_pixel_offset(hp, x, y) + t * hp.stride;
  s.ps.dim = 0;
  s.centre = t == 0;
  if (s.centre) {
    vp_start(cam, (float)x + 0.5f, (float)y + 0.5f, s);
    return;
  }
  s.fx = (float)x + s.ps.next(tab, hp);
  s.fy = (float)y + s.ps.next(tab, hp);
  vp_start(cam, s.fx, s.fy, s);
}

// A film task's path has completed: the centre path goes on to sample 1
// (false: a path is running again); any other path completes the task.
FLG_D bool vp_film_next(const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, VPath& s) {
  if (!s.centre) return true;
  const int x = (int)(s.pix % (uint32_t)cam.w), y = (int)(s.pix / (uint32_t)cam.w);
  s.centre = false;
  s.fx = (float)x + s.ps.next(tab, hp);
  s.fy = (float)y + s.ps.next(tab, hp);
  vp_start(cam, s.fx, s.fy, s);
  return false;
}

// One trip of the reference's loop; returns true when the path is complete.
template <bool COMPAT, bool COUNT>
FLG_D bool vp_step(const MeshGeom& g, const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                   const HaltonParams& hp, int max_depth, VPath& s, VpCount& cnt) {
  float t;
  V3 ng;
  const bool hit = mesh_hit<COMPAT, COUNT>(g, s.o, s.d, kInf, &t, &ng, cnt) >= 0;
  if (s.bounces == 0 && !hit && L.env.present) add_scaled(s.L, s.beta, L.env.radiance);
  if (!hit || s.bounces >= max_depth) return true;
  bool success = false;
  V3 pos;
  if (s.inside) {
    HaltonDraw rng{s.ps, tab, hp};
    s.beta *= vol_sample<COUNT>(v, rng, s.o, s.d, t, &success, &pos, cnt.look);
  }
  if (success) {
    const V3 wo = -s.d;
    add_scaled(s.L, s.beta, sample_one_light(L, v.g, tab, hp, s.ps, pos, wo, [&](V3 p, V3 d, float tmax) {
                 HaltonDraw rng{s.ps, tab, hp};
                 return shadow_tr<COMPAT, COUNT>(g, v, rng, p, d, tmax, true, cnt);
               }));
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

// Adds the wave's counts into counts[0] (lookups), counts[2] (rays),
// counts[3] (node visits) and counts[4] (triangle tests).
template <bool COUNT>
FLG_D void vp_count(const VpCount& cnt, unsigned long long* counts) {
  if (COUNT) {
    unsigned long long c[4] = {cnt.look, cnt.rays, cnt.nodes, cnt.tris};
    for (int k = 0; k < 4; ++k)
      for (int o = 32; o >= 1; o >>= 1) c[k] += __shfl_xor(c[k], o, 64);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(counts, c[0]);
      atomicAdd(counts + 2, c[1]);
      atomicAdd(counts + 3, c[2]);
      atomicAdd(counts + 4, c[3]);
    }
  }
}

// Task p: path sample p (pixel p / tasks, sample p % tasks), or with FILM
// film task p of the render loop.
template <bool FILM>
FLG_D void vp_task_begin(const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam, int32_t tasks,
                         uint32_t p, VPath& s) {
  if (FILM) vp_film_begin(tab, hp, cam, tasks, p, s);
  else vp_begin(tab, hp, cam, tasks, p, s);
}

template <bool FILM>
FLG_D void vp_task_finish(uint32_t p, const VPath& s, float* lpath, float2* fpos) {
  float* o = lpath + 3 * (size_t)p;
  o[0] = s.L.x;
  o[1] = s.L.y;
  o[2] = s.L.z;
  if (FILM) fpos[p] = make_float2(s.fx, s.fy);
}

// One lane per task.
template <bool COMPAT, bool COUNT, bool FILM>
__global__ __launch_bounds__(kMeshBlock, kMeshWaves) void k_volpath(TraceGeom g, VolDev v, LightsDev L,
                                                                    HaltonTables tab_g, HaltonParams hp,
                                                                    CamDevV cam, int32_t tasks, int32_t max_depth,
                                                                    float* lpath, float2* fpos,
                                                                    unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  __shared__ uint2 ring[kMeshBlock * kMeshRing];
  const MeshGeom mg{g, ring};
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ntasks = (int64_t)cam.w * cam.h * tasks;
  VpCount cnt;
  if (p < ntasks) {
    VPath s;
    vp_task_begin<FILM>(tab, hp, cam, tasks, (uint32_t)p, s);
    // one trip per loop iteration; a film task's centre path goes on to
    // sample 1 (written without a side effect inside the loop condition: the
    // short-circuit form `!step() || !film_next()` was compiled wrongly for
    // lanes of one wave that reach the end of the centre path on different
    // trips, which then traced a path twice and kept the centre radiance)
    bool more = true;
    while (more) {
      if (vp_step<COMPAT, COUNT>(mg, v, L, tab, hp, max_depth, s, cnt)) more = FILM && !vp_film_next(tab, hp, cam, s);
    }
    vp_task_finish<FILM>((uint32_t)p, s, lpath, fpos);
  }
  vp_count<COUNT>(cnt, nlookups);
}

// Tiles of the order in which the regenerating mesh kernels start paths
// (tile_order): 3.9% faster forward and 3.4% faster adjoint than scanline
// order on the 8M-triangle interface-mesh scene at 1920x1080.
constexpr uint32_t kMeshTileW = 16, kMeshTileH = 16;

// Persistent lanes with path regeneration (regen_loop), paths started in
// tile_order; same per-task code.
template <bool COMPAT, bool COUNT, bool FILM>
__global__ __launch_bounds__(kMeshBlock, kMeshWaves) void k_volpath_regen(TraceGeom g, VolDev v, LightsDev L,
                                                                          HaltonTables tab_g, HaltonParams hp,
                                                                          CamDevV cam, int32_t tasks,
                                                                          int32_t max_depth, float* lpath,
                                                                          float2* fpos, unsigned int* next_path,
                                                                          unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  __shared__ uint2 ring[kMeshBlock * kMeshRing];
  const MeshGeom mg{g, ring};
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t ntasks = (uint32_t)((int64_t)cam.w * cam.h * tasks);
  VpCount cnt;
  regen_loop<VPath>(
      ntasks, next_path,
      [&](uint32_t q, VPath& s) {
        vp_task_begin<FILM>(tab, hp, cam, tasks, tile_order<kMeshTileW, kMeshTileH>(q, tasks, cam.w, cam.h), s);
        return false;
      },
      [&](VPath& s) {
        return vp_step<COMPAT, COUNT>(mg, v, L, tab, hp, max_depth, s, cnt) &&
               (!FILM || vp_film_next(tab, hp, cam, s));
      },
      [&](uint32_t q, const VPath& s) {
        vp_task_finish<FILM>(tile_order<kMeshTileW, kMeshTileH>(q, tasks, cam.w, cam.h), s, lpath, fpos);
      });
  vp_count<COUNT>(cnt, nlookups);
}

__global__ void k_vp_zero(unsigned long long* p) {
  if (threadIdx.x < 5) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// The differentiable estimator (diff.hpp) over the same scene: VolPath's path
// structure and draw order (light sample, then phase sample; no roulette),
// differential delta tracking inside the medium up to the next interface,
// and each shadow ray's ratio track through the interfaces as VolEstimateTr
// walks it, on one walk stream whose draws continue from segment to segment.

// The ratio track of walk `key` from a medium vertex along (o, d) up to tmax
// (product of its factors; nothing after the last interface, as
// VolEstimateTr).
template <bool COMPAT>
FLG_D Track mesh_track_measure(const MeshGeom& g, const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax) {
  Track r{1.0f, 0};
  uint32_t n = 0;
  bool inside = true;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t;
    V3 ng;
    if (mesh_hit<COMPAT>(g, o, d, t_max, &t, &ng) < 0) return r;
    if (inside && !seg_measure(D, key, n, o, d, t, r)) return r;
    t_max -= t;
    if (t_max < 0.0f) return r;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);
    d = normalize(d);
    inside = dot(d, ng) < 0.0f;
  }
  return r;
}

// Second walk of the same track: its derivative, times scale, deposited.
template <bool COMPAT>
FLG_D void mesh_track_deposit(const MeshGeom& g, const DiffDev& D, uint32_t key, V3 o, V3 d, float tmax,
                              const Track& r, float scale, float* grad) {
  uint32_t n = 0;
  bool inside = true;
  float t_max = tmax;
  for (int k = 0; k < kMaxCrossings; ++k) {
    float t;
    V3 ng;
    if (mesh_hit<COMPAT>(g, o, d, t_max, &t, &ng) < 0) return;
    if (inside && !seg_deposit(D, key, n, o, d, t, r, scale, grad)) return;
    t_max -= t;
    if (t_max < 0.0f) return;
    const V3 ph = o + d * t;
    o = offset_ray_origin(ph, ng, d);
    d = normalize(d);
    inside = dot(d, ng) < 0.0f;
  }
}

// The path state of the differentiable estimator over the mesh; rgb, acc,
// delta, L_total, wp and we as in the volume-only estimator's DiffPath.
struct DiffVPath {
  PathSampler ps;
  V3 o, d, rgb;
  float beta, acc, delta, L_total, wp, we;
  int bounces, crossings;
  bool inside;
  uint32_t path;
  uint64_t mask;  // tape bits of the path's walks with exactly one zero factor
};

// Camera ray of path p; the adjoint (ADJ) also takes the path's radiance and
// adjoint weight.  Returns true when the path has nothing to do.
template <bool ADJ>
FLG_D bool vpd_begin(const DiffDev& D, const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp,
                     const CamDevV& cam, int32_t spp, uint32_t p, const float* lpath, const float* dl_drgb,
                     DiffVPath& s) {
  const uint32_t pix = p / (uint32_t)spp, smp = p - pix * (uint32_t)spp;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.acc = 0.0f;
  s.rgb = v3(0.0f, 0.0f, 0.0f);
  s.path = p;
  s.mask = 0;
  s.delta = s.L_total = s.wp = s.we = 0.0f;
  if (ADJ) {
    const float* gp = dl_drgb + 3 * (int64_t)pix;
    light_weights(L, gp, &s.wp, &s.we);
    if (s.wp == 0.0f && s.we == 0.0f) return true;
    s.delta = 1.0f / (float)spp;
    const float* l = lpath + 3 * (int64_t)p;
    s.L_total = gp[0] * l[0] + gp[1] * l[1] + gp[2] * l[2];
    if (D.tape) s.mask = ((uint64_t)D.mask_hi[p] << 32) | D.mask_lo[p];
  }
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  gen_ray_v(cam, (float)x + ux, (float)y + uy, &s.o, &s.d);
  s.beta = 1.0f;
  s.bounces = 0;
  s.crossings = 0;
  s.inside = false;
  return false;
}

// One trip of VolPath's loop; returns true when the path is complete.
template <bool ADJ, bool COMPAT>
FLG_D bool vpd_step(const MeshGeom& g, const DiffDev& D, const LightsDev& L, const HaltonTables& tab,
                    const HaltonParams& hp, int max_depth, float* grad, float* galpha, DiffVPath& s) {
  float t;
  V3 ng;
  const bool hit = mesh_hit<COMPAT>(g, s.o, s.d, kInf, &t, &ng) >= 0;
  if (s.bounces == 0 && !hit && L.env.present) {
    if (ADJ) s.acc += s.beta * s.we;
    else add_scaled(s.rgb, s.beta, L.env.radiance);
  }
  if (!hit || s.bounces >= max_depth) return true;
  V3 pos;
  if (s.inside && diff_delta_track<ADJ>(D, tab, hp, s.ps, s.o, s.d, t, s.beta, s.acc, s.delta, s.L_total, grad,
                                        galpha, &pos)) {
    const V3 wo = -s.d;
    const NeeSample ns = nee_sample(D, L, tab, hp, s.ps, pos, wo);
    if (ns.valid) {
      const V3 diff = ns.p_light - pos;
      const float nrm = sqrtf(sqnorm(diff));
      const V3 dd = diff / nrm;
      const uint32_t w = (uint32_t)s.bounces + 1;
      const uint32_t key = walk_key(s.path, w);
      const Track tr = track<ADJ>(D, s.path, w, s.mask, [&] {
        return mesh_track_measure<COMPAT>(g, D, key, pos, dd, nrm - SHADOW_EPS);
      });
      if (ADJ) {
        const float K = nee_k(ns, nee_li_w(ns, s.wp, s.we));
        if (track_live(tr) && s.delta * s.beta * K != 0.0f)
          mesh_track_deposit<COMPAT>(g, D, key, pos, dd, nrm - SHADOW_EPS, tr, s.delta * s.beta * K, grad);
        s.acc += s.beta * K * track_value(tr);
      } else {
        add_nee(s.rgb, L, ns, s.beta, track_value(tr));
      }
    }
    const float u = s.ps.next(tab, hp);
    const float vv = s.ps.next(tab, hp);
    const V3 wi = hg_sample(wo, u, vv, D.g);
    if (s.beta == 0.0f) return true;
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

// Forward: the path's RGB radiance and its tape bits.
FLG_D void vpd_finish(const DiffDev& D, uint32_t p, const DiffVPath& s, float* lpath) {
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
template <bool ADJ, bool COMPAT>
__global__ __launch_bounds__(kMeshBlock, kMeshWaves) void k_vpdiff(TraceGeom g, DiffDev D, LightsDev L, HaltonTables tab_g,
                                                HaltonParams hp, CamDevV cam, int32_t spp, int32_t max_depth,
                                                float* lpath, const float* dl_drgb, float* grad, float* galpha) {
  __shared__ int2 lds_primes[kLocalPrimes];
  __shared__ uint2 ring[kMeshBlock * kMeshRing];
  const MeshGeom mg{g, ring};
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npaths = (int64_t)cam.w * cam.h * spp;
  float ga = 0.0f;
  if (p < npaths) {
    DiffVPath s;
    bool done = vpd_begin<ADJ>(D, L, tab, hp, cam, spp, (uint32_t)p, lpath, dl_drgb, s);
    while (!done) done = vpd_step<ADJ, COMPAT>(mg, D, L, tab, hp, max_depth, grad, &ga, s);
    if (!ADJ) vpd_finish(D, (uint32_t)p, s, lpath);
  }
  reduce_galpha<ADJ>(ga, galpha);
}

// Persistent lanes with path regeneration (regen_loop), paths started in
// tile_order; same per-path code.
template <bool ADJ, bool COMPAT>
__global__ __launch_bounds__(kMeshBlock, kMeshWaves) void k_vpdiff_regen(TraceGeom g, DiffDev D, LightsDev L, HaltonTables tab_g,
                                                      HaltonParams hp, CamDevV cam, int32_t spp, int32_t max_depth,
                                                      float* lpath, const float* dl_drgb, float* grad,
                                                      float* galpha, unsigned int* next_path) {
  __shared__ int2 lds_primes[kLocalPrimes];
  __shared__ uint2 ring[kMeshBlock * kMeshRing];
  const MeshGeom mg{g, ring};
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t npaths = (uint32_t)((int64_t)cam.w * cam.h * spp);
  float ga = 0.0f;
  regen_loop<DiffVPath>(
      npaths, next_path,
      [&](uint32_t q, DiffVPath& s) {
        return vpd_begin<ADJ>(D, L, tab, hp, cam, spp, tile_order<kMeshTileW, kMeshTileH>(q, spp, cam.w, cam.h), lpath, dl_drgb, s);
      },
      [&](DiffVPath& s) { return vpd_step<ADJ, COMPAT>(mg, D, L, tab, hp, max_depth, grad, &ga, s); },
      [&](uint32_t q, const DiffVPath& s) {
        if (!ADJ) vpd_finish(D, tile_order<kMeshTileW, kMeshTileH>(q, spp, cam.w, cam.h), s, lpath);
      });
  reduce_galpha<ADJ>(ga, galpha);
}

template <bool ADJ, bool COMPAT>
int launch_vpdiff(flg_ctx* ctx, const flg_render_cfg* cfg, const TraceGeom& g, const DiffDev& D,
                  const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam,
                  float* lpath, const float* dl, float* grad, float* galpha) {
  const int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
  const int sched = cfg->schedule ? cfg->schedule : 2;
  if (sched == 1) {
    k_vpdiff<ADJ, COMPAT><<<div_up(npaths, kMeshBlock), kMeshBlock, 0, ctx->stream>>>(g, D, L, tab, hp, cam, cfg->spp,
                                                                        cfg->max_depth, lpath, dl, grad, galpha);
    return FLG_OK;
  }
  int st = FLG_OK;
  unsigned long long* counter =
      (unsigned long long*)scratch(ctx, ADJ ? "vpdiff_counter_adj" : "vpdiff_counter_fwd", 256, &st);
  if (!counter) return st;
  k_vp_zero<<<1, 64, 0, ctx->stream>>>(counter);
  const int grid = (int)std::min<int64_t>(div_up(npaths, kMeshBlock), (int64_t)ctx->num_cus * 8);
  k_vpdiff_regen<ADJ, COMPAT><<<grid, kMeshBlock, 0, ctx->stream>>>(g, D, L, tab, hp, cam, cfg->spp, cfg->max_depth,
                                                             lpath, dl, grad, galpha, (unsigned int*)counter);
  return FLG_OK;
}

// The mesh's checks shared by the entry points over an interface mesh.
int check_mesh(flg_ctx* ctx, const flg_bvh* bvh, const char* who) {
  if (!bvh || bvh->n <= 0) return set_error(ctx, FLG_ERR_INVALID, who);
  const float* mb = bvh->bound;
  if (!(mb[3] > mb[0] && mb[4] > mb[1] && mb[5] > mb[2]))
    return set_error(ctx, FLG_ERR_INVALID, "interface mesh: the mesh's bounding box is flat");
  return FLG_OK;
}

template <bool COMPAT, bool COUNT, bool FILM>
void launch_volpath(int sched, int blocks1, int blocks2, hipStream_t st, const TraceGeom& g, const VolDev& v,
                    const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp, const CamDevV& cam,
                    int32_t tasks, int32_t D, float* lpath, float2* fpos, unsigned int* next_path,
                    unsigned long long* nl) {
  if (sched == 1)
    k_volpath<COMPAT, COUNT, FILM><<<blocks1, kMeshBlock, 0, st>>>(g, v, L, tab, hp, cam, tasks, D, lpath, fpos, nl);
  else
    k_volpath_regen<COMPAT, COUNT, FILM><<<blocks2, kMeshBlock, 0, st>>>(g, v, L, tab, hp, cam, tasks, D, lpath, fpos,
                                                                        next_path, nl);
}

template <bool COMPAT, bool COUNT>
void launch_volpath_film(bool film, int sched, int blocks1, int blocks2, hipStream_t st, const TraceGeom& g,
                         const VolDev& v, const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp,
                         const CamDevV& cam, int32_t tasks, int32_t D, float* lpath, float2* fpos,
                         unsigned int* next_path, unsigned long long* nl) {
  if (film)
    launch_volpath<COMPAT, COUNT, true>(sched, blocks1, blocks2, st, g, v, L, tab, hp, cam, tasks, D, lpath, fpos,
                                        next_path, nl);
  else
    launch_volpath<COMPAT, COUNT, false>(sched, blocks1, blocks2, st, g, v, L, tab, hp, cam, tasks, D, lpath, fpos,
                                         next_path, nl);
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_volpath_mesh_render(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol, const flg_vol_scene* scene,
                            const flg_render_cfg* cfg, float* d_rgb, uint64_t* h_stats) {
  return host_guard(ctx, "flg_volpath_mesh_render", [&]() -> int {
    if (!ctx || !bvh || !vol || !scene || !cfg || !d_rgb || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
        cfg->max_depth < 1)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: bad arguments");
    if (cfg->film < 0 || cfg->film > 1)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: film must be 0 or 1");
    if ((int64_t)cfg->width * cfg->height * cfg->spp >= ((int64_t)1 << 31))
      return set_error(ctx, FLG_ERR_RANGE, "flg_volpath_mesh_render: too many paths");
    const int sched = cfg->schedule ? cfg->schedule : 2;
    if (sched < 1 || sched > 2)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_render: schedule must be 0, 1 or 2");
    int st = check_mesh(ctx, bvh, "flg_volpath_mesh_render: empty mesh");
    if (st) return st;
    const float* mb = bvh->bound;
    HaltonParams hp = make_halton_params(cfg->width, cfg->height);
    if ((uint64_t)hp.stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
      return set_error(ctx, FLG_ERR_RANGE, "flg_volpath_mesh_render: spp * Halton stride must stay at or below 2^32");
    const bool film = cfg->film == 1;
    const int64_t npix = (int64_t)cfg->width * cfg->height;
    // film mode: spp - 1 splatted samples per pixel (the first task also runs
    // the pixel-centre path); with spp == 1 nothing reaches the film
    const int32_t tasks = film ? cfg->spp - 1 : cfg->spp;
    const int64_t ntasks = npix * tasks;
    // [256 B counters][ntasks RGB radiances][pad to 8 B][ntasks film positions]
    char* sc = (char*)scratch(ctx, "volpath_mesh", (size_t)ntasks * (film ? 20 : 12) + 264, &st);
    if (!sc) return st;
    // nl[0]: lookups, nl[1]: the path counter, nl[2..4]: rays, node visits,
    // triangle tests (cleared together)
    unsigned long long* nl = (unsigned long long*)sc;
    unsigned int* next_path = (unsigned int*)(nl + 1);
    float* lpath = (float*)(sc + 256);
    float2* fpos = film ? (float2*)(sc + 256 + (size_t)ntasks * 12 + ((ntasks & 1) ? 4 : 0)) : nullptr;
    HaltonTables tab;
    if ((st = get_halton_tables(ctx, &tab))) return st;
    LightsDev L;
    if ((st = make_lights(ctx, vol, scene, &L, mb))) return st;
    // the medium on the mesh's bounding box (setBound)
    const VolDev v =
        voldev_of(vol->grid, vol->bmap, vol->nx, vol->ny, vol->nz, mb, vol->medium, vol->max_density);
    const CamDevV cam = make_cam_v(&scene->camera, cfg->width, cfg->height);
    const TraceGeom g = geom_of(bvh);
    const bool compat = bvh->flags & FLG_NORMAL_COMPAT;
    const bool count = h_stats != nullptr;
    const int blocks1 = (int)div_up(ntasks, kMeshBlock);
    const int blocks2 = (int)std::min<int64_t>(div_up(ntasks, kMeshBlock), (int64_t)ctx->num_cus * 8);
    k_vp_zero<<<1, 64, 0, ctx->stream>>>(nl);
    if (ntasks > 0) {
      ScopedTimer tm(ctx, "volpath_mesh_paths");
      if (compat && count)
        launch_volpath_film<true, true>(film, sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, tasks,
                                        cfg->max_depth, lpath, fpos, next_path, nl);
      else if (compat)
        launch_volpath_film<true, false>(film, sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, tasks,
                                         cfg->max_depth, lpath, fpos, next_path, nl);
      else if (count)
        launch_volpath_film<false, true>(film, sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, tasks,
                                         cfg->max_depth, lpath, fpos, next_path, nl);
      else
        launch_volpath_film<false, false>(film, sched, blocks1, blocks2, ctx->stream, g, v, L, tab, hp, cam, tasks,
                                          cfg->max_depth, lpath, fpos, next_path, nl);
    }
    if (film)
      film_resolve(ctx->stream, ntasks > 0 ? lpath : nullptr, fpos, cfg->width, cfg->height, tasks, d_rgb);
    else
      pixel_mean(ctx->stream, lpath, npix, cfg->spp, d_rgb);
    FLG_CHECK_LAUNCH(ctx);
    if (h_stats) {
      unsigned long long c[5] = {0, 0, 0, 0, 0};
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(c, nl, 40, hipMemcpyDeviceToHost, ctx->stream));
      FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
      h_stats[0] = (uint64_t)(ntasks > 0 ? npix * cfg->spp : 0);
      h_stats[1] = c[0];
      h_stats[2] = c[2];
      h_stats[3] = c[3];
      h_stats[4] = c[4];
    }
    return FLG_OK;
  });
}

int flg_volpath_mesh_diff_forward(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol, const flg_volume* proposal,
                                  const flg_vol_scene* scene, const flg_render_cfg* cfg,
                                  const flg_diff_params* params, float* d_rgb, float* d_lpath, void* d_tape) {
  return host_guard(ctx, "flg_volpath_mesh_diff_forward", [&]() -> int {
    int st = check_mesh(ctx, bvh, "flg_volpath_mesh_diff_forward: no mesh");
    if (st) return st;
    DiffDev D;
    LightsDev L;
    HaltonTables tab;
    HaltonParams hp;
    CamDevV cam;
    if ((st = diff_prepare(ctx, vol, proposal, scene, cfg, params, d_tape, bvh->bound, &D, &L, &tab, &hp, &cam)))
      return st;
    if (!d_rgb && !d_lpath) return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_diff_forward: no output");
    const int64_t npix = (int64_t)cfg->width * cfg->height, npaths = npix * cfg->spp;
    float* lpath = d_lpath;
    if (!lpath) {
      lpath = (float*)scratch(ctx, "vpdiff_lpath", (size_t)npaths * 12, &st);
      if (!lpath) return st;
    }
    const TraceGeom g = geom_of(bvh);
    {
      ScopedTimer tm(ctx, "vpdiff_forward");
      st = (bvh->flags & FLG_NORMAL_COMPAT)
               ? launch_vpdiff<false, true>(ctx, cfg, g, D, L, tab, hp, cam, lpath, nullptr, nullptr, nullptr)
               : launch_vpdiff<false, false>(ctx, cfg, g, D, L, tab, hp, cam, lpath, nullptr, nullptr, nullptr);
      if (st) return st;
      if (d_rgb) pixel_mean(ctx->stream, lpath, npix, cfg->spp, d_rgb);
    }
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

int flg_volpath_mesh_adjoint(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol, const flg_volume* proposal,
                             const flg_vol_scene* scene, const flg_render_cfg* cfg, const flg_diff_params* params,
                             const float* d_lpath, const void* d_tape, const float* d_dl_drgb,
                             const flg_brick_layout* bricks, float* d_grad, float* d_grad_albedo) {
  return host_guard(ctx, "flg_volpath_mesh_adjoint", [&]() -> int {
    int st = check_mesh(ctx, bvh, "flg_volpath_mesh_adjoint: no mesh");
    if (st) return st;
    DiffDev D;
    LightsDev L;
    HaltonTables tab;
    HaltonParams hp;
    CamDevV cam;
    if ((st = diff_prepare(ctx, vol, proposal, scene, cfg, params, d_tape, bvh->bound, &D, &L, &tab, &hp, &cam)))
      return st;
    if ((st = diff_set_bricks(ctx, bricks, &D))) return st;
    if (!d_lpath || !d_dl_drgb || !d_grad || !d_grad_albedo)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volpath_mesh_adjoint: null buffer");
    const TraceGeom g = geom_of(bvh);
    {
      ScopedTimer tm(ctx, "vpdiff_adjoint");
      float* lp = (float*)d_lpath;
      st = (bvh->flags & FLG_NORMAL_COMPAT)
               ? launch_vpdiff<true, true>(ctx, cfg, g, D, L, tab, hp, cam, lp, d_dl_drgb, d_grad, d_grad_albedo)
               : launch_vpdiff<true, false>(ctx, cfg, g, D, L, tab, hp, cam, lp, d_dl_drgb, d_grad, d_grad_albedo);
      if (st) return st;
    }
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

}  // extern "C"
