// Volume-only path integrator on gfx950: SVolIntegrator::Li
// (src/integrator.cpp:328-405) with delta tracking and ratio tracking
// (src/volume.cpp:75-148), next-event estimation (integrator.cpp:41-120),
// HG phase sampling, a constant environment light (light.cpp:93-159) and a
// point light.  One thread per path sample (pixel, s); the per-pixel value is
// the in-order float sum over s divided by spp, as in eval_pixel
// (integrator.cpp:136-159) under the per-sample Halton schedule, or the
// reference's render loop resolved through its film (Film::addSample /
// commitSamples, film.hpp:143-174).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "volume.hpp"

namespace flg {

namespace {

constexpr int kEnvN = 512;
__global__ __launch_bounds__(256) void k_pad(const float* src, float* dst, int nx, int ny, int nz) {
  const int px = nx + 2, py = ny + 2, pz = nz + 2;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t tot = (int64_t)px * py * pz;
  if (i >= tot) return;
  int x = (int)(i % px), y = (int)((i / px) % py), z = (int)(i / ((int64_t)px * py));
  float v = 0.0f;
  if (x >= 1 && x <= nx && y >= 1 && y <= ny && z >= 1 && z <= nz)
    v = src[((int64_t)(z - 1) * ny + (y - 1)) * nx + (x - 1)];
  dst[i] = v;
}

__global__ __launch_bounds__(256) void k_max(const float* src, int64_t n, float* out) {
  __shared__ float red[4];
  float m = -3.402823466e+38f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = smax(m, src[i]);
  for (int o = 32; o >= 1; o >>= 1) m = smax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = smax(smax(red[0], red[1]), smax(red[2], red[3]));
    out[blockIdx.x] = r;
  }
}

// Dist1D ctor (distribution.hpp:14-22)
void dist_init(const float* f, int n, float* cdf, float* integral) {
  cdf[0] = 0.0f;
  for (int i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / (float)n;
  *integral = cdf[n];
  if (*integral == 0.0f)
    for (int i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
  else
    for (int i = 1; i < n + 1; ++i) cdf[i] /= *integral;
}

// The scene's bounding sphere as InfiniteAreaLight::preprocess uses it
// (AABB::boundSphere, aabb.h:67-70): centre, and radius * 1000 capped at
// FLT_MAX / 2.
void env_sphere(const float* bounds, float* center, float* radius) {
  V3 lo = v3(bounds[0], bounds[1], bounds[2]);
  V3 hi = v3(bounds[3], bounds[4], bounds[5]);
  V3 c = (lo + hi) / 2.0f;
  V3 dd = hi - c;
  float r = sqrtf(sqnorm(dd));
  r *= 1000.0f;
  r = smin(3.402823466e+38f / 2, r);
  center[0] = c.x;
  center[1] = c.y;
  center[2] = c.z;
  *radius = r;
}

// InfiniteAreaLight(texture) with a ConstTexture (light.cpp:93-108) and
// preprocess (light.cpp:110-117) over the scene bound (= the volume box).
int build_env(flg_ctx* ctx, flg_volume* vol, V3 radiance) {
  // |texture| of a constant colour (Vector3f::norm: x*x + y*y + z*z, then
  // the square root); the tables depend on the colour through it only
  const float nrm = sqrtf(radiance.x * radiance.x + radiance.y * radiance.y + radiance.z * radiance.z);
  if (vol->env && vol->env_norm == nrm) return FLG_OK;
  const int NU = kEnvN, NV = kEnvN;
  std::vector<float> host((size_t)513 + 512 + 512 * 513 + 512 * 512 + 512, 0.0f);
  float* marg_cdf = host.data();
  float* marg_f = marg_cdf + 513;
  float* row_cdf = marg_f + 512;
  float* row_f = row_cdf + 512 * 513;
  float* row_int = row_f + 512 * 512;
  for (int v = 0; v < NV; ++v) {
    float sin_theta = sinf(PI * (float)(v + 0.5f) / (float)NV);
    for (int u = 0; u < NU; ++u) {
      float f = nrm;
      f *= sin_theta;
      row_f[v * NU + u] = f;
    }
    dist_init(row_f + v * NU, NU, row_cdf + v * 513, &row_int[v]);
    marg_f[v] = row_int[v];
  }
  float marg_int;
  dist_init(marg_f, NV, marg_cdf, &marg_int);
  if (!vol->env) FLG_CHECK_HIP(ctx, hipMalloc(&vol->env, host.size() * 4));
  FLG_CHECK_HIP(ctx, hipMemcpyAsync(vol->env, host.data(), host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  vol->env_norm = nrm;
  vol->env_marg_int = marg_int;
  env_sphere(vol->bounds, vol->env_center, &vol->env_radius);
  return FLG_OK;
}

// SVolIntegrator::Li (integrator.cpp:328-405), m_rrThreshold = 0.01, split
// into the path's start (camera ray, entry into the box, primary
// transmittance) and one step per scattering event, so that a persistent
// kernel can interleave paths; the arithmetic and the draw order are those
// of one uninterrupted call.
struct SvolPath {
  PathSampler ps;
  V3 o, d;
  V3 Lr;  // RGB
  float beta;
  int bounces;
  // film mode: the pixel, the film position of the splatted sample, and
  // whether the path in flight is the pixel-centre path (not splatted)
  uint32_t pix;
  float fx, fy;
  bool centre;
};

// The camera ray at film position (fx, fy) and its entry into the box.
// Returns true when the path is already complete (the ray misses the
// volume: the environment is its radiance); otherwise s.o is the entry point,
// s.d the direction, and the primary transmittance walk (from the camera,
// draws from s.ps) is next.
FLG_D bool svol_enter(const VolDev& v, const LightsDev& L, const CamDevV& cam, float fx, float fy, SvolPath& s) {
  V3 o, d;
  gen_ray_v(cam, fx, fy, &o, &d);
  s.Lr = v3(0.0f, 0.0f, 0.0f);
  s.beta = 1.0f;
  float t_min, t_max;
  if (!box_enter(v.bmin, v.bmax, o, d, &t_min, &t_max)) {
    if (L.env.present) add_scaled(s.Lr, s.beta, L.env.radiance);
    return true;
  }
  s.o = o + d * t_min;
  s.d = d;
  s.bounces = 1;
  return false;
}

// The primary transmittance tr's contribution.
FLG_D void svol_primary(const LightsDev& L, float tr, SvolPath& s) {
  if (L.env.present) add_scaled(s.Lr, tr, L.env.radiance);
}

// Starts a path at film position (fx, fy) with the sampler state in s.ps:
// camera ray, entry into the box, primary transmittance.  Returns true when
// the path is already complete (its camera ray misses the volume).
template <bool COUNT>
FLG_D bool svol_start(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                      const HaltonParams& hp, const CamDevV& cam, float fx, float fy, SvolPath& s,
                      uint32_t& nlook) {
  if (svol_enter(v, L, cam, fx, fy, s)) return true;
  HaltonDraw rng{s.ps, tab, hp};
  svol_primary(L, vol_tr<COUNT>(v, rng, cam.pos, s.d, kInf, nlook), s);
  return false;
}

// Starts path p (pixel p / spp, sample p % spp) of the per-sample schedule:
// Halton index offset(pixel) + s * stride, dimensions 0 and 1 for the film
// position, the estimator from dimension 2.
template <bool COUNT>
FLG_D bool svol_begin(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                      const HaltonParams& hp, const CamDevV& cam, int32_t spp, uint32_t p,
                      SvolPath& s, uint32_t& nlook) {
  const uint32_t pix = p / (uint32_t)spp, smp = p - pix * (uint32_t)spp;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
  s.ps.dim = 0;
  const float ux = s.ps.next(tab, hp);
  const float uy = s.ps.next(tab, hp);
  return svol_start<COUNT>(v, L, tab, hp, cam, (float)x + ux, (float)y + uy, s, nlook);
}

// The reference's render loop (eval_pixel, integrator.cpp:136-159, with
// HaltonSampler::setPixel / getPixelSample / reset, sampler.h:164-186).  Per
// pixel: a path from the pixel centre (x + 0.5, y + 0.5) on Halton index
// offset + 0 * stride from dimension 0, whose radiance is discarded; sample 1,
// whose film position takes the next two dimensions of the same index and
// whose path continues from there; sample s >= 2 on index offset +
// (s - 1) * stride from dimension 0, which is the per-sample schedule's sample
// s - 1.  Task t of a pixel (spp - 1 tasks) is sample t + 1; task 0 runs the
// centre path first.
template <bool COUNT>
FLG_D bool film_sample1(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                        const HaltonParams& hp, const CamDevV& cam, SvolPath& s, uint32_t& nlook) {
  const int x = (int)(s.pix % (uint32_t)cam.w), y = (int)(s.pix / (uint32_t)cam.w);
  s.centre = false;
  s.fx = (float)x + s.ps.next(tab, hp);
  s.fy = (float)y + s.ps.next(tab, hp);
  return svol_start<COUNT>(v, L, tab, hp, cam, s.fx, s.fy, s, nlook);
}

template <bool COUNT>
FLG_D bool film_begin(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                      const HaltonParams& hp, const CamDevV& cam, int32_t tasks, uint32_t p,
                      SvolPath& s, uint32_t& nlook) {
  const uint32_t pix = p / (uint32_t)tasks, t = p - pix * (uint32_t)tasks;
  const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
  s.pix = pix;
  s.ps.index = halton_pixel_offset(hp, x, y) + t * hp.stride;
  s.ps.dim = 0;
  if (t == 0) {
    s.centre = true;
    if (!svol_start<COUNT>(v, L, tab, hp, cam, (float)x + 0.5f, (float)y + 0.5f, s, nlook)) return false;
    return film_sample1<COUNT>(v, L, tab, hp, cam, s, nlook);
  }
  s.centre = false;
  s.fx = (float)x + s.ps.next(tab, hp);
  s.fy = (float)y + s.ps.next(tab, hp);
  return svol_start<COUNT>(v, L, tab, hp, cam, s.fx, s.fy, s, nlook);
}

// The first half of a scattering event: delta tracking to the next
// collision, then at the collision the HG draws and the light sample up to
// its shadow ray.  Returns false when the path is complete (no collision).
// s.o, s.d become the collision point and the scattered direction (the
// termination test that decides whether they are used comes after the light).
template <bool COUNT>
FLG_D bool svol_collide(const VolDev& v, const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp,
                        SvolPath& s, LightSample& ls, uint32_t& nlook) {
  HaltonDraw rng{s.ps, tab, hp};
  bool success = false;
  V3 pos;
  const float f = vol_sample<COUNT>(v, rng, s.o, s.d, kInf, &success, &pos, nlook);
  if (!success) return false;
  s.beta = s.beta * f;
  const V3 wo = -s.d;
  const float u = s.ps.next(tab, hp);
  const float vv = s.ps.next(tab, hp);
  const V3 wi = hg_sample(wo, u, vv, v.g);
  ls = light_pre(L, v.g, tab, hp, s.ps, pos, wo);
  s.o = pos;
  s.d = normalize(wi);
  return true;
}

// The second half, given the light estimate ld: radiance, the depth and
// throughput tests, Russian roulette after bounce 10.  Returns true when the
// path is complete.
FLG_D bool svol_continue(const HaltonTables& tab, const HaltonParams& hp, int max_depth, V3 ld, SvolPath& s) {
  add_scaled(s.Lr, s.beta, ld);
  if (s.bounces >= max_depth || s.beta * s.beta + s.beta * s.beta + s.beta * s.beta == 0.0f) return true;
  if (s.bounces > 10) {
    const float rr = 0.01f;
    if (s.ps.next(tab, hp) < rr) return true;
    s.beta /= (1.0f - rr);
  }
  ++s.bounces;
  return false;
}

// One scattering event of the path: delta tracking to the next collision, HG
// sampling, next-event estimation, termination tests.  Returns true when the
// path is complete.
template <bool COUNT>
FLG_D bool svol_step(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                     const HaltonParams& hp, int max_depth, SvolPath& s, uint32_t& nlook) {
  LightSample ls;
  if (!svol_collide<COUNT>(v, L, tab, hp, s, ls, nlook)) return true;
  V3 ld;
  if (ls.kind == 2) {
    V3 d;
    float tmax;
    light_ray(ls, s.o, &d, &tmax);
    HaltonDraw rng{s.ps, tab, hp};
    ld = light_post(L, ls, vol_tr<COUNT>(v, rng, s.o, d, tmax, nlook));
  } else {
    ld = light_unshadowed(ls);
  }
  return svol_continue(tab, hp, max_depth, ld, s);
}

// svol_step for film tasks: the end of the centre path starts sample 1.
template <bool COUNT>
FLG_D bool film_step(const VolDev& v, const LightsDev& L, const HaltonTables& tab,
                     const HaltonParams& hp, const CamDevV& cam, int max_depth, SvolPath& s,
                     uint32_t& nlook) {
  if (!svol_step<COUNT>(v, L, tab, hp, max_depth, s, nlook)) return false;
  if (!s.centre) return true;
  return film_sample1<COUNT>(v, L, tab, hp, cam, s, nlook);
}

template <bool COUNT>
FLG_D void count_lookups(uint32_t nlook, unsigned long long* nlookups) {
  if (COUNT) {
    unsigned long long c = nlook;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(nlookups, c);
  }
}

// Radiance of path sample p: three floats at 3 p.
FLG_D void put_rgb(float* lpath, uint32_t p, V3 L) {
  float* o = lpath + 3 * (size_t)p;
  o[0] = L.x;
  o[1] = L.y;
  o[2] = L.z;
}

// Per-sample outputs: radiance, and in film mode the film position.
template <bool FILM>
FLG_D void svol_finish(uint32_t p, const SvolPath& s, float* lpath, float2* fpos) {
  put_rgb(lpath, p, s.Lr);
  if (FILM) fpos[p] = make_float2(s.fx, s.fy);
}

// One lane per path sample (FILM: per film task; `tasks` per pixel).
template <bool COUNT, bool FILM>
__global__ __launch_bounds__(256) void k_svol(VolDev v, LightsDev L, HaltonTables tab_g,
                                              HaltonParams hp, CamDevV cam, int32_t tasks,
                                              int32_t max_depth, float* lpath, float2* fpos,
                                              unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npaths = (int64_t)cam.w * cam.h * tasks;
  uint32_t nlook = 0;
  if (p < npaths) {
    SvolPath s;
    bool done = FILM ? film_begin<COUNT>(v, L, tab, hp, cam, tasks, (uint32_t)p, s, nlook)
                     : svol_begin<COUNT>(v, L, tab, hp, cam, tasks, (uint32_t)p, s, nlook);
    while (!done)
      done = FILM ? film_step<COUNT>(v, L, tab, hp, cam, max_depth, s, nlook)
                  : svol_step<COUNT>(v, L, tab, hp, max_depth, s, nlook);
    svol_finish<FILM>((uint32_t)p, s, lpath, fpos);
  }
  count_lookups<COUNT>(nlook, nlookups);
}

// Persistent lanes with path regeneration: a lane whose path completes takes
// the next unstarted path (wave-aggregated fetch from a global counter), so
// short and long paths no longer idle each other's wave.  Every path runs
// the same per-path code as k_svol: results are identical.  Compiled for 5
// waves per SIMD (96 registers and a few spilled): 13.7 against 14.0 ms on
// the 256^3 fBm frame at the unconstrained 4.
template <bool COUNT, bool FILM>
__global__ __launch_bounds__(256, 5) void k_svol_regen(VolDev v, LightsDev L, HaltonTables tab_g,
                                                    HaltonParams hp, CamDevV cam, int32_t tasks,
                                                    int32_t max_depth, float* lpath, float2* fpos,
                                                    unsigned int* next_path,
                                                    unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t npaths = (uint32_t)((int64_t)cam.w * cam.h * tasks);
  uint32_t nlook = 0;
  regen_loop<SvolPath, kVolChunk>(
      npaths, next_path,
      [&](uint32_t p, SvolPath& s) {
        return FILM ? film_begin<COUNT>(v, L, tab, hp, cam, tasks, p, s, nlook)
                    : svol_begin<COUNT>(v, L, tab, hp, cam, tasks, p, s, nlook);
      },
      [&](SvolPath& s) {
        return FILM ? film_step<COUNT>(v, L, tab, hp, cam, max_depth, s, nlook)
                    : svol_step<COUNT>(v, L, tab, hp, max_depth, s, nlook);
      },
      [&](uint32_t p, const SvolPath& s) { svol_finish<FILM>(p, s, lpath, fpos); });
  count_lookups<COUNT>(nlook, nlookups);
}

// ------------------------------------------------ queue-based wavefront ---
// Schedule 3: SVolIntegrator::Li as stages over queues of path states (film
// = 1: the centre path and sample 1 of a task are one entry after the
// other).  A path's state lives in HBM between stages, in the queue that
// holds it.  Every stage is one persistent launch over its input queue and
// appends to the other queue:
//   begin   every task: Halton index, film position, camera ray, box entry;
//           a ray that misses completes its path; the others queue for the
//           primary transmittance walk;
//   walk    ratio tracking of each queued slot: the primary transmittance
//           (from the camera) or the shadow ray of its last collision, then
//           the rest of that step (radiance, depth and throughput tests,
//           roulette); live paths queue for a collision;
//   collide delta tracking to the next collision, the HG draws and the light
//           sample up to its shadow ray; paths that escape complete.
// A film task whose centre path completes starts sample 1 in the same slot
// (its primary walk runs at once when the walk stage completes the centre
// path, as that stage's output is the collision queue).  Each path runs the
// draws and arithmetic of svol_begin/svol_step, so the bits equal the other
// schedules'.  The launch sequence is fixed (begin, then walk and collide
// pairs enough for the deepest path, then a last walk); a stage reads its
// queue's length from device memory, so the host never waits between stages
// and the whole sequence can be captured once as a hipGraph.
enum : uint32_t { kWfCamera = 0, kWfShadow = 1, kWfUnshadowed = 2 };

// A queue of path states, one 88-byte record per entry in six arrays, in
// queue order: a stage reads its input queue's records at consecutive
// positions and writes its output records at the positions it claims, so
// both are coalesced.
struct WfQueue {
  float4* od;   // o.xyz, Lr.x
  float4* dir;  // d.xyz, beta
  float4* lw;   // pending light: f, d2, pdf, light_pdf (kWfShadow); estimate in .x (kWfUnshadowed)
  float4* sh;   // shadow ray: direction, tmax (kWfShadow)
  uint4* meta;  // Halton index, next dimension, bounces | kind << 16 | centre << 24, task
  float2* gb;   // Lr.y, Lr.z
};

FLG_D void wf_load(const WfQueue& Q, uint32_t i, SvolPath& s, uint32_t* kind, uint32_t* task) {
  const float4 a = Q.od[i], b = Q.dir[i];
  const uint4 m = Q.meta[i];
  const float2 c = Q.gb[i];
  s.o = v3(a.x, a.y, a.z);
  s.Lr = v3(a.w, c.x, c.y);
  s.d = v3(b.x, b.y, b.z);
  s.beta = b.w;
  s.ps.index = m.x;
  s.ps.dim = m.y;
  s.bounces = (int)(m.z & 0xFFFFu);
  *kind = (m.z >> 16) & 0xFFu;
  s.centre = (m.z >> 24) != 0;
  *task = m.w;
}

FLG_D void wf_store(const WfQueue& Q, uint32_t i, const SvolPath& s, uint32_t kind, uint32_t task) {
  Q.od[i] = make_float4(s.o.x, s.o.y, s.o.z, s.Lr.x);
  Q.gb[i] = make_float2(s.Lr.y, s.Lr.z);
  Q.dir[i] = make_float4(s.d.x, s.d.y, s.d.z, s.beta);
  Q.meta[i] = make_uint4(s.ps.index, s.ps.dim, (uint32_t)s.bounces | (kind << 16) | ((s.centre ? 1u : 0u) << 24),
                         task);
}

// The queue position of each lane that pushes (wave-aggregated: one atomic
// per wave on the queue's length).
FLG_D uint32_t wf_claim(bool push, uint32_t* n) {
  const uint64_t m = __ballot(push);
  if (!m) return 0;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(m);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(n, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Task p's path (film: its centre path, or sample 1 after it) is complete:
// a film task goes on to sample 1 (returns true: its primary walk is next);
// otherwise the task's outputs are written (false).
template <bool FILM>
FLG_D bool wf_complete(const VolDev& v, const LightsDev& L, const HaltonTables& tab, const HaltonParams& hp,
                       const CamDevV& cam, int32_t tasks, uint32_t p, SvolPath& s, float* lpath, float2* fpos) {
  while (FILM && s.centre) {
    const uint32_t pix = p / (uint32_t)tasks;
    const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
    s.centre = false;
    s.fx = (float)x + s.ps.next(tab, hp);
    s.fy = (float)y + s.ps.next(tab, hp);
    fpos[p] = make_float2(s.fx, s.fy);
    if (!svol_enter(v, L, cam, s.fx, s.fy, s)) return true;
  }
  put_rgb(lpath, p, s.Lr);
  return false;
}

// begin: task p's first path up to its primary walk (svol_begin,
// film_begin); queued unless it completes at once.
template <bool FILM>
__global__ __launch_bounds__(256) void k_wf_begin(VolDev v, LightsDev L, HaltonTables tab_g, HaltonParams hp,
                                                  CamDevV cam, int32_t tasks, float* lpath, float2* fpos,
                                                  WfQueue out, uint32_t* nout) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t ntasks = (uint32_t)((int64_t)cam.w * cam.h * tasks);
  for (uint32_t base = blockIdx.x * blockDim.x; base < ntasks; base += gridDim.x * blockDim.x) {
    const uint32_t p = base + threadIdx.x;
    bool push = false;
    SvolPath s;
    if (p < ntasks) {
      const uint32_t pix = p / (uint32_t)tasks, t = p - pix * (uint32_t)tasks;
      const int x = (int)(pix % (uint32_t)cam.w), y = (int)(pix / (uint32_t)cam.w);
      s.ps.index = halton_pixel_offset(hp, x, y) + t * hp.stride;
      s.ps.dim = 0;
      s.centre = FILM && t == 0;
      float fx, fy;
      if (s.centre) {
        fx = (float)x + 0.5f;
        fy = (float)y + 0.5f;
      } else {
        fx = (float)x + s.ps.next(tab, hp);
        fy = (float)y + s.ps.next(tab, hp);
        if (FILM) fpos[p] = make_float2(fx, fy);
      }
      push = !svol_enter(v, L, cam, fx, fy, s) || wf_complete<FILM>(v, L, tab, hp, cam, tasks, p, s, lpath, fpos);
    }
    const uint32_t at = wf_claim(push, nout);
    if (push) wf_store(out, at, s, kWfCamera, p);
  }
}

// walk: the pending ratio-tracking walk of each queued path and the rest of
// its step; paths that go on queue for a collision.
template <bool COUNT, bool FILM>
__global__ __launch_bounds__(256) void k_wf_walk(VolDev v, LightsDev L, HaltonTables tab_g, HaltonParams hp,
                                                 CamDevV cam, int32_t tasks, int32_t max_depth, float* lpath,
                                                 float2* fpos, WfQueue in, const uint32_t* nin, WfQueue out,
                                                 uint32_t* nout, unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t n = *nin;
  uint32_t nlook = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool push = false;
    uint32_t p = 0;
    SvolPath s;
    if (i < n) {
      uint32_t kind;
      wf_load(in, i, s, &kind, &p);
      bool walking = true;
      while (walking) {
        walking = false;
        if (kind == kWfCamera) {
          HaltonDraw rng{s.ps, tab, hp};
          svol_primary(L, vol_tr<COUNT>(v, rng, cam.pos, s.d, kInf, nlook), s);
          push = true;
        } else {
          V3 ld;
          if (kind == kWfShadow) {
            const float4 a = in.lw[i], b = in.sh[i];
            LightSample ls;
            ls.f = a.x;
            ls.d2 = a.y;
            ls.pdf = a.z;
            ls.light_pdf = a.w;
            HaltonDraw rng{s.ps, tab, hp};
            ld = light_post(L, ls, vol_tr<COUNT>(v, rng, s.o, v3(b.x, b.y, b.z), b.w, nlook));
          } else {
            const float z = in.lw[i].x;
            ld = v3(z, z, z);
          }
          if (!svol_continue(tab, hp, max_depth, ld, s)) {
            push = true;
          } else if (wf_complete<FILM>(v, L, tab, hp, cam, tasks, p, s, lpath, fpos)) {
            kind = kWfCamera;  // sample 1's primary walk, here: this stage feeds the collision queue
            walking = true;
          }
        }
      }
    }
    const uint32_t at = wf_claim(push, nout);
    if (push) wf_store(out, at, s, kWfCamera, p);
  }
  count_lookups<COUNT>(nlook, nlookups);
}

// collide: delta tracking of each queued path to its next collision and the
// light sample there; live paths queue for their shadow walk.
template <bool COUNT, bool FILM>
__global__ __launch_bounds__(256) void k_wf_collide(VolDev v, LightsDev L, HaltonTables tab_g, HaltonParams hp,
                                                    CamDevV cam, int32_t tasks, float* lpath, float2* fpos,
                                                    WfQueue in, const uint32_t* nin, WfQueue out, uint32_t* nout,
                                                    unsigned long long* nlookups) {
  __shared__ int2 lds_primes[kLocalPrimes];
  const HaltonTables tab = stage_halton(tab_g, lds_primes);
  const uint32_t n = *nin;
  uint32_t nlook = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool push = false;
    uint32_t p = 0, kind = kWfCamera;
    SvolPath s;
    LightSample ls;
    V3 shd = v3(0.0f, 0.0f, 0.0f);
    float shmax = 0.0f;
    if (i < n) {
      wf_load(in, i, s, &kind, &p);
      if (svol_collide<COUNT>(v, L, tab, hp, s, ls, nlook)) {
        push = true;
        if (ls.kind == 2) {
          light_ray(ls, s.o, &shd, &shmax);
          kind = kWfShadow;
        } else {
          ls.f = light_unshadowed(ls).x;  // the same value in every channel
          kind = kWfUnshadowed;
        }
      } else if (wf_complete<FILM>(v, L, tab, hp, cam, tasks, p, s, lpath, fpos)) {
        push = true;
        kind = kWfCamera;
      }
    }
    const uint32_t at = wf_claim(push, nout);
    if (push) {
      wf_store(out, at, s, kind, p);
      if (kind == kWfShadow) {
        out.lw[at] = make_float4(ls.f, ls.d2, ls.pdf, ls.light_pdf);
        out.sh[at] = make_float4(shd.x, shd.y, shd.z, shmax);
      } else if (kind == kWfUnshadowed) {
        out.lw[at].x = ls.f;
      }
    }
  }
  count_lookups<COUNT>(nlook, nlookups);
}

__global__ void k_wf_zero(uint32_t* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Pixel = the in-order float mean of its spp samples, per channel.  A block
// takes kMeanPix pixels; their samples are staged through LDS kMeanTile at a
// time, read as the contiguous rows they are in HBM (each pixel's samples are
// 3 spp consecutive floats), and thread 3 q + c then adds channel c of pixel
// q in sample order (the order of the per-sample schedule's mean).  The odd
// row pitch spreads the summing threads of a wave over the LDS banks.
constexpr int kMeanPix = 64, kMeanTile = 16, kMeanRow = 3 * kMeanTile + 1;
__global__ __launch_bounds__(256) void k_pixel_mean(const float* lpath, int64_t npix, int32_t spp,
                                                    float* rgb) {
  __shared__ float tile[kMeanPix * kMeanRow];
  const int64_t pix0 = (int64_t)blockIdx.x * kMeanPix;
  const int npx = (int)(npix - pix0 < kMeanPix ? npix - pix0 : kMeanPix);
  const int t = threadIdx.x, q = t / 3, c = t - 3 * q;
  float acc = 0.0f;
  for (int s0 = 0; s0 < spp; s0 += kMeanTile) {
    const int ns = spp - s0 < kMeanTile ? spp - s0 : kMeanTile;
    __syncthreads();
    for (int e = t; e < npx * 3 * ns; e += blockDim.x) {
      const int r = e / (3 * ns), o = e - r * 3 * ns;
      tile[r * kMeanRow + o] = lpath[((pix0 + r) * spp + s0) * 3 + o];
    }
    __syncthreads();
    if (q < npx)
      for (int k = 0; k < ns; ++k) acc += tile[q * kMeanRow + 3 * k + c];
  }
  if (q < npx) rgb[3 * (pix0 + q) + c] = acc / (float)spp;
}

// MitchellFilter::mitchell1D (common/filter.h:44-53) of the film's filter
// (film.hpp:55: radius 2, B = C = 1 / 3 in integer arithmetic, i.e. 0), for
// an offset d in pixels: x = |2 (d / 2)|; 0 for x > 1, else
// (12 x^3 - 18 x^2 + 6) / 6 in the reference's operation order.
FLG_D float film_mitchell1d(float d) {
  const float B = 0.0f, C = 0.0f;
  const float x = fabsf(2.0f * (d * 0.5f));
  if (x > 1.0f)
    return ((-B - 6.0f * C) * x * x * x + (6.0f * B + 30.0f * C) * x * x + (-12.0f * B - 48.0f * C) * x +
            (8.0f * B + 24.0f * C)) *
           (1.0f / 6.0f);
  return ((12.0f - 9.0f * B - 6.0f * C) * x * x * x + (-18.0f + 12.0f * B + 6.0f * C) * x * x +
          (6.0f - 2.0f * B)) *
         (1.0f / 6.0f);
}

// Film::addSample + commitSamples (film.hpp:143-174) as a gather: pixel (x, y)
// sums w * L and w over the samples of the 3x3 pixels around it (a sample of
// pixel q lies in [q, q + 1)^2, so only those reach |offset| < 1 where the
// filter is non-zero), in the order (qx, qy, sample); w = filter(sample - 0.5
// - pixel).  Pixels with sum(w) == 0 stay 0.
__global__ __launch_bounds__(256) void k_film_resolve(const float* lpath, const float2* fpos, int32_t w,
                                                      int32_t h, int32_t tasks, float* rgb) {
  const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= (int64_t)w * h) return;
  const int x = (int)(pix % w), y = (int)(pix / w);
  float I = 0.0f;
  V3 Ls = v3(0.0f, 0.0f, 0.0f);
  for (int qx = x - 1; qx <= x + 1; ++qx) {
    if (qx < 0 || qx >= w) continue;
    for (int qy = y - 1; qy <= y + 1; ++qy) {
      if (qy < 0 || qy >= h) continue;
      const int64_t base = ((int64_t)qy * w + qx) * tasks;
      for (int t = 0; t < tasks; ++t) {
        const float2 uv = fpos[base + t];
        const float wgt = film_mitchell1d((uv.x - 0.5f) - (float)x) * film_mitchell1d((uv.y - 0.5f) - (float)y);
        I = I + wgt;
        const float* l = lpath + 3 * (base + t);
        Ls.x = Ls.x + l[0] * wgt;
        Ls.y = Ls.y + l[1] * wgt;
        Ls.z = Ls.z + l[2] * wgt;
      }
    }
  }
  rgb[3 * pix] = I == 0.0f ? 0.0f : Ls.x / I;
  rgb[3 * pix + 1] = I == 0.0f ? 0.0f : Ls.y / I;
  rgb[3 * pix + 2] = I == 0.0f ? 0.0f : Ls.z / I;
}

__global__ __launch_bounds__(256) void k_fill(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------- sparse volumes ---
__global__ __launch_bounds__(256) void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Brick cell (x fastest, nbx * nby per layer) of each leaf: cell_leaf[cell] =
// leaf.  err bit 0: an origin outside the grid or not a multiple of 8; bit 1:
// two leaves with one origin.  covered: voxels of the grid inside a leaf.
__global__ __launch_bounds__(256) void k_leaf_cells(const int32_t* origin, int32_t n, int32_t nx, int32_t ny,
                                                    int32_t nz, int32_t nbx, int32_t nby, int32_t* cell_leaf,
                                                    int32_t* err, unsigned long long* covered) {
  const int32_t l = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (l >= n) return;
  const int32_t ox = origin[3 * l], oy = origin[3 * l + 1], oz = origin[3 * l + 2];
  if (ox < 0 || oy < 0 || oz < 0 || ox >= nx || oy >= ny || oz >= nz || ((ox | oy | oz) & (kPoolBrick - 1))) {
    atomicOr(err, 1);
    return;
  }
  const int64_t cell = ((int64_t)(oz >> 3) * nby + (oy >> 3)) * nbx + (ox >> 3);
  if (atomicCAS(&cell_leaf[cell], -1, l) != -1) {
    atomicOr(err, 2);
    return;
  }
  const int cx = min(kPoolBrick, nx - ox), cy = min(kPoolBrick, ny - oy), cz = min(kPoolBrick, nz - oz);
  atomicAdd(covered, (unsigned long long)(cx * cy * cz));
}

// Slot map: entry (cell, class c) = (leaf << 3) | d for the first offset d
// (0..7, a subset of c) whose brick cell holds a leaf, -1 if none.  Map cells
// run from -1 to nb - 1 per axis.
__global__ __launch_bounds__(256) void k_pool_map(const int32_t* cell_leaf, int32_t nbx, int32_t nby, int32_t nbz,
                                                  int32_t* bmap) {
  const int32_t bcx = nbx + 1, bcy = nby + 1;
  const int64_t total = (int64_t)bcx * bcy * (nbz + 1) * kPoolClasses;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int c = (int)(t & (kPoolClasses - 1));
  const int64_t cell = t >> 3;
  const int cx = (int)(cell % bcx) - 1, cy = (int)((cell / bcx) % bcy) - 1, cz = (int)(cell / ((int64_t)bcx * bcy)) - 1;
  int32_t e = -1;
  for (int d = 0; d < kPoolClasses; ++d) {
    if (d & ~c) continue;
    const int x = cx + (d & 1), y = cy + ((d >> 1) & 1), z = cz + ((d >> 2) & 1);
    if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
    const int32_t l = cell_leaf[((int64_t)z * nby + y) * nbx + x];
    if (l >= 0) {
      e = (l << 3) | d;
      break;
    }
  }
  bmap[t] = e;
}

// Pool slot of leaf l (one block per leaf): its 8^3 voxels and the apron,
// voxel (ox - 1 + a) per axis, read from the leaf that holds it, 0 outside
// the grid or in no leaf.
__global__ __launch_bounds__(256) void k_pool_fill(const int32_t* origin, const float* values, const int32_t* cell_leaf,
                                                   int32_t nx, int32_t ny, int32_t nz, int32_t nbx, int32_t nby,
                                                   float* pool) {
  const int32_t l = (int32_t)blockIdx.x;
  const int32_t ox = origin[3 * l], oy = origin[3 * l + 1], oz = origin[3 * l + 2];
  for (int a = threadIdx.x; a < kPoolFloats; a += blockDim.x) {
    const int x = ox - 1 + a % kPoolEdge, y = oy - 1 + (a / kPoolEdge) % kPoolEdge, z = oz - 1 + a / (kPoolEdge * kPoolEdge);
    float v = 0.0f;
    if (x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz) {
      const int32_t src = cell_leaf[((int64_t)(z >> 3) * nby + (y >> 3)) * nbx + (x >> 3)];
      if (src >= 0) v = values[((int64_t)src << 9) + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7))];
    }
    pool[(int64_t)l * kPoolFloats + a] = v;
  }
}

// Per-block maxima of the leaves' voxels inside the grid.
__global__ __launch_bounds__(256) void k_leaf_max(const int32_t* origin, const float* values, int32_t n,
                                                  int32_t nx, int32_t ny, int32_t nz, float* out) {
  __shared__ float red[4];
  float m = -3.402823466e+38f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ((int64_t)n << 9);
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t l = (int32_t)(t >> 9), q = (int32_t)(t & 511);
    const int x = origin[3 * l] + (q & 7), y = origin[3 * l + 1] + ((q >> 3) & 7), z = origin[3 * l + 2] + (q >> 6);
    if (x < nx && y < ny && z < nz) m = smax(m, values[t]);
  }
  for (int o = 32; o >= 1; o >>= 1) m = smax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = smax(smax(red[0], red[1]), smax(red[2], red[3]));
}

}  // namespace

void pixel_mean(hipStream_t st, const float* lpath, int64_t npix, int32_t spp, float* rgb) {
  if (npix > 0) k_pixel_mean<<<div_up(npix, kMeanPix), 256, 0, st>>>(lpath, npix, spp, rgb);
}

void film_resolve(hipStream_t st, const float* lpath, const float2* fpos, int32_t w, int32_t h, int32_t tasks,
                  float* rgb) {
  const int64_t npix = (int64_t)w * h;
  if (lpath && tasks > 0)
    k_film_resolve<<<div_up(npix, 256), 256, 0, st>>>(lpath, fpos, w, h, tasks, rgb);
  else
    k_fill<<<div_up(npix * 3, 256), 256, 0, st>>>(rgb, npix * 3, 0.0f);
}

VolDev make_voldev(const flg_volume* vol) {
  return voldev_of(vol->grid, vol->bmap, vol->nx, vol->ny, vol->nz, vol->bounds, vol->medium,
                   vol->max_density);
}

int make_lights(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* sc, LightsDev* L,
                const float* scene_bounds, bool may_build) {
  *L = LightsDev{};
  L->has_point = sc->has_point != 0;
  L->point_pos = v3(sc->point_pos[0], sc->point_pos[1], sc->point_pos[2]);
  L->point_intensity = v3(sc->point_intensity[0], sc->point_intensity[1], sc->point_intensity[2]);
  L->env.present = sc->has_env != 0;
  if (L->env.present) {
    const V3 rad = v3(sc->env_radiance[0], sc->env_radiance[1], sc->env_radiance[2]);
    const float nrm = sqrtf(rad.x * rad.x + rad.y * rad.y + rad.z * rad.z);
    if (!may_build && (!vol->env || vol->env_norm != nrm))
      return set_error(ctx, FLG_ERR_INVALID,
                       "the volume's environment-light table is not built for a radiance of norm " +
                           std::to_string(nrm) + ": render the scene once with flg_volpath_render first");
    int st = build_env(ctx, vol, rad);
    if (st) return st;
    const float* e = vol->env;
    L->env.marg_cdf = e;
    L->env.marg_f = e + 513;
    L->env.row_cdf = e + 513 + 512;
    L->env.row_f = e + 513 + 512 + 512 * 513;
    L->env.row_int = e + 513 + 512 + 512 * 513 + 512 * 512;
    L->env.marg_int = vol->env_marg_int;
    L->env.radiance = rad;
    L->env.center = v3(vol->env_center[0], vol->env_center[1], vol->env_center[2]);
    L->env.radius = vol->env_radius;
    if (scene_bounds) {
      float c[3];
      env_sphere(scene_bounds, c, &L->env.radius);
      L->env.center = v3(c[0], c[1], c[2]);
    }
  }
  return FLG_OK;
}

int refresh_max(flg_ctx* ctx, flg_volume* vol, const float* d_density) {
  int st = FLG_OK;
  const int64_t tot = (int64_t)vol->nx * vol->ny * vol->nz;
  const int blocks = 512;
  float* part = (float*)scratch(ctx, "vol_max", blocks * 4, &st);
  if (!part) return st;
  k_max<<<blocks, 256, 0, ctx->stream>>>(d_density, tot, part);
  FLG_CHECK_LAUNCH(ctx);
  std::vector<float> h(blocks);
  FLG_CHECK_HIP(ctx, hipMemcpyAsync(h.data(), part, blocks * 4, hipMemcpyDeviceToHost, ctx->stream));
  FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float m = -3.402823466e+38f;
  for (float x : h) m = smax(m, x);
  vol->max_density = m;
  return FLG_OK;
}

// Refills a sparse volume's pool from the leaf values and recomputes the
// maximum density: the leaves' voxels inside the grid, and the 0 of the
// voxels in no leaf when there are any (the dense grid's maximum).
int sparse_fill(flg_ctx* ctx, flg_volume* vol, const float* d_values, bool covers_all) {
  if (vol->n_leaves > 0) {
    const int32_t nbx = (vol->nx + kPoolBrick - 1) / kPoolBrick, nby = (vol->ny + kPoolBrick - 1) / kPoolBrick;
    k_pool_fill<<<vol->n_leaves, 256, 0, ctx->stream>>>(vol->leaf_origin, d_values, vol->cell_leaf, vol->nx, vol->ny,
                                                        vol->nz, nbx, nby, vol->grid);
    FLG_CHECK_LAUNCH(ctx);
  }
  float m = covers_all ? -3.402823466e+38f : 0.0f;
  if (vol->n_leaves > 0) {
    int st = FLG_OK;
    const int blocks = 512;
    float* part = (float*)scratch(ctx, "vol_max", blocks * 4, &st);
    if (!part) return st;
    k_leaf_max<<<blocks, 256, 0, ctx->stream>>>(vol->leaf_origin, d_values, vol->n_leaves, vol->nx, vol->ny,
                                                vol->nz, part);
    FLG_CHECK_LAUNCH(ctx);
    std::vector<float> h(blocks);
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(h.data(), part, blocks * 4, hipMemcpyDeviceToHost, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (float x : h) m = smax(m, x);
  }
  vol->max_density = m;
  return FLG_OK;
}

// Whether the leaves of a sparse volume cover every voxel of its grid.
static bool sparse_covers_all(const flg_volume* vol, unsigned long long covered) {
  return covered == (unsigned long long)vol->nx * (unsigned long long)vol->ny * (unsigned long long)vol->nz;
}

// ------------------------------------------------------ frame render ---
// The per-frame work of flg_volpath_render and flg_volpath_render_noalloc.
// Scratch: [256 B counters][npaths RGB radiances][pad to 8 B][npaths film
// positions], from the context's grow-only pool, or (noalloc) from the
// library's fixed per-GPU array, in which case nothing is allocated: the
// Halton tables come from flg_ctx_create and the environment table must
// already hold the scene's radiance.
static int volpath_render(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* scene, const flg_render_cfg* cfg,
                          float* d_rgb, uint64_t* h_stats, bool noalloc) {
  const char* what = noalloc ? "flg_volpath_render_noalloc" : "flg_volpath_render";
  if (!ctx || !vol || !scene || !cfg || !d_rgb || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
      cfg->max_depth < 1)
    return set_error(ctx, FLG_ERR_INVALID, std::string(what) + ": bad arguments");
  if ((int64_t)cfg->width * cfg->height * cfg->spp >= ((int64_t)1 << 31))
    return set_error(ctx, FLG_ERR_RANGE, std::string(what) + ": too many paths");
  if (cfg->film < 0 || cfg->film > 1) return set_error(ctx, FLG_ERR_INVALID, std::string(what) + ": film must be 0 or 1");
  const int sched = cfg->schedule ? cfg->schedule : 2;
  if (sched < 1 || sched > 3)
    return set_error(ctx, FLG_ERR_INVALID, std::string(what) + ": schedule must be 0, 1, 2 or 3");
  if (sched == 3 && noalloc)
    return set_error(ctx, FLG_ERR_INVALID,
                     std::string(what) + ": schedule 3 keeps the path state in allocated scratch; use 1 or 2");
  HaltonParams hp = make_halton_params(cfg->width, cfg->height);
  if ((uint64_t)hp.stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
    return set_error(ctx, FLG_ERR_RANGE, std::string(what) + ": spp * Halton stride must stay at or below 2^32");
  const bool film = cfg->film == 1;
  const int64_t npix = (int64_t)cfg->width * cfg->height;
  // film mode: spp - 1 splatted samples per pixel (the first task also runs
  // the pixel-centre path); with spp == 1 nothing reaches the film
  const int32_t tasks = film ? cfg->spp - 1 : cfg->spp;
  const int64_t npaths = npix * tasks;
  const size_t need = (size_t)npaths * (film ? 20 : 12) + 264;
  int st = FLG_OK;
  char* sc = nullptr;
  if (noalloc) {
    if (need > kNoAllocScratchBytes)
      return set_error(ctx, FLG_ERR_RANGE,
                       std::string(what) + ": the frame needs " + std::to_string(need) +
                           " bytes of scratch, more than the fixed " + std::to_string(kNoAllocScratchBytes) +
                           " (at most " + std::to_string((kNoAllocScratchBytes - 264) / 12) +
                           " paths, or " + std::to_string((kNoAllocScratchBytes - 264) / 20) +
                           " tasks with film = 1)");
    sc = (char*)ctx->noalloc_scratch;
    if (!sc) return set_error(ctx, FLG_ERR_HIP, std::string(what) + ": no device address for the fixed scratch");
  } else {
    sc = (char*)scratch(ctx, "vol_paths", need, &st);
    if (!sc) return st;
  }
  unsigned long long* nl = (unsigned long long*)sc;
  unsigned int* next_path = (unsigned int*)(nl + 1);  // cleared with the lookup count
  float* lpath = (float*)(sc + 256);
  float2* fpos = film ? (float2*)(sc + 256 + (size_t)npaths * 12 + ((npaths & 1) ? 4 : 0)) : nullptr;
  const bool count = h_stats != nullptr;
  HaltonTables tab;
  LightsDev L;
  if (npaths > 0) {
    if (noalloc) {
      if (!halton_tables_ready(ctx, &tab))
        return set_error(ctx, FLG_ERR_INVALID, std::string(what) + ": the context has no Halton tables");
    } else if ((st = get_halton_tables(ctx, &tab))) {
      return st;
    }
    if ((st = make_lights(ctx, vol, scene, &L, nullptr, !noalloc))) return st;
  }
  k_zero_u64<<<1, 64, 0, ctx->stream>>>(nl, 2);  // lookup count and path counter
  uint32_t* wf_counts = nullptr;
  int wf_ncounts = 0;
  if (npaths > 0 && sched == 3) {
    // [two queues of npaths 88-byte records, six arrays each][stage counts]
    const int pairs = film ? 2 * cfg->max_depth + 1 : cfg->max_depth + 1;
    wf_ncounts = 2 * pairs + 2;
    const size_t rec = (size_t)npaths * 16, qbytes = 5 * rec + (size_t)npaths * 8;
    char* w = (char*)scratch(ctx, "vol_wavefront", 2 * qbytes + (size_t)wf_ncounts * 4, &st);
    if (!w) return st;
    auto queue = [&](int k) {
      char* b = w + (size_t)k * qbytes;
      return WfQueue{(float4*)b, (float4*)(b + rec), (float4*)(b + 2 * rec), (float4*)(b + 3 * rec),
                     (uint4*)(b + 4 * rec), (float2*)(b + 5 * rec)};
    };
    const WfQueue q0 = queue(0), q1 = queue(1);
    wf_counts = (uint32_t*)(w + 2 * qbytes);
    const VolDev v = make_voldev(vol);
    const CamDevV cam = make_cam_v(&scene->camera, cfg->width, cfg->height);
    const int grid = (int)std::min<int64_t>(div_up(npaths, 256), (int64_t)ctx->num_cus * 8);
    const int32_t D = cfg->max_depth;
    uint32_t* counts = wf_counts;
    auto enqueue = [&](hipStream_t s) {
      k_wf_zero<<<1, 256, 0, s>>>(counts, wf_ncounts);
#define FLG_WF_STAGES(C, F)                                                                                      \
  k_wf_begin<F><<<grid, 256, 0, s>>>(v, L, tab, hp, cam, tasks, lpath, fpos, q0, counts);                        \
  for (int k = 0; k <= pairs; ++k) {                                                                             \
    k_wf_walk<C, F><<<grid, 256, 0, s>>>(v, L, tab, hp, cam, tasks, D, lpath, fpos, q0, counts + 2 * k, q1,      \
                                         counts + 2 * k + 1, nl);                                                \
    if (k < pairs)                                                                                               \
      k_wf_collide<C, F><<<grid, 256, 0, s>>>(v, L, tab, hp, cam, tasks, lpath, fpos, q1, counts + 2 * k + 1,    \
                                              q0, counts + 2 * k + 2, nl);                                       \
  }
      if (count && film) {
        FLG_WF_STAGES(true, true)
      } else if (count) {
        FLG_WF_STAGES(true, false)
      } else if (film) {
        FLG_WF_STAGES(false, true)
      } else {
        FLG_WF_STAGES(false, false)
      }
#undef FLG_WF_STAGES
    };
    ScopedTimer tm(ctx, "volpath_paths");
    if (cfg->use_graph) {
      std::string key = "volume wavefront:";
      auto put = [&key](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
      put(&v, sizeof v);
      put(&L, sizeof L);
      put(&tab, sizeof tab);
      put(&hp, sizeof hp);
      put(&cam, sizeof cam);
      put(&q0, sizeof q0);
      put(&q1, sizeof q1);
      const void* ptrs[4] = {lpath, fpos, counts, nl};
      put(ptrs, sizeof ptrs);
      const int32_t misc[6] = {tasks, D, grid, pairs, count ? 1 : 0, film ? 1 : 0};
      put(misc, sizeof misc);
      if ((st = graph_launch(ctx, key, enqueue, "volume wavefront"))) return st;
    } else {
      enqueue(ctx->stream);
    }
  }
  if (npaths > 0 && sched != 3) {
    VolDev v = make_voldev(vol);
    CamDevV cam = make_cam_v(&scene->camera, cfg->width, cfg->height);
    ScopedTimer tm(ctx, "volpath_paths");
    const int blocks1 = (int)div_up(npaths, 256);
    const int blocks2 = (int)std::min<int64_t>(div_up(npaths, 256), (int64_t)ctx->num_cus * 8);
#define FLG_SVOL_LAUNCH(C, F)                                                                              \
  if (sched == 1)                                                                                          \
    k_svol<C, F><<<blocks1, 256, 0, ctx->stream>>>(v, L, tab, hp, cam, tasks, cfg->max_depth, lpath, fpos, \
                                                   nl);                                                    \
  else                                                                                                     \
    k_svol_regen<C, F><<<blocks2, 256, 0, ctx->stream>>>(v, L, tab, hp, cam, tasks, cfg->max_depth, lpath, \
                                                         fpos, next_path, nl);
    if (count && film) {
      FLG_SVOL_LAUNCH(true, true)
    } else if (count) {
      FLG_SVOL_LAUNCH(true, false)
    } else if (film) {
      FLG_SVOL_LAUNCH(false, true)
    } else {
      FLG_SVOL_LAUNCH(false, false)
    }
#undef FLG_SVOL_LAUNCH
  }
  {
    ScopedTimer tm(ctx, "volpath_mean");
    if (film)
      film_resolve(ctx->stream, npaths > 0 ? lpath : nullptr, fpos, cfg->width, cfg->height, tasks, d_rgb);
    else
      pixel_mean(ctx->stream, lpath, npix, cfg->spp, d_rgb);
  }
  FLG_CHECK_LAUNCH(ctx);
  if (h_stats) {
    unsigned long long c = 0;
    uint32_t left = 0;  // schedule 3: slots the last walk stage queued (none may be)
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(&c, nl, 8, hipMemcpyDeviceToHost, ctx->stream));
    if (wf_counts)
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(&left, wf_counts + wf_ncounts - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (left)
      return set_error(ctx, FLG_ERR_HIP, std::string(what) + ": " + std::to_string(left) +
                                             " paths outlived the wavefront's stage sequence");
    h_stats[0] = (uint64_t)(npaths > 0 ? npix * cfg->spp : 0);
    h_stats[1] = c;
  }
  return FLG_OK;
}

}  // namespace flg

// ------------------------------------------------ fixed no-alloc scratch ---
// The scratch of flg_volpath_render_noalloc: a file-scope device array of
// kNoAllocScratchBytes, one copy on every GPU the library runs on, held until
// the process ends.  In host code the array's name is not a device address;
// its address on each device comes from hipGetSymbolAddress with that device
// current (flg_ctx_create calls noalloc_scratch_address once per device) and
// only that address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[flg::kNoAllocScratchBytes];

namespace flg {

int noalloc_scratch_address(int device, void** out) {
  static std::mutex mu;
  static void* addr[kMaxNoAllocDevices] = {};
  *out = nullptr;
  if (device < 0 || device >= kMaxNoAllocDevices) return FLG_ERR_RANGE;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) return FLG_ERR_HIP;
    addr[device] = p;
  }
  *out = addr[device];
  return FLG_OK;
}

}  // namespace flg

using namespace flg;

extern "C" {

int flg_volume_create(flg_ctx* ctx, const float* d_density, int32_t nx, int32_t ny, int32_t nz,
                      const float* h_bounds, const flg_medium* medium, flg_volume** out) {
  return host_guard(ctx, "flg_volume_create", [&]() -> int {
    if (!ctx || !d_density || !h_bounds || !medium || !out || nx <= 0 || ny <= 0 || nz <= 0)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volume_create: bad arguments");
    *out = nullptr;
    if ((int64_t)(nx + 2) * (ny + 2) * (nz + 2) >= ((int64_t)1 << 31))
      return set_error(ctx, FLG_ERR_RANGE, "flg_volume_create: grid too large");
    for (int k = 0; k < 3; ++k)
      if (!(h_bounds[3 + k] > h_bounds[k])) return set_error(ctx, FLG_ERR_INVALID, "flg_volume_create: empty bounds");
    FLG_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    flg_volume* vol = new flg_volume;
    vol->ctx = ctx;
    vol->nx = nx;
    vol->ny = ny;
    vol->nz = nz;
    for (int k = 0; k < 6; ++k) vol->bounds[k] = h_bounds[k];
    vol->medium = *medium;
    const int64_t tot = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
    hipError_t e = hipMalloc(&vol->grid, tot * 4);
    if (e != hipSuccess) {
      delete vol;
      return check_hip(ctx, e, "hipMalloc(volume grid)");
    }
    *out = vol;
    int st = flg_volume_update(ctx, vol, d_density);
    if (st) {
      flg_volume_destroy(vol);
      *out = nullptr;
    }
    return st;
  });
}

int flg_volume_update(flg_ctx* ctx, flg_volume* vol, const float* d_density) {
  return host_guard(ctx, "flg_volume_update", [&]() -> int {
    if (!ctx || !vol || !d_density) return set_error(ctx, FLG_ERR_INVALID, "flg_volume_update: bad arguments");
    if (vol->bmap)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volume_update: a sparse volume takes flg_volume_update_sparse");
    const int64_t tot = (int64_t)(vol->nx + 2) * (vol->ny + 2) * (vol->nz + 2);
    k_pad<<<div_up(tot, 256), 256, 0, ctx->stream>>>(d_density, vol->grid, vol->nx, vol->ny, vol->nz);
    FLG_CHECK_LAUNCH(ctx);
    return refresh_max(ctx, vol, d_density);
  });
}

int flg_volume_create_sparse(flg_ctx* ctx, int32_t n_leaves, const int32_t* d_leaf_origin,
                             const float* d_leaf_values, int32_t nx, int32_t ny, int32_t nz, const float* h_bounds,
                             const flg_medium* medium, flg_volume** out) {
  return host_guard(ctx, "flg_volume_create_sparse", [&]() -> int {
    if (!ctx || !h_bounds || !medium || !out || nx <= 0 || ny <= 0 || nz <= 0 || n_leaves < 0 ||
        (n_leaves > 0 && (!d_leaf_origin || !d_leaf_values)))
      return set_error(ctx, FLG_ERR_INVALID, "flg_volume_create_sparse: bad arguments");
    *out = nullptr;
    for (int k = 0; k < 3; ++k)
      if (!(h_bounds[3 + k] > h_bounds[k]))
        return set_error(ctx, FLG_ERR_INVALID, "flg_volume_create_sparse: empty bounds");
    const int64_t nbx = (nx + kPoolBrick - 1) / kPoolBrick, nby = (ny + kPoolBrick - 1) / kPoolBrick,
                  nbz = (nz + kPoolBrick - 1) / kPoolBrick;
    const int64_t map_entries = (nbx + 1) * (nby + 1) * (nbz + 1) * kPoolClasses;
    // lookups index the map and the pool with 32-bit integers
    if (map_entries >= ((int64_t)1 << 31) || (int64_t)n_leaves * kPoolFloats >= ((int64_t)1 << 31))
      return set_error(ctx, FLG_ERR_RANGE, "flg_volume_create_sparse: grid or leaf count too large");
    FLG_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    flg_volume* vol = new flg_volume;
    vol->ctx = ctx;
    vol->nx = nx;
    vol->ny = ny;
    vol->nz = nz;
    for (int k = 0; k < 6; ++k) vol->bounds[k] = h_bounds[k];
    vol->medium = *medium;
    vol->n_leaves = n_leaves;
    const int64_t ncells = nbx * nby * nbz;
    const int64_t nl = n_leaves > 0 ? n_leaves : 1;
    hipError_t e = hipMalloc(&vol->grid, nl * kPoolFloats * 4);
    if (e == hipSuccess) e = hipMalloc(&vol->bmap, map_entries * 4);
    if (e == hipSuccess) e = hipMalloc(&vol->cell_leaf, ncells * 4);
    if (e == hipSuccess) e = hipMalloc(&vol->leaf_origin, nl * 3 * 4);
    if (e != hipSuccess) {
      flg_volume_destroy(vol);
      return check_hip(ctx, e, "hipMalloc(sparse volume)");
    }
    *out = vol;
    auto fail = [&](int st) {
      flg_volume_destroy(vol);
      *out = nullptr;
      return st;
    };
    int st = FLG_OK;
    int32_t* check = (int32_t*)flg::scratch(ctx, "sparse_check", 256, &st);
    if (!check) return fail(st);
    unsigned long long* covered = (unsigned long long*)(check + 2);
    k_zero_u64<<<1, 64, 0, ctx->stream>>>((unsigned long long*)check, 2);
    k_fill_i32<<<div_up(ncells, 256), 256, 0, ctx->stream>>>(vol->cell_leaf, ncells, -1);
    if (n_leaves > 0) {
      if (hipMemcpyAsync(vol->leaf_origin, d_leaf_origin, (size_t)n_leaves * 3 * 4, hipMemcpyDeviceToDevice,
                         ctx->stream) != hipSuccess)
        return fail(set_error(ctx, FLG_ERR_HIP, "flg_volume_create_sparse: copy of the leaf origins failed"));
      k_leaf_cells<<<div_up(n_leaves, 256), 256, 0, ctx->stream>>>(vol->leaf_origin, n_leaves, nx, ny, nz,
                                                                    (int32_t)nbx, (int32_t)nby, vol->cell_leaf,
                                                                    check, covered);
    }
    int32_t h[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(h, check, 16, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      return fail(check_hip(ctx, hipGetLastError(), "flg_volume_create_sparse"));
    if (h[0] & 1)
      return fail(set_error(ctx, FLG_ERR_INVALID,
                            "flg_volume_create_sparse: a leaf origin is outside the grid or not a multiple of 8"));
    if (h[0] & 2) return fail(set_error(ctx, FLG_ERR_INVALID, "flg_volume_create_sparse: two leaves share an origin"));
    unsigned long long cov = 0;
    std::memcpy(&cov, h + 2, 8);
    k_pool_map<<<div_up(map_entries, 256), 256, 0, ctx->stream>>>(vol->cell_leaf, (int32_t)nbx, (int32_t)nby,
                                                                   (int32_t)nbz, vol->bmap);
    if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, FLG_ERR_HIP, "flg_volume_create_sparse: launch"));
    vol->sparse_covers_all = sparse_covers_all(vol, cov);
    st = sparse_fill(ctx, vol, d_leaf_values, vol->sparse_covers_all);
    if (st) return fail(st);
    return FLG_OK;
  });
}

int flg_volume_update_sparse(flg_ctx* ctx, flg_volume* vol, const float* d_leaf_values) {
  return host_guard(ctx, "flg_volume_update_sparse", [&]() -> int {
    if (!ctx || !vol || (!d_leaf_values && vol->n_leaves > 0))
      return set_error(ctx, FLG_ERR_INVALID, "flg_volume_update_sparse: bad arguments");
    if (!vol->bmap)
      return set_error(ctx, FLG_ERR_INVALID, "flg_volume_update_sparse: a dense volume takes flg_volume_update");
    return sparse_fill(ctx, vol, d_leaf_values, vol->sparse_covers_all);
  });
}

int flg_volume_brick_layout(const flg_volume* vol, flg_brick_layout* out) {
  return host_guard(vol ? vol->ctx : nullptr, "flg_volume_brick_layout", [&]() -> int {
    if (!vol || !out) return FLG_ERR_INVALID;
    if (!vol->bmap) return set_error(vol->ctx, FLG_ERR_INVALID, "flg_volume_brick_layout: not a sparse volume");
    out->brick = kPoolBrick;
    out->n_active = vol->n_leaves;
    out->nx = vol->nx;
    out->ny = vol->ny;
    out->nz = vol->nz;
    out->d_slot = vol->cell_leaf;
    return FLG_OK;
  });
}

int flg_volume_destroy(flg_volume* vol) {
  return host_guard(nullptr, "flg_volume_destroy", [&]() -> int {
    if (!vol) return FLG_ERR_INVALID;
    (void)hipSetDevice(vol->ctx->device);
    (void)hipStreamSynchronize(vol->ctx->stream);
    (void)hipFree(vol->grid);
    if (vol->bmap) (void)hipFree(vol->bmap);
    if (vol->cell_leaf) (void)hipFree(vol->cell_leaf);
    if (vol->leaf_origin) (void)hipFree(vol->leaf_origin);
    if (vol->env) (void)hipFree(vol->env);
    delete vol;
    return FLG_OK;
  });
}

int flg_volpath_render(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* scene,
                       const flg_render_cfg* cfg, float* d_rgb, uint64_t* h_stats) {
  return host_guard(ctx, "flg_volpath_render",
                    [&]() -> int { return volpath_render(ctx, vol, scene, cfg, d_rgb, h_stats, false); });
}

int flg_volpath_render_noalloc(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* scene,
                               const flg_render_cfg* cfg, float* d_rgb) {
  return host_guard(ctx, "flg_volpath_render_noalloc",
                    [&]() -> int { return volpath_render(ctx, vol, scene, cfg, d_rgb, nullptr, true); });
}

}  // extern "C"
