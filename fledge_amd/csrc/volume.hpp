// Device description of the dense density volume and the per-sample helpers
// of the volume-only integrator (shared by the forward and adjoint kernels).
#pragma once
#include "common.hpp"
#include "halton.hpp"

struct flg_volume {
  flg_ctx* ctx = nullptr;
  int32_t nx = 0, ny = 0, nz = 0;
  float bounds[6];
  flg_medium medium;
  float* grid = nullptr;  // dense: padded grid; sparse: the brick pool
  float max_density = 0;
  // sparse-occupancy volumes (flg_volume_create_sparse), else null / 0
  int32_t* bmap = nullptr;       // lookup slot map (flg::kPoolClasses entries per cell)
  int32_t* cell_leaf = nullptr;  // brick cell -> leaf index, -1 for none (x fastest)
  int32_t* leaf_origin = nullptr;
  int32_t n_leaves = 0;
  bool sparse_covers_all = false;  // the leaves cover every voxel of the grid
  // environment-light tables, rebuilt when the norm of the radiance changes
  // (the only way the colour enters them)
  float* env = nullptr;
  float env_norm = -1.0f;
  float env_marg_int = 0;
  float env_center[3] = {0, 0, 0};
  float env_radius = 0;
};

namespace flg {

// Sparse-occupancy grids keep each active 8^3 brick in a pool slot of 10^3
// floats: the brick and a one-voxel apron on every side holding the
// neighbouring voxels (0 outside the grid or in bricks that are not stored).
// A trilinear lookup whose lower corner is voxel (i, j, k) reads voxels
// {i, i+1} x {j, j+1} x {k, k+1}; every brick that holds one of them holds
// all eight in its apron, so the lookup reads one slot.  The slot map has, for
// each brick cell (cells -1 .. nb-1 per axis, so that corner -1 of the zero
// border has a cell) and each border class c (bit a set when the lower corner
// is the last voxel of its brick along axis a, so that the eight corners
// reach into the next brick along a), the entry (slot << 3) | d of a stored
// brick among the cells the corners touch, d its offset from the cell (a
// subset of c), or -1 when none is stored: all eight corners are then 0.
constexpr int kPoolBrick = 8;
constexpr int kPoolEdge = kPoolBrick + 2;
constexpr int kPoolFloats = kPoolEdge * kPoolEdge * kPoolEdge;
constexpr int kPoolClasses = 8;

// Dense grids are in HBM with a one-voxel zero border: (nx+2)(ny+2)(nz+2)
// floats, x fastest; voxel (i,j,k) of the caller's grid is at
// ((k+1)*py + j+1)*px + i+1.  Sparse grids: the brick pool and its slot map.
struct VolDev {
  const float* grid;    // dense padded grid, or the brick pool when bmap is set
  const int32_t* bmap;  // null: dense
  int32_t nx, ny, nz;
  int32_t px, pxy;  // padded row and slice strides (dense)
  int32_t bcx, bcy;  // slot-map cells per row and per layer (sparse)
  V3 bmin, bmax, scale;
  float sigma_s, sigma_a, sigma_t, g;
  float inv_max_density, max_mu;
};

// Device (or host-mirror) description of a grid: `data` is the padded dense
// grid (bmap null) or the brick pool of the slot map bmap.  OpenVDBVolume
// ctor (volume.cpp:33-59): sigma_t = sigma_a + sigma_s, the inverse max
// density through a double division.
inline VolDev voldev_of(const float* data, const int32_t* bmap, int32_t nx, int32_t ny, int32_t nz,
                        const float* bounds, flg_medium m, float max_density) {
  VolDev v;
  v.grid = data;
  v.bmap = bmap;
  v.nx = nx;
  v.ny = ny;
  v.nz = nz;
  v.px = nx + 2;
  v.pxy = (nx + 2) * (ny + 2);
  v.bcx = (nx + kPoolBrick - 1) / kPoolBrick + 1;
  v.bcy = (ny + kPoolBrick - 1) / kPoolBrick + 1;
  v.bmin = v3(bounds[0], bounds[1], bounds[2]);
  v.bmax = v3(bounds[3], bounds[4], bounds[5]);
  v.scale = v3((float)nx / (v.bmax.x - v.bmin.x), (float)ny / (v.bmax.y - v.bmin.y),
               (float)nz / (v.bmax.z - v.bmin.z));
  v.sigma_s = m.sigma_s;
  v.sigma_a = m.sigma_a;
  v.sigma_t = v.sigma_a + v.sigma_s;
  v.g = m.g;
  v.inv_max_density = (float)(1.0 / (double)max_density);
  v.max_mu = v.sigma_t * max_density;
  return v;
}

struct Tri8 {  // one trilinear lookup: lower corner, weights, value
  int32_t i, j, k;  // lower corner voxel, each >= -1
  bool inside;      // false when the lookup is outside the grid (density 0)
  float fx, fy, fz;
  float rho;
};

// The eight voxels' trilinear blend: rows g0 (y, z), g0 + sy (y + 1), g0 + sz
// (z + 1), g0 + sz + sy, each two adjacent floats.
FLG_HD float trilerp(const float* g0, int sy, int sz, float fx, float fy, float fz) {
  const float* g1 = g0 + sy;
  const float* g2 = g0 + sz;
  const float* g3 = g2 + sy;
  // all eight loads first: under the kernels' register caps the scheduler
  // would otherwise load a row, wait, blend it and only then load the next.
  // The empty asm takes the eight values at once, so the four row loads are
  // in flight together before the first blend.
  const float a0 = g0[0], a1 = g0[1], b0 = g1[0], b1 = g1[1];
  const float e0 = g2[0], e1 = g2[1], h0 = g3[0], h1 = g3[1];
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" ::"v"(a0), "v"(a1), "v"(b0), "v"(b1), "v"(e0), "v"(e1), "v"(h0), "v"(h1));
#endif
  float wx0 = 1.0f - fx, wx1 = fx;
  float wy0 = 1.0f - fy, wy1 = fy;
  float wz0 = 1.0f - fz, wz1 = fz;
  float c00 = a0 * wx0 + a1 * wx1;
  float c10 = b0 * wx0 + b1 * wx1;
  float c01 = e0 * wx0 + e1 * wx1;
  float c11 = h0 * wx0 + h1 * wx1;
  float c0 = c00 * wy0 + c10 * wy1;
  float c1 = c01 * wy0 + c11 * wy1;
  return c0 * wz0 + c1 * wz1;
}

// Trilinear lookup: index space g = (p - bmin) * scale - 0.5, zero outside the
// grid, lerps along x, then y, then z (the library's definition of the dense
// grid's interpolation; OpenVDB's BoxSampler is not in the reference tree).
// A sparse grid gives the same bits as the dense grid holding the same
// values: the same eight voxels, the same operations.
FLG_HD Tri8 lookup(const VolDev& v, V3 p) {
  float gx = (p.x - v.bmin.x) * v.scale.x - 0.5f;
  float gy = (p.y - v.bmin.y) * v.scale.y - 0.5f;
  float gz = (p.z - v.bmin.z) * v.scale.z - 0.5f;
  float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
  Tri8 r;
  r.fx = gx - flx;
  r.fy = gy - fly;
  r.fz = gz - flz;
  int i = (int)smin(smax(flx, -2.0f), (float)v.nx + 1.0f);
  int j = (int)smin(smax(fly, -2.0f), (float)v.ny + 1.0f);
  int k = (int)smin(smax(flz, -2.0f), (float)v.nz + 1.0f);
  r.i = i;
  r.j = j;
  r.k = k;
  r.rho = 0.0f;
  r.inside = !(i < -1 || j < -1 || k < -1 || i > v.nx - 1 || j > v.ny - 1 || k > v.nz - 1);
  if (!r.inside) return r;
  if (!v.bmap) {  // dense (a uniform branch: one grid per launch)
    r.rho = trilerp(v.grid + (((k + 1) * (v.ny + 2) + (j + 1)) * v.px + (i + 1)), v.px, v.pxy, r.fx, r.fy, r.fz);
    return r;
  }
  const int cls = ((i & 7) == 7 ? 1 : 0) | ((j & 7) == 7 ? 2 : 0) | ((k & 7) == 7 ? 4 : 0);
  const int cell = (((k >> 3) + 1) * v.bcy + ((j >> 3) + 1)) * v.bcx + ((i >> 3) + 1);
  const int e = v.bmap[cell * kPoolClasses + cls];
  if (e < 0) return r;  // no stored brick holds a corner: all eight are 0
  const int lx = (i & 7) + 1 - ((e & 1) << 3);
  const int ly = (j & 7) + 1 - ((e & 2) << 2);
  const int lz = (k & 7) + 1 - ((e & 4) << 1);
  r.rho = trilerp(v.grid + ((e >> 3) * kPoolFloats + (lz * kPoolEdge + ly) * kPoolEdge + lx), kPoolEdge,
                  kPoolEdge * kPoolEdge, r.fx, r.fy, r.fz);
  return r;
}

// AABB::intersect_pbrt (src/common/aabb.h:38-59)
FLG_HD bool box_pbrt(V3 lo, V3 hi, V3 o, V3 d, float tmax, float* t_min, float* t_max) {
  float t0 = 0.0f, t1 = tmax;
  float los[3] = {lo.x, lo.y, lo.z}, his[3] = {hi.x, hi.y, hi.z};
  float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / ds[i];
    float tn = (los[i] - os[i]) * inv;
    float tf = (his[i] - os[i]) * inv;
    if (tn > tf) { float t = tn; tn = tf; tf = t; }
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  *t_min = t0;
  *t_max = t1;
  return true;
}

// AABB::intersect (aabb.h:20-36) with the inverse through the double literal
FLG_HD bool box_enter(V3 lo, V3 hi, V3 o, V3 d, float* t_min, float* t_max) {
  V3 inv = v3((float)(1.0 / (double)d.x), (float)(1.0 / (double)d.y), (float)(1.0 / (double)d.z));
  V3 t1 = (lo - o) * inv;
  V3 t2 = (hi - o) * inv;
  V3 vmn = vmin(t1, t2), vmx = vmax(t1, t2);
  float lmin = smax(smax(vmn.x, vmn.y), vmn.z);
  float lmax = smin(smin(vmx.x, vmx.y), vmx.z);
  if (lmin > lmax) { float t = lmin; lmin = lmax; lmax = t; }
  if (lmin < lmax && lmax >= 0) {
    *t_min = lmin;
    *t_max = lmax;
    return true;
  }
  return false;
}

// HGP (math_utils.h:103-107)
FLG_HD float hg_p(V3 wi, V3 wo, float g) {
  float cos_theta = dot(wi, wo);
  float denom = 1.0f + g * g + 2.0f * g * cos_theta;
  return (1.0f - g * g) / (denom * sqrtf(denom)) * INV_4PI;
}

// CoordinateSystem + HGSampleP (math_utils.h:57-66, 112-128); u is the first
// draw (left-to-right argument evaluation at integrator.cpp:378).
FLG_HD V3 hg_sample(V3 wo, float u, float vv, float g) {
  float cos_theta;
  if ((double)fabsf(g) < 1e-3) {
    cos_theta = 1.0f - 2.0f * u;
  } else {
    float tmp = (1.0f - g * g) / (1.0f + g - 2.0f * g * u);
    cos_theta = -(1.0f + g * g - tmp * tmp) / (2.0f * g);
  }
  float phi = 2.0f * PI * vv;
  float x = 1.0f - cos_theta * cos_theta;
  // sqrt(fmax(0, x)) in double rounded to float equals sqrtf of the clamped value
  float sin_theta = sqrtf(x > 0.0f ? x : 0.0f);
  V3 v1, v2;
  if (fabsf(wo.x) > fabsf(wo.y))
    v1 = v3(-wo.z, 0.0f, wo.x) / sqrtf(wo.x * wo.x + wo.z * wo.z);
  else
    v1 = v3(0.0f, wo.z, -wo.y) / sqrtf(wo.y * wo.y + wo.z * wo.z);
  v2 = cross(wo, v1);
  float sp, cp;
  flg_sincosf(phi, &sp, &cp);
  return (v1 * (sin_theta * cp) + v2 * (sin_theta * sp)) + wo * cos_theta;
}

// Environment light tables (Dist2D of InfiniteAreaLight, light.cpp:93-108,
// distribution.hpp:12-92), 512 x 512.
// Free-flight walks.  A Halton index of 0 (pixel offset 0, sample 0) makes
// every dimension >= 2 equal to 0, so t never advances and the reference's
// loops never end.  A walk therefore ends at a zero draw (only index 0 gives
// one), and in any case after kMaxTrackSteps steps: ratio tracking returns the
// current transmittance, delta tracking reports no collision.
constexpr int kMaxTrackSteps = 4096;
constexpr float kInf = __builtin_huge_valf();

// Draw source of the trackers on the device: the path's Halton stream.  The
// host drop-ins instantiate the same trackers over a host-side sampler.
struct HaltonDraw {
  PathSampler& ps;
  const HaltonTables& tab;
  const HaltonParams& hp;
  FLG_D float next() { return ps.next(tab, hp); }
};

// Ratio tracking with Russian roulette (OpenVDBVolume::tr, volume.cpp:75-107)
template <bool COUNT, class Draw>
FLG_HD float vol_tr(const VolDev& v, Draw& rng, V3 o, V3 d, float ray_tmax, uint32_t& nlook) {
  float t_min, t_max;
  if (!box_pbrt(v.bmin, v.bmax, o, d, ray_tmax, &t_min, &t_max)) return 1.0f;
  float tr = 1.0f, t = smax(t_min, 0.0f);
  t_max = smin(t_max, ray_tmax);
  const float max_mu = v.max_mu;
  if (!(max_mu > 0.0f)) return 1.0f;  // empty medium: no interaction
  for (int step = 0; step < kMaxTrackSteps; ++step) {
    const float u = rng.next();
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / max_mu;
    if (t >= t_max) break;
    V3 pos = o + d * t;
    float density = lookup(v, pos).rho;
    if (COUNT) ++nlook;
    tr *= 1.0f - smax(0.0f, density * v.inv_max_density);
    if (tr < 0.1f) {
      float q = smax(0.05f, 1.0f - tr);
      if (rng.next() < q) return 0.0f;
      tr /= 1.0f - q;
    }
  }
  return tr;
}

// Delta tracking (OpenVDBVolume::sample, volume.cpp:109-148)
template <bool COUNT, class Draw>
FLG_HD float vol_sample(const VolDev& v, Draw& rng, V3 o, V3 d, float ray_tmax, bool* success,
                        V3* pos_out, uint32_t& nlook) {
  float t_min, t_max;
  if (!box_pbrt(v.bmin, v.bmax, o, d, ray_tmax, &t_min, &t_max)) {
    *success = false;
    return 1.0f;
  }
  float t = smax(t_min, 0.0f);
  const float max_mu = v.max_mu;
  t_max = smin(t_max, ray_tmax);
  if (!(max_mu > 0.0f)) {
    *success = false;
    return 1.0f;
  }
  for (int step = 0; step < kMaxTrackSteps; ++step) {
    const float u = rng.next();
    if (u == 0.0f) break;  // a zero draw never advances t (Halton index 0): the walk ends
    t += -flg_logf(1.0f - u) / max_mu;
    if (t >= t_max) break;
    V3 pos = o + d * t;
    float density = lookup(v, pos).rho;
    if (COUNT) ++nlook;
    float p_s = v.sigma_s * density / max_mu;
    float p_a = v.sigma_a * density / max_mu;
    float p_n = 1.0f - p_s - p_a;
    if (rng.next() > p_n) {
      *success = true;
      *pos_out = pos;
      return v.sigma_s / v.sigma_t;
    }
  }
  *success = false;
  return 1.0f;
}

// Chunk of the volume-only kernels (svol and differentiable): their paths
// are short and many, so the shared counter's atomics show; 128 measured
// 7% faster than per-refill claiming on a 512^2 x 64 spp fBm frame, while
// the mesh kernels (long, uneven paths) keep per-refill claiming.
constexpr uint32_t kVolChunk = 128;

// Persistent lanes with path regeneration.  Each lane runs one path at a
// time as begin(p, state) followed by step(state) until either returns true
// (path complete), then finish(p, state) and the next unstarted path.  Path
// indices come from *next_path.  With CHUNK == 0 a wave claims exactly the
// indices its idle lanes need, one atomic per refill; with CHUNK >= 64 it
// claims CHUNK consecutive indices at a time and hands them out over several
// refills (fewer atomics on the shared counter, coarser balance at the end of
// the launch).  Every wave leaves the loop once its lanes have drawn indices
// past npaths and their paths are complete; each path runs exactly the code of
// a one lane per path launch, so results do not depend on the schedule.
template <class State, uint32_t CHUNK = 0, class Begin, class Step, class Finish>
FLG_D void regen_loop(uint32_t npaths, unsigned int* next_path, Begin begin, Step step, Finish finish) {
  static_assert(CHUNK == 0 || CHUNK >= 64, "a chunk must cover every lane of a refill");
  const int lane = threadIdx.x & 63;
  uint32_t cursor = 0, end = 0;  // CHUNK > 0: the wave's unclaimed indices [cursor, end), wave-uniform
  State s;
  uint32_t p = 0;
  bool need = true, exhausted = false;
  while (true) {
    const uint64_t want = __ballot(need && !exhausted);
    if (want) {
      const uint32_t nw = (uint32_t)__popcll(want);
      const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
      const uint32_t r = (uint32_t)__popcll(want & below);  // this lane's rank among those that want
      const int leader = __builtin_ctzll(want);
      uint32_t mine;
      if constexpr (CHUNK == 0) {
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(next_path, nw);
        base = __shfl(base, leader, 64);
        mine = base + r;
      } else {
        // the rest of the current chunk first, then a fresh chunk (nw <= 64 <= CHUNK)
        const uint32_t rem = end - cursor;
        uint32_t fresh = 0;
        if (rem < nw) {
          if (lane == leader) fresh = atomicAdd(next_path, CHUNK);
          fresh = __shfl(fresh, leader, 64);
        }
        mine = r < rem ? cursor + r : fresh + (r - rem);
        if (rem < nw) {
          cursor = fresh + (nw - rem);
          end = fresh + CHUNK;
        } else {
          cursor += nw;
        }
      }
      if (need && !exhausted) {
        p = mine;
        if (p >= npaths) {
          exhausted = true;
        } else {
          need = false;
          if (begin(p, s)) {
            finish(p, s);
            need = true;
          }
        }
      }
    }
    if (__ballot(!need) == 0 && __ballot(!exhausted) == 0) break;
    if (!need && step(s)) {
      finish(p, s);
      need = true;
    }
  }
}

struct EnvDev {
  const float* marg_cdf;  // 513
  const float* marg_f;    // 512
  const float* row_cdf;   // 512 x 513
  const float* row_f;     // 512 x 512
  const float* row_int;   // 512
  float marg_int;
  V3 radiance;  // RGB
  V3 center;
  float radius;
  int32_t present;
};

// Dist1D::inverseCDF / sampleC (distribution.hpp:26-50), r = mid - 1 kept.
FLG_D float dist_sample(const float* cdf, const float* f, float integral, int n, float u,
                        float* pdf, int* off) {
  int l = 0, r = n;
  while (r - l > 1) {
    int mid = (l + r) / 2;
    if (cdf[mid] <= u) l = mid;
    else r = mid - 1;
  }
  *off = l;
  float du = u - cdf[l];
  if ((cdf[l + 1] - cdf[l]) > 0) du /= (cdf[l + 1] - cdf[l]);
  *pdf = f[l] / integral;
  return ((float)l + du) / (float)n;
}

// InfiniteAreaLight::sampleLi (light.cpp:119-145) with a constant texture,
// up to its radiance (the texture's colour, the same in every direction).
FLG_D void env_sample_dir(const EnvDev& e, V3 ref, float u0, float u1, V3* wi, float* pdf, V3* p_light) {
  float p1, p0;
  int o1, o0;
  float d1 = dist_sample(e.marg_cdf, e.marg_f, e.marg_int, 512, u1, &p1, &o1);
  float d0 = dist_sample(e.row_cdf + 513 * o1, e.row_f + 512 * o1, e.row_int[o1], 512, u0, &p0, &o0);
  *pdf = p0 * p1;
  float phi = d0 * 2.0f * PI;
  float theta = d1 * PI;
  float st, ct, sp, cp;
  flg_sincosf(theta, &st, &ct);
  flg_sincosf(phi, &sp, &cp);
  V3 dir = v3(st * cp, st * sp, ct);
  *pdf /= (2.0f * PI * PI * smax(st, 1e-4f));
  dir = v3(-dir.y, dir.z, -dir.x);
  V3 p = e.center + (dir * 2.0f) * e.radius;
  *p_light = p;
  *wi = normalize(p - ref);
}

struct LightsDev {
  EnvDev env;
  int32_t has_point;
  V3 point_pos;
  V3 point_intensity;  // RGB
};

// UniformSampleOneLight + EstimateDirect for a volume vertex
// (integrator.cpp:69-120): lights in the order [point, environment], one
// picked with p = 1/n, u_light drawn before sampleLi, the HG phase as f.
// Radiance is RGB and every channel runs the reference's Vector3f arithmetic
// (vector.inc: component-wise products and quotients).  light_pre runs it up
// to the shadow ray: the light choice and its sample.  kind 0: no light (the
// estimate is 0); 1: no shadow ray (pdf or the phase function is 0, or Li is
// 0 in every channel; the estimate is 0 / light_pdf); 2: the estimate needs
// the transmittance along the shadow ray from p towards p_light (light_post).
// The sample keeps the point light's squared distance rather than Li: Li is
// I / d2 per channel (or the environment's colour), recomputed where it is
// used, so that a path carries one float for it instead of three.
struct LightSample {
  V3 p_light;
  float f, d2, pdf, light_pdf;  // d2 < 0: the environment light
  int kind;
};

// Li of the sampled light, per channel.
FLG_D V3 light_li(const LightsDev& L, float d2) {
  if (d2 < 0.0f) return L.env.radiance;
  return v3(L.point_intensity.x / d2, L.point_intensity.y / d2, L.point_intensity.z / d2);
}

FLG_D LightSample light_pre(const LightsDev& L, float g, const HaltonTables& tab, const HaltonParams& hp,
                            PathSampler& ps, V3 p, V3 wo) {
  LightSample ls;
  ls.kind = 0;
  const int n_lights = (L.has_point ? 1 : 0) + (L.env.present ? 1 : 0);
  if (n_lights == 0) return ls;
  int light_num = (int)(ps.next(tab, hp) * (float)n_lights);
  if (light_num > n_lights - 1) light_num = n_lights - 1;
  ls.light_pdf = (float)(1.0 / (double)n_lights);
  const bool use_point = L.has_point && light_num == 0;
  float u0 = ps.next(tab, hp);
  float u1 = ps.next(tab, hp);
  V3 wi;
  if (use_point) {
    V3 diff = L.point_pos - p;
    ls.d2 = sqnorm(diff);
    wi = normalize(diff);
    ls.pdf = 1.0f;
    ls.p_light = L.point_pos;
  } else {
    env_sample_dir(L.env, p, u0, u1, &wi, &ls.pdf, &ls.p_light);
    ls.d2 = -1.0f;
  }
  ls.kind = 1;
  const V3 Li = light_li(L, ls.d2);
  if (ls.pdf == 0.0f || (Li.x == 0.0f && Li.y == 0.0f && Li.z == 0.0f)) return ls;
  ls.f = hg_p(wi, wo, g);
  if (ls.f != 0.0f) ls.kind = 2;
  return ls;
}

// The estimate without a shadow ray (kinds 0 and 1).
FLG_D V3 light_unshadowed(const LightSample& ls) {
  const float z = ls.kind == 0 ? 0.0f : 0.0f / ls.light_pdf;
  return v3(z, z, z);
}

// The shadow ray of a kind-2 sample from p: direction and length.
FLG_D void light_ray(const LightSample& ls, V3 p, V3* d, float* tmax) {
  V3 diff = ls.p_light - p;
  float nrm = sqrtf(sqnorm(diff));
  *d = diff / nrm;
  *tmax = nrm - SHADOW_EPS;
}

// The estimate of a kind-2 sample given the shadow ray's transmittance:
// Li *= tr; L += f * Li / pdf; L / light_pdf, per channel.
FLG_D float light_post1(float Li, float f, float pdf, float light_pdf, float tr) {
  Li *= tr;
  float out = 0.0f;
  out += f * Li / pdf;
  return out / light_pdf;
}

FLG_D V3 light_post(const LightsDev& L, const LightSample& ls, float tr) {
  const V3 Li = light_li(L, ls.d2);
  return v3(light_post1(Li.x, ls.f, ls.pdf, ls.light_pdf, tr), light_post1(Li.y, ls.f, ls.pdf, ls.light_pdf, tr),
            light_post1(Li.z, ls.f, ls.pdf, ls.light_pdf, tr));
}

// L += beta * ld, per channel.
FLG_D void add_scaled(V3& L, float beta, V3 ld) {
  L.x += beta * ld.x;
  L.y += beta * ld.y;
  L.z += beta * ld.z;
}

// The whole estimate; `shadow(p, d, tmax)` is the transmittance along the
// shadow ray: ratio tracking in the medium (SVolIntegrator) or VolEstimateTr
// through interfaces (VolPathIntegrator).
template <class Shadow>
FLG_D V3 sample_one_light(const LightsDev& L, float g, const HaltonTables& tab, const HaltonParams& hp,
                          PathSampler& ps, V3 p, V3 wo, Shadow shadow) {
  const LightSample ls = light_pre(L, g, tab, hp, ps, p, wo);
  if (ls.kind != 2) return light_unshadowed(ls);
  V3 d;
  float tmax;
  light_ray(ls, p, &d, &tmax);
  return light_post(L, ls, shadow(p, d, tmax));
}

// Camera (src/common/camera.h:11-31): ctor on the host in float, per-ray
// offsets through the double literal as in generateRay.
struct CamDevV {
  V3 pos, forward, right, up2;
  float fov_ratio, aspect_fov;
  int32_t w, h;
};

FLG_D void gen_ray_v(const CamDevV& c, float x, float y, V3* o, V3* d) {
  float y_offset = (float)((double)c.fov_ratio * ((double)((y / (float)c.h) * 2.0f) - 1.0));
  float x_offset = (float)((double)c.aspect_fov * ((double)((x / (float)c.w) * 2.0f) - 1.0));
  V3 dir = (c.forward + c.right * x_offset) + c.up2 * y_offset;
  *o = c.pos;
  *d = normalize(dir);
}

inline CamDevV make_cam_v(const flg_camera* c, int w, int h) {
  CamDevV cd;
  cd.pos = v3(c->pos[0], c->pos[1], c->pos[2]);
  V3 look = v3(c->look_at[0], c->look_at[1], c->look_at[2]);
  V3 up = v3(c->up[0], c->up[1], c->up[2]);
  cd.forward = normalize(look - cd.pos);
  cd.right = normalize(cross(up, cd.forward));
  cd.up2 = normalize(cross(cd.forward, cd.right));
  cd.fov_ratio = tanf(PI * c->fov_deg / 360.0f);
  float aspect = (float)w / (float)h;
  cd.aspect_fov = cd.fov_ratio * aspect;
  cd.w = w;
  cd.h = h;
  return cd;
}

VolDev make_voldev(const flg_volume* vol);
// Pixel = the in-order float mean of its spp RGB path samples (three floats
// per sample, pixel-major), per channel, into rgb (npix*3).
void pixel_mean(hipStream_t st, const float* lpath, int64_t npix, int32_t spp, float* rgb);
// The film of the reference's render loop from the film tasks' RGB radiance
// (three floats per task) and film positions (tasks per pixel, pixel-major),
// into rgb (w*h*3); lpath NULL
// or tasks 0: nothing reached the film, rgb is 0.
void film_resolve(hipStream_t st, const float* lpath, const float2* fpos, int32_t w, int32_t h, int32_t tasks,
                  float* rgb);
// scene_bounds (may be NULL: the volume's box): the scene bound that places the
// environment light's sample points (InfiniteAreaLight::preprocess).
// may_build false: an environment table not yet built for the scene's
// radiance is an error (FLG_ERR_INVALID) instead of being built (which
// allocates and uploads).
int make_lights(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* sc, LightsDev* L,
                const float* scene_bounds = nullptr, bool may_build = true);

}  // namespace flg
