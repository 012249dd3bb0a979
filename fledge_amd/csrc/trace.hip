// Closest-hit traversal, primary rays and the wavefront diffuse path tracer.
//
// Traversal order and tie rule are those of RadixBVHBuilder::recursiveIntersect
// (src/experimental/radix_bvh.cpp:202-245): depth first from the root, a
// node's leaf children are tested first (left then right, an equal t
// replaces the current hit), then the left internal subtree, then the right
// one; a node is skipped when its box is missed or best < tnear.  Child boxes
// are stored in the parent (64-byte node), so a child's box test runs when its
// parent is visited; a child that would be skipped at its visit is not pushed
// (best only decreases, so it would be skipped later too) and pushed entries
// are re-checked against best when popped.  The result is therefore
// identical to the recursive traversal.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "bvh.hpp"
#include "halton.hpp"
#include "traverse.hpp"

namespace flg {

namespace {

// The traversal stack of a lane: RING > 0 keeps its RING most recent entries
// in LDS (ring slot k of the lane at ring[k * stride]) and spills older ones
// to priv; RING == 0 keeps all of them in priv.
template <int RING>
using StackOf = typename std::conditional<(RING > 0), RingStack<(RING > 0 ? RING : 1)>, PrivStack>::type;
template <int RING>
FLG_D StackOf<RING> make_stack(uint2* ring, int stride, uint2* priv) {
  if constexpr (RING > 0) return RingStack<RING>(ring, stride, priv);
  else return PrivStack(priv);
}

// Traversal is latency-bound: k_trace and k_pt_bounce ask for 8 waves per SIMD
// (at most 64 VGPRs); the few values that no longer fit spill to scratch, which
// measured faster than the 4-5 waves the unconstrained allocation allows.
constexpr int kTraceWaves = 8;

template <bool COMPAT, int BS, int RING>
__global__ __launch_bounds__(BS, kTraceWaves) void k_trace(TraceGeom g, const flg_ray* rays, flg_hit* hits,
                                              int64_t n) {
  __shared__ uint2 ring[RING > 0 ? BS * RING : 1];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* r = reinterpret_cast<const float4*>(rays + i);
  float4 a = r[0], b = r[1];
  float best = a.w;
  Counters c{0, 0};
  int32_t p = trace_lane<COMPAT, false, BS, RING>(g, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), best, c, ring);
  flg_hit h;
  h.id = p >= 0 ? (int32_t)f2u(g.tris[p].v0.w) : -1;
  h.t = p >= 0 ? best : a.w;
  hits[i] = h;
}

template <bool COMPAT>
__global__ __launch_bounds__(256) void k_trace_count(TraceGeom g, const flg_ray* rays, int64_t n,
                                                     unsigned long long* counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Counters c{0, 0};
  if (i < n) {
    const float4* r = reinterpret_cast<const float4*>(rays + i);
    float4 a = r[0], b = r[1];
    float best = a.w;
    (void)trace_one<COMPAT, true>(g, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), best, c);
  }
  unsigned long long cn = c.nodes, ct = c.tris;
  for (int o = 32; o >= 1; o >>= 1) {
    cn += __shfl_xor(cn, o, 64);
    ct += __shfl_xor(ct, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(counts, cn);
    atomicAdd(counts + 1, ct);
  }
}

// ----------------------------------------------------------------- camera --
struct CamDev {
  V3 pos, forward, right, up2;
  float fov_ratio, aspect_fov;  // fov_ratio, fov_ratio * aspect
  int32_t w, h;
};

// Camera ctor (src/common/camera.h:12-21) evaluated on the host in float.
CamDev make_cam(const flg_camera* c, int w, int h) {
  CamDev cd;
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

// Camera::generateRay (camera.h:21-31): the offsets are products with the
// double literal 1.0, rounded back to float.
FLG_D void gen_ray(const CamDev& c, float x, float y, V3* o, V3* d) {
  float y_offset = (float)((double)c.fov_ratio * ((double)((y / (float)c.h) * 2.0f) - 1.0));
  float x_offset = (float)((double)c.aspect_fov * ((double)((x / (float)c.w) * 2.0f) - 1.0));
  V3 dir = (c.forward + c.right * x_offset) + c.up2 * y_offset;
  *o = c.pos;
  *d = normalize(dir);
}

__global__ __launch_bounds__(256) void k_camera(CamDev cam, HaltonTables tab, HaltonParams hp,
                                                uint32_t sample, flg_ray* rays) {
  int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= (int64_t)cam.w * cam.h) return;
  int x = (int)(pix % cam.w), y = (int)(pix / cam.w);
  PathSampler ps;
  ps.index = halton_pixel_offset(hp, x, y) + sample * hp.stride;
  ps.dim = 0;
  float ux = ps.next(tab, hp);
  float uy = ps.next(tab, hp);
  V3 o, d;
  gen_ray(cam, (float)x + ux, (float)y + uy, &o, &d);
  float4* r = reinterpret_cast<float4*>(rays + pix);
  r[0] = make_float4(o.x, o.y, o.z, __builtin_inff());
  r[1] = make_float4(d.x, d.y, d.z, 0.0f);
}

// ------------------------------------------------------------ path tracer --
// ConcentricSampleDisk / CosineSampleHemisphere (math_utils.h:69-88)
FLG_D V3 cosine_hemisphere(float u0, float u1) {
  float ux = (float)((double)u0 * 2.0 - 1.0), uy = (float)((double)u1 * 2.0 - 1.0);
  float dx = 0.0f, dy = 0.0f;
  if (!(ux == 0 && uy == 0)) {
    float theta, r;
    if (fabsf(ux) > fabsf(uy)) {
      r = ux;
      theta = PI_OVER4 * (uy / ux);
    } else {
      r = uy;
      theta = PI_OVER2 - (PI_OVER4 * (ux / uy));
    }
    float s, c;
    flg_sincosf(theta, &s, &c);
    dx = c * r;
    dy = s * r;
  }
  float z = sqrtf(smax(0.0f, 1.0f - dx * dx - dy * dy));
  return v3(dx, dy, z);
}

struct QRay {  // 32 bytes: origin + path id, direction + Halton index
  float4 o;    // w: path id bits
  float4 d;    // w: Halton index bits
};

struct PtState {
  QRay* q[2];
  uint32_t* count;  // [max_depth + 2] queue sizes
  uint32_t* taken;  // [max_depth + 1] chunks of each bounce's queue taken so far
  uint2* hits;      // split bounces: closest hit (sorted position, t bits) per queued ray
  int32_t* nseg;
  int32_t* seg_id;
  float* seg_t;
  int32_t stride_seg;
};

__global__ __launch_bounds__(256) void k_pt_zero(uint32_t* count, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) count[i] = 0;
}

// CoordinateTransition (src/materials/material_base.hpp:14-41)
struct Frame {
  V3 n, t, b;
  FLG_D explicit Frame(V3 normal) {
    n = normal / sqrtf(sqnorm(normal));
    if (fabsf(n.x) > fabsf(n.z)) b = v3(-n.y, n.x, 0.0f);
    else b = v3(0.0f, -n.z, n.y);
    b = b / sqrtf(sqnorm(b));
    t = cross(b, n);
  }
  FLG_D V3 to_local(V3 p) const { return v3(dot(p, t), dot(p, b), dot(p, n)); }
  FLG_D V3 to_world(V3 p) const { return (t * p.x + b * p.y) + n * p.z; }
};

// After the closest hit (p, t) of segment `bounce` of a path: the segment
// record, and on a hit the diffuse sampling in the shading frame
// (PathIntegrator::Li with a diffuse material, integrator.cpp:252-326) that
// makes the next segment.  Returns true when nr holds that segment.
template <bool COMPAT>
FLG_D bool pt_finish(const TraceGeom& g, const PtState& s, const HaltonTables& tab, const HaltonParams& hp,
                     int bounce, int max_depth, V3 o, V3 d, uint32_t path, uint32_t hidx, int32_t p,
                     float t, QRay& nr) {
  BTri tri;
  if (p >= 0) tri = g.tris[p];
  if (s.seg_id) {
    int64_t k = (int64_t)path * s.stride_seg + bounce;
    s.seg_id[k] = p >= 0 ? (int32_t)f2u(tri.v0.w) : -1;
    s.seg_t[k] = t;
  }
  bool spawn = false;
  if (p >= 0 && bounce < max_depth) {
    const V3 v0 = v3(tri.v0.x, tri.v0.y, tri.v0.z) - o, v1 = v3(tri.v1.x, tri.v1.y, tri.v1.z) - o,
             v2 = v3(tri.v2.x, tri.v2.y, tri.v2.z) - o;
    const V3 ng = stable_normalize(stable_triangle_normal(v2 - v0, v0 - v1, v1 - v2), COMPAT);
    const V3 hp3 = o + d * t;
    Frame fr(ng);
    PathSampler ps;
    ps.index = hidx;
    ps.dim = 2 + 2 * (uint32_t)bounce;
    float u0 = ps.next(tab, hp);
    float u1 = ps.next(tab, hp);
    V3 wo_l = fr.to_local(-d);
    V3 wi_l = cosine_hemisphere(u0, u1);
    float pdf = wo_l.z * wi_l.z > 0 ? fabsf(wi_l.z) * INV_PI : 0.0f;
    V3 wi = fr.to_world(wi_l);
    if (pdf != 0.0f) {
      V3 no = offset_ray_origin(hp3, ng, wi);
      V3 nd = normalize(wi);
      nr.o = make_float4(no.x, no.y, no.z, u2f(path));
      nr.d = make_float4(nd.x, nd.y, nd.z, u2f(hidx));
      spawn = true;
    }
  }
  if (!spawn && s.nseg) s.nseg[path] = bounce + 1;
  return spawn;
}

// The first bounce takes the camera paths in tile order (tile_pixel) of 8 x 64
// tiles: the paths in flight then cover a compact part of the image, whose
// rays cross the same nodes, and the next bounce's queue, appended in that
// order, follows it.  Measured on 1920 x 1080 at 16 spp: 10.0 ms per frame
// against 10.5 ms in scanline order (8 x 32, 16 x 64 and 16 x 256 tiles
// 10.1-10.3 ms; tiles 128 rows high no faster than scanlines).
constexpr uint32_t kTileW = 8, kTileH = 64;

// Wave-aggregated append of the lanes' new segments to the next queue.
FLG_D void pt_append(const PtState& s, int bounce, bool spawn, const QRay& nr) {
  const uint64_t m = __ballot(spawn);
  if (m) {
    const int lane = threadIdx.x & 63;
    uint32_t wbase = 0;
    if (lane == 0) wbase = atomicAdd(&s.count[bounce + 1], (uint32_t)__popcll(m));
    wbase = __shfl(wbase, 0, 64);
    if (spawn) {
      const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
      s.q[(bounce + 1) & 1][wbase + (uint32_t)__popcll(m & lt)] = nr;
    }
  }
}

// One bounce of every live path: (bounce 0) camera ray, closest hit, segment
// record, diffuse sampling in the shading frame and the spawn of the next
// segment into the other queue (wave-aggregated append).  Persistent grid:
// the queue length is read from device memory, so the host never waits.
template <bool COMPAT, bool FIRST, bool COUNT, int BS, int RING>
__global__ __launch_bounds__(BS, kTraceWaves) void k_pt_bounce(TraceGeom g, PtState s, CamDev cam,
                                                  HaltonTables tab, HaltonParams hp, int32_t spp,
                                                  int bounce, int max_depth,
                                                  unsigned long long* visits) {
  __shared__ uint2 ring[RING > 0 ? BS * RING : 1];
  const uint32_t n = FIRST ? (uint32_t)((int64_t)cam.w * cam.h * spp) : s.count[bounce];
  if (FIRST && blockIdx.x == 0 && threadIdx.x == 0) s.count[0] = n;
  const QRay* qin = s.q[bounce & 1];
  Counters c{0, 0};
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool spawn = false;
    QRay nr;
    if (i < n) {
      V3 o, d;
      uint32_t path, hidx;
      if (FIRST) {
        const uint32_t k = i / (uint32_t)spp, smp = i - k * (uint32_t)spp;
        uint32_t ux0, uy0;
        tile_pixel<kTileW, kTileH>(k, (uint32_t)cam.w, (uint32_t)cam.h, ux0, uy0);
        const int x = (int)ux0, y = (int)uy0;
        path = (uy0 * (uint32_t)cam.w + ux0) * (uint32_t)spp + smp;  // pixel-major, as the outputs
        PathSampler ps;
        ps.index = halton_pixel_offset(hp, x, y) + smp * hp.stride;
        ps.dim = 0;
        float ux = ps.next(tab, hp);
        float uy = ps.next(tab, hp);
        gen_ray(cam, (float)x + ux, (float)y + uy, &o, &d);
        hidx = ps.index;
      } else {
        const QRay r = qin[i];
        o = v3(r.o.x, r.o.y, r.o.z);
        d = v3(r.d.x, r.d.y, r.d.z);
        path = f2u(r.o.w);
        hidx = f2u(r.d.w);
      }
      float t = __builtin_inff();
      const int32_t p = trace_lane<COMPAT, COUNT, BS, RING>(g, o, d, t, c, ring);
      spawn = pt_finish<COMPAT>(g, s, tab, hp, bounce, max_depth, o, d, path, hidx, p, t, nr);
    }
    pt_append(s, bounce, spawn, nr);
  }
  if (COUNT) {
    unsigned long long cn = c.nodes, ct = c.tris;
    for (int o = 32; o >= 1; o >>= 1) {
      cn += __shfl_xor(cn, o, 64);
      ct += __shfl_xor(ct, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(visits, cn);
      atomicAdd(visits + 1, ct);
    }
  }
}

// Bounces after the first, split in two launches (ray_fetch 2).  The
// traversal kernel hands out the queue (its length is read on the device) in
// chunks of kChunk rays, in queue order, through one counter per bounce: the
// rays in flight at any moment then lie close together in the queue, which
// follows the image, so they cross the same nodes and share the caches.
// Within a chunk a lane whose ray has ended takes the next entry whenever
// enough of the wave's lanes are idle, so a wave does not idle behind its
// longest traversal.  It only writes each ray's closest hit.  The shading
// kernel then records the segments and spawns the next ones, one lane per
// ray, with the code of the fused kernel: results are the same bits.
constexpr int kRefillIdle = 16;   // refill once this many lanes of a wave are idle
constexpr uint32_t kChunk = 128;  // rays per counter increment

// Takes the wave's next chunk [next, end) of a queue of n rays; false when
// none is left.  Called by every lane of the wave together.
FLG_D bool take_chunk(uint32_t* taken, uint32_t n, uint32_t& next, uint32_t& end) {
  uint32_t c = 0;
  if ((threadIdx.x & 63) == 0) c = atomicAdd(taken, 1u);
  c = __builtin_amdgcn_readfirstlane(c);  // a scalar: the chunk bounds stay in SGPRs
  if (c >= (uint32_t)(((uint64_t)n + kChunk - 1) / kChunk)) return false;
  next = c * kChunk;
  end = next + kChunk < n ? next + kChunk : n;
  return true;
}

// The split traversal fetches a node's two leaf triangles together
// (trace_visit's PAIR), which needs more than 64 registers.  On the 2M-
// triangle frame the bounces after the first took 5.77-5.83 ms at 7 waves per
// SIMD (72 registers) against 6.16-6.22 ms at 8 waves without PAIR (6 waves
// with it: 5.94-5.98 ms; 8 waves with it spilled: 7.67 ms).
constexpr int kSplitTraceWaves = 7;

template <bool COMPAT, int BS, int RING>
__global__ __launch_bounds__(BS, kSplitTraceWaves) void k_pt_trace(TraceGeom g, PtState s, int bounce) {
  __shared__ uint2 ring[RING > 0 ? BS * RING : 1];
  const uint32_t n = s.count[bounce];
  const QRay* qin = s.q[bounce & 1];
  const int lane = threadIdx.x & 63;
  uint32_t next = 0, end = 0;  // the wave's current chunk; the same in every lane
  bool more = true;            // chunks may be left
  uint2 priv[kStack];
  auto st = make_stack<RING>(ring + threadIdx.x, BS, priv);
  Counters c{0, 0};
  TraceRay r;
  uint32_t cur = 0;
  bool have = false, active = false;
  // the counter only grows, so a wave's takes end; every wave leaves once no
  // chunk is left and all its lanes are idle; each traversal ends (a node is
  // visited at most once)
  while (true) {
    if (next >= end && more) more = take_chunk(s.taken + bounce, n, next, end);
    const uint64_t idle = __ballot(!have);
    const uint64_t live = ~idle & __ballot(true);
    if (next < end && idle != 0 && (live == 0 || __popcll(idle) >= kRefillIdle)) {
      if (!have) {
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint32_t i = next + (uint32_t)__popcll(idle & below);
        if (i < end) {
          const QRay q = qin[i];
          st.reset();
          cur = i;
          have = true;
          active = trace_begin<COMPAT, false>(g, v3(q.o.x, q.o.y, q.o.z), v3(q.d.x, q.d.y, q.d.z),
                                              __builtin_inff(), r, c);
        }
      }
      next += (uint32_t)__popcll(idle);
    } else if (live == 0) {
      break;
    }
    if (have) {
      if (active) active = trace_visit<COMPAT, false, true>(g, r, st, c);
      if (!active) {
        s.hits[cur] = make_uint2((uint32_t)r.hit, f2u(r.best));
        have = false;
      }
    }
  }
}

// Shading of the split bounces.  The new segments of a block's 1024 lanes are
// appended with one atomic per block (per-wave counts summed in LDS), since
// with nothing to overlap them one atomic per wave on the one queue counter
// would serialise.  Only a hit before the last bounce samples a new segment
// (about a fifth of bounce 1's rays): those are gathered onto the block's
// first lanes, which keeps the sampling on full waves (0.06 ms less per
// frame on the 2M-triangle frame; lane activity had been 0.08).
constexpr int kShadeBlock = 1024;

template <bool COMPAT>
__global__ __launch_bounds__(kShadeBlock) void k_pt_shade(TraceGeom g, PtState s, HaltonTables tab,
                                                          HaltonParams hp, int bounce, int max_depth) {
  __shared__ uint32_t wave_n[kShadeBlock / 64];
  __shared__ uint32_t block_base;
  __shared__ uint32_t gather_n[kShadeBlock / 64];
  __shared__ uint32_t hit_idx[kShadeBlock];
  const uint32_t n = s.count[bounce];
  const QRay* qin = s.q[bounce & 1];
  QRay* qout = s.q[(bounce + 1) & 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
    const uint32_t i = base + threadIdx.x;
    // a segment that cannot spawn (a miss, or the last bounce) is only
    // recorded; the others are gathered, in queue order, onto the block's
    // first lanes, so the costly sampling runs on full waves
    bool gather = false;
    if (i < n) {
      const uint2 h = s.hits[i];
      gather = (int32_t)h.x >= 0 && bounce < max_depth;
      if (!gather) {
        const QRay q = qin[i];
        QRay none;
        (void)pt_finish<COMPAT>(g, s, tab, hp, bounce, max_depth, v3(q.o.x, q.o.y, q.o.z),
                                v3(q.d.x, q.d.y, q.d.z), f2u(q.o.w), f2u(q.d.w), (int32_t)h.x, u2f(h.y), none);
      }
    }
    const uint64_t gm = __ballot(gather);
    if (lane == 0) gather_n[w] = (uint32_t)__popcll(gm);
    __syncthreads();
    uint32_t goff = 0, ngather = 0;
    for (int k = 0; k < kShadeBlock / 64; ++k) {
      const uint32_t c = gather_n[k];
      goff += k < w ? c : 0u;
      ngather += c;
    }
    if (gather) hit_idx[goff + (uint32_t)__popcll(gm & lt)] = i;
    __syncthreads();
    bool spawn = false;
    QRay nr;
    if (threadIdx.x < ngather) {
      const uint32_t j = hit_idx[threadIdx.x];
      const QRay q = qin[j];
      const uint2 h = s.hits[j];
      spawn = pt_finish<COMPAT>(g, s, tab, hp, bounce, max_depth, v3(q.o.x, q.o.y, q.o.z),
                                v3(q.d.x, q.d.y, q.d.z), f2u(q.o.w), f2u(q.d.w), (int32_t)h.x, u2f(h.y), nr);
    }
    const uint64_t m = __ballot(spawn);
    if (lane == 0) wave_n[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int k = 0; k < kShadeBlock / 64; ++k) tot += wave_n[k];
      block_base = tot ? atomicAdd(&s.count[bounce + 1], tot) : 0u;
    }
    __syncthreads();
    if (spawn) {
      uint32_t off = block_base;
      for (int k = 0; k < w; ++k) off += wave_n[k];
      qout[off + (uint32_t)__popcll(m & lt)] = nr;
    }
    __syncthreads();  // gather_n, hit_idx, wave_n and block_base are rewritten by the next trip
  }
}

// Bounces from bounce0 on in one persistent launch (ray_fetch 3: from the
// second segment of every path; ray_fetch 4: from the third, after the
// second's traversal and shading kernels).  Each wave owns a contiguous slice
// of bounce0's queue (at least kMinSlice rays: the tail's paths are long and
// scattered, and many small slices spread over the blocks balance them better
// than chunks taken in queue order); a lane takes a queued
// ray and follows its path to the end: when its traversal ends it records the
// segment and samples the diffuse bounce (pt_finish, the code of the other
// schedules, so the same bits) and traces the next segment itself, with no
// queue or launch in between.  Lanes whose traversals have ended wait until
// kFinishBatch of the wave's lanes (or all its traversing ones) are done, so
// the costly shading runs for many lanes at once; idle lanes refill from the
// slice as in k_pt_trace.  The segments traced after the queue's are counted
// into count[bounce + 1] for the frame's statistics.
constexpr int kFinishBatch = 16;
// rays per wave at least: short queues keep full waves (measured on the 2M-
// triangle frame's tail: 32 rays 6.53, 64 6.29-6.37, 128 6.19-6.22, 256
// 6.35-6.36, 512 6.93 ms for the bounces after the first)
constexpr uint32_t kMinSlice = 128;
// The kernel holds a path's whole state besides the traversal: it asks for 4
// waves per SIMD (128 VGPRs), which keeps the shading out of scratch.
constexpr int kPathWaves = 4;

template <bool COMPAT, int BS, int RING>
__global__ __launch_bounds__(BS, kPathWaves) void k_pt_paths(TraceGeom g, PtState s, HaltonTables tab,
                                                 HaltonParams hp, int bounce0, int max_depth) {
  __shared__ uint2 ring[RING > 0 ? BS * RING : 1];
  const uint32_t n = s.count[bounce0];
  const QRay* qin = s.q[bounce0 & 1];
  const int lane = threadIdx.x & 63;
  const uint32_t nwaves = gridDim.x * (BS / 64);
  const uint32_t wave = blockIdx.x * (BS / 64) + threadIdx.x / 64;
  uint32_t per = (uint32_t)(((uint64_t)n + nwaves - 1) / nwaves);
  per = per < kMinSlice ? kMinSlice : per;
  const uint64_t lo = (uint64_t)wave * per;
  const uint32_t end = (uint32_t)(lo + per < n ? lo + per : n);
  uint32_t next = (uint32_t)(lo < n ? lo : n);  // the same in every lane of the wave
  uint2 priv[kStack];
  auto st = make_stack<RING>(ring + threadIdx.x, BS, priv);
  Counters c{0, 0};
  TraceRay r;
  V3 o = v3(0.0f, 0.0f, 0.0f), d = v3(0.0f, 0.0f, 1.0f);
  uint32_t path = 0, hidx = 0, spawned = 0;
  int b = bounce0;
  bool have = false, active = false;
  // a wave leaves once its slice is taken and all its lanes are idle; every
  // trip traverses, shades or refills some lane, each traversal ends (a node
  // is visited at most once) and each path ends by max_depth
  while (true) {
    const uint64_t busy = __ballot(have && active);
    const uint64_t done = __ballot(have && !active);
    if (done != 0 && (busy == 0 || __popcll(done) >= kFinishBatch)) {
      if (have && !active) {
        QRay nr;
        if (pt_finish<COMPAT>(g, s, tab, hp, b, max_depth, o, d, path, hidx, r.hit, r.best, nr)) {
          o = v3(nr.o.x, nr.o.y, nr.o.z);
          d = v3(nr.d.x, nr.d.y, nr.d.z);
          ++b;
          ++spawned;
          st.reset();
          active = trace_begin<COMPAT, false>(g, o, d, __builtin_inff(), r, c);
        } else {
          have = false;
        }
      }
      continue;
    }
    const uint64_t idle = __ballot(!have);
    if (next < end && idle != 0 && (busy == 0 || __popcll(idle) >= kRefillIdle)) {
      if (!have) {
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint32_t i = next + (uint32_t)__popcll(idle & below);
        if (i < end) {
          const QRay q = qin[i];
          o = v3(q.o.x, q.o.y, q.o.z);
          d = v3(q.d.x, q.d.y, q.d.z);
          path = f2u(q.o.w);
          hidx = f2u(q.d.w);
          b = bounce0;
          st.reset();
          have = true;
          active = trace_begin<COMPAT, false>(g, o, d, __builtin_inff(), r, c);
        }
      }
      next += (uint32_t)__popcll(idle);
      continue;
    }
    if (busy == 0) break;  // nothing traversing, nothing to shade, nothing left to take
    if (have && active) active = trace_visit<COMPAT, false>(g, r, st, c);
  }
  // segments after the queue's own, for the frame's count
  uint32_t tot = spawned;
  for (int k = 32; k >= 1; k >>= 1) tot += __shfl_xor(tot, k, 64);
  if (lane == 0 && tot) atomicAdd(&s.count[bounce0 + 1], tot);
}

// Launch shapes: threads per block in {64, 128, 256, 512}; the counting
// variants (instrumentation only) exist at 256.
bool valid_block(int b) { return b == 64 || b == 128 || b == 256 || b == 512; }

// Stack policy of a launch: stack_mode 2 keeps the `ring` (4, 8, 16 or 32)
// most recent entries of each lane in LDS (RingStack), anything else the
// whole stack in scratch.  A block's ring takes block * ring * 8 bytes of
// LDS, at most 64 KiB: a ring of 32 needs blocks of at most 256 threads.
constexpr int kRingDefault = 8;
int ring_of(int mode, int ring) { return mode == 2 ? (ring ? ring : kRingDefault) : 0; }
bool valid_ring(int block, int ring) {
  return ring == 0 || ((ring == 4 || ring == 8 || ring == 16 || ring == 32) && block * ring * 8 <= 65536);
}

// Calls f with the ring depth as a compile-time constant (0: private stack).
template <class F>
void with_ring(int ring, F&& f) {
  switch (ring) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: f(std::integral_constant<int, 0>{}); break;
  }
}

template <bool COMPAT, int RING>
void launch_trace_r(int block, hipStream_t st, const TraceGeom& g, const flg_ray* rays, flg_hit* hits,
                    int64_t n) {
  const int grid = (int)div_up(n, block);
  switch (block) {
    case 64: k_trace<COMPAT, 64, RING><<<grid, 64, 0, st>>>(g, rays, hits, n); break;
    case 128: k_trace<COMPAT, 128, RING><<<grid, 128, 0, st>>>(g, rays, hits, n); break;
    case 512:
      if constexpr (RING <= 16) k_trace<COMPAT, 512, RING><<<grid, 512, 0, st>>>(g, rays, hits, n);
      break;
    default: k_trace<COMPAT, 256, RING><<<grid, 256, 0, st>>>(g, rays, hits, n); break;
  }
}

template <bool COMPAT>
void launch_trace(int block, int ring, hipStream_t st, const TraceGeom& g, const flg_ray* rays,
                  flg_hit* hits, int64_t n) {
  with_ring(ring, [&](auto R) { launch_trace_r<COMPAT, decltype(R)::value>(block, st, g, rays, hits, n); });
}

template <bool COMPAT, bool FIRST, int RING>
void launch_bounce_bs(int grid, int block, hipStream_t st, const TraceGeom& g, const PtState& s,
                      const CamDev& cam, const HaltonTables& tab, const HaltonParams& hp, int spp,
                      int b, int D, unsigned long long* visits) {
  switch (block) {
    case 64:
      k_pt_bounce<COMPAT, FIRST, false, 64, RING><<<grid, 64, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
      break;
    case 128:
      k_pt_bounce<COMPAT, FIRST, false, 128, RING><<<grid, 128, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
      break;
    case 512:
      if constexpr (RING <= 16)
        k_pt_bounce<COMPAT, FIRST, false, 512, RING><<<grid, 512, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
      break;
    default:
      k_pt_bounce<COMPAT, FIRST, false, 256, RING><<<grid, 256, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
      break;
  }
}

template <bool COMPAT, int RING>
void launch_split_bs(int grid, int block, hipStream_t st, const TraceGeom& g, const PtState& s, int b) {
  switch (block) {
    case 64: k_pt_trace<COMPAT, 64, RING><<<grid, 64, 0, st>>>(g, s, b); break;
    case 128: k_pt_trace<COMPAT, 128, RING><<<grid, 128, 0, st>>>(g, s, b); break;
    case 512:
      if constexpr (RING <= 16) k_pt_trace<COMPAT, 512, RING><<<grid, 512, 0, st>>>(g, s, b);
      break;
    default: k_pt_trace<COMPAT, 256, RING><<<grid, 256, 0, st>>>(g, s, b); break;
  }
}

template <bool COMPAT, int RING>
void launch_paths_bs(int grid, int block, hipStream_t st, const TraceGeom& g, const PtState& s,
                     const HaltonTables& tab, const HaltonParams& hp, int b, int D) {
  switch (block) {
    case 64: k_pt_paths<COMPAT, 64, RING><<<grid, 64, 0, st>>>(g, s, tab, hp, b, D); break;
    case 128: k_pt_paths<COMPAT, 128, RING><<<grid, 128, 0, st>>>(g, s, tab, hp, b, D); break;
    case 512:
      if constexpr (RING <= 16) k_pt_paths<COMPAT, 512, RING><<<grid, 512, 0, st>>>(g, s, tab, hp, b, D);
      break;
    default: k_pt_paths<COMPAT, 256, RING><<<grid, 256, 0, st>>>(g, s, tab, hp, b, D); break;
  }
}

// One bounce of the frame (fetch 1: the fused kernel; fetch 2: traversal
// with refill, then shading) or every remaining bounce in one launch (fetch 3
// at bounce 1, fetch 4 at bounce 2).
template <bool COMPAT>
void launch_bounce(bool first, bool count, int fetch, int grid, int block, int ring, hipStream_t st,
                   const TraceGeom& g, const PtState& s, const CamDev& cam, const HaltonTables& tab,
                   const HaltonParams& hp, int spp, int b, int D, unsigned long long* visits) {
  // the first bounce's camera rays are coherent: a wave's lanes stay in step
  // in the fused kernel, which refilling them one by one would undo
  if (((fetch == 3 && b >= 1) || (fetch == 4 && b >= 2)) && !count && !first) {
    with_ring(ring, [&](auto R) { launch_paths_bs<COMPAT, decltype(R)::value>(grid, block, st, g, s, tab, hp, b, D); });
    return;
  }
  if ((fetch == 2 || fetch == 4) && !count && !first) {
    with_ring(ring, [&](auto R) { launch_split_bs<COMPAT, decltype(R)::value>(grid, block, st, g, s, b); });
    const int sgrid = grid * block / kShadeBlock > 0 ? grid * block / kShadeBlock : 1;
    k_pt_shade<COMPAT><<<sgrid, kShadeBlock, 0, st>>>(g, s, tab, hp, b, D);
    return;
  }
  if (count) {  // instrumentation: one shape
    if (first)
      k_pt_bounce<COMPAT, true, true, 256, 0><<<grid, 256, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
    else
      k_pt_bounce<COMPAT, false, true, 256, 0><<<grid, 256, 0, st>>>(g, s, cam, tab, hp, spp, b, D, visits);
    return;
  }
  with_ring(ring, [&](auto R) {
    if (first)
      launch_bounce_bs<COMPAT, true, decltype(R)::value>(grid, block, st, g, s, cam, tab, hp, spp, b, D, visits);
    else
      launch_bounce_bs<COMPAT, false, decltype(R)::value>(grid, block, st, g, s, cam, tab, hp, spp, b, D, visits);
  });
}

// Milliseconds of one call of fn on the context stream: one warm-up, then the
// best of `reps` timed calls.
template <class F>
int time_call(flg_ctx* ctx, F fn, int reps, float* ms) {
  hipEvent_t a, b;
  FLG_CHECK_HIP(ctx, hipEventCreate(&a));
  FLG_CHECK_HIP(ctx, hipEventCreate(&b));
  int st = fn();
  float best = 1e30f;
  for (int r = 0; r < reps && st == FLG_OK; ++r) {
    (void)hipEventRecord(a, ctx->stream);
    st = fn();
    (void)hipEventRecord(b, ctx->stream);
    if (hipEventSynchronize(b) != hipSuccess) st = set_error(ctx, FLG_ERR_HIP, "autotune: event wait failed");
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    best = std::min(best, t);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = best;
  return st;
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_trace_closest(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays, flg_hit* d_hits,
                      int64_t n, const flg_trace_cfg* cfg) {
  return host_guard(ctx, "flg_trace_closest", [&]() -> int {
    if (!ctx || !bvh || n < 0 || (n > 0 && (!d_rays || !d_hits)))
      return set_error(ctx, FLG_ERR_INVALID, "flg_trace_closest: bad arguments");
    int block = cfg && cfg->block_size ? cfg->block_size : ctx->tuned_trace_block;
    if (block == 0) block = 256;
    int smode = cfg && cfg->stack_mode ? cfg->stack_mode : ctx->tuned_trace_stack;
    const int rdepth = cfg && cfg->stack_mode ? cfg->ring_depth : ctx->tuned_trace_ring;
    if (!valid_block(block) || smode < 0 || smode > 2)
      return set_error(ctx, FLG_ERR_INVALID,
                       "flg_trace_closest: block_size must be 64, 128, 256 or 512, stack_mode 0..2");
    const int ring = ring_of(smode, rdepth);
    if (!valid_ring(block, ring))
      return set_error(ctx, FLG_ERR_INVALID,
                       "flg_trace_closest: ring_depth must be 0, 4, 8, 16 or 32, with block * ring_depth * 8 "
                       "<= 65536 bytes of LDS");
    if (n == 0) return FLG_OK;
    if (div_up(n, 64) > (int64_t)INT32_MAX) return set_error(ctx, FLG_ERR_RANGE, "flg_trace_closest: too many rays");
    TraceGeom g = geom_of(bvh);
    {
      ScopedTimer tm(ctx, "trace_closest");
      if (bvh->flags & FLG_NORMAL_COMPAT)
        launch_trace<true>(block, ring, ctx->stream, g, d_rays, d_hits, n);
      else
        launch_trace<false>(block, ring, ctx->stream, g, d_rays, d_hits, n);
    }
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

int flg_trace_count(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays, int64_t n,
                    uint64_t* d_counts) {
  return host_guard(ctx, "flg_trace_count", [&]() -> int {
    if (!ctx || !bvh || n < 0 || !d_counts || (n > 0 && !d_rays))
      return set_error(ctx, FLG_ERR_INVALID, "flg_trace_count: bad arguments");
    TraceGeom g = geom_of(bvh);
    k_pt_zero<<<1, 256, 0, ctx->stream>>>((uint32_t*)d_counts, 4);
    if (n == 0) return FLG_OK;
    int grid = div_up(n, 256);
    if (bvh->flags & FLG_NORMAL_COMPAT)
      k_trace_count<true><<<grid, 256, 0, ctx->stream>>>(g, d_rays, n, (unsigned long long*)d_counts);
    else
      k_trace_count<false><<<grid, 256, 0, ctx->stream>>>(g, d_rays, n, (unsigned long long*)d_counts);
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

int flg_camera_rays(flg_ctx* ctx, const flg_camera* cam, int32_t width, int32_t height, int32_t spp,
                    int32_t sample, flg_ray* d_rays) {
  return host_guard(ctx, "flg_camera_rays", [&]() -> int {
    if (!ctx || !cam || width <= 0 || height <= 0 || spp <= 0 || sample < 0 || sample >= spp || !d_rays)
      return set_error(ctx, FLG_ERR_INVALID, "flg_camera_rays: bad arguments");
    HaltonParams hp = make_halton_params(width, height);
    if ((uint64_t)hp.stride * (uint64_t)spp > kMaxHaltonIndex)
      return set_error(ctx, FLG_ERR_RANGE, "flg_camera_rays: spp * Halton stride must stay at or below 2^32");
    HaltonTables tab;
    int st = get_halton_tables(ctx, &tab);
    if (st) return st;
    CamDev cd = make_cam(cam, width, height);
    int64_t npix = (int64_t)width * height;
    {
      ScopedTimer tm(ctx, "camera_rays");
      k_camera<<<div_up(npix, 256), 256, 0, ctx->stream>>>(cd, tab, hp, (uint32_t)sample, d_rays);
    }
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

int flg_pathtrace_diffuse(flg_ctx* ctx, const flg_bvh* bvh, const flg_camera* cam,
                          const flg_pt_cfg* cfg, int32_t* d_nseg, int32_t* d_seg_id,
                          float* d_seg_t, uint64_t* h_stats) {
  return host_guard(ctx, "flg_pathtrace_diffuse", [&]() -> int {
    if (!ctx || !bvh || !cam || !cfg || cfg->width <= 0 || cfg->height <= 0 || cfg->spp <= 0 ||
        cfg->max_depth < 0 || cfg->max_depth > 63 || ((d_seg_id == nullptr) != (d_seg_t == nullptr)))
      return set_error(ctx, FLG_ERR_INVALID, "flg_pathtrace_diffuse: bad arguments");
    int tb = cfg->trace_block ? cfg->trace_block : ctx->tuned_pt_block;
    if (tb == 0) tb = 256;
    int bpc = cfg->blocks_per_cu ? cfg->blocks_per_cu : ctx->tuned_pt_bpc;
    if (bpc == 0) bpc = 8;
    const int smode = cfg->stack_mode ? cfg->stack_mode : ctx->tuned_pt_stack;
    const int rdepth = cfg->stack_mode ? cfg->ring_depth : ctx->tuned_pt_ring;
    int fmode = cfg->ray_fetch ? cfg->ray_fetch : ctx->tuned_pt_fetch;
    if (fmode == 0) fmode = 1;
    if (!valid_block(tb) || bpc < 1 || bpc > 64 || smode < 0 || smode > 2 || fmode < 1 || fmode > 4)
      return set_error(ctx, FLG_ERR_INVALID,
                       "flg_pathtrace_diffuse: trace_block must be 64..512 (power of two), blocks_per_cu 1..64, "
                       "stack_mode 0..2 and ray_fetch 0..4");
    const int ring = ring_of(smode, rdepth);
    if (!valid_ring(tb, ring))
      return set_error(ctx, FLG_ERR_INVALID,
                       "flg_pathtrace_diffuse: ring_depth must be 0, 4, 8, 16 or 32, with trace_block * "
                       "ring_depth * 8 <= 65536 bytes of LDS");
    HaltonParams hp = make_halton_params(cfg->width, cfg->height);
    if ((uint64_t)hp.stride * (uint64_t)cfg->spp > kMaxHaltonIndex)
      return set_error(ctx, FLG_ERR_RANGE, "flg_pathtrace_diffuse: spp * Halton stride must stay at or below 2^32");
    int64_t npaths = (int64_t)cfg->width * cfg->height * cfg->spp;
    if (npaths >= (int64_t)1 << 31) return set_error(ctx, FLG_ERR_RANGE, "flg_pathtrace_diffuse: too many paths");
    HaltonTables tab;
    int st = get_halton_tables(ctx, &tab);
    if (st) return st;
    CamDev cd = make_cam(cam, cfg->width, cfg->height);
    const int D = cfg->max_depth;
    size_t qbytes = (size_t)npaths * sizeof(QRay);
    char* sc = (char*)scratch(ctx, "pt_queues", 2 * qbytes + (size_t)npaths * 8 + 4352, &st);
    if (!sc) return st;
    PtState s;
    s.q[0] = (QRay*)sc;
    s.q[1] = (QRay*)(sc + qbytes);
    s.hits = (uint2*)(sc + 2 * qbytes);
    s.count = (uint32_t*)(sc + 2 * qbytes + (size_t)npaths * 8);
    unsigned long long* visits = (unsigned long long*)(sc + 2 * qbytes + (size_t)npaths * 8 + 2048);
    s.taken = (uint32_t*)(sc + 2 * qbytes + (size_t)npaths * 8 + 4096);  // 64 entries
    const bool count_visits = cfg->count_visits != 0;
    s.nseg = d_nseg;
    s.seg_id = d_seg_id;
    s.seg_t = d_seg_t;
    s.stride_seg = D + 1;
    TraceGeom g = geom_of(bvh);
    const bool compat = bvh->flags & FLG_NORMAL_COMPAT;
    const int pgrid = ctx->num_cus * bpc;
    // the whole frame: clear the queue counts, then one fused launch per bounce
    // s.count[0 .. D+1]: queue sizes
    auto enqueue = [&](hipStream_t stream, bool timed) {
      k_pt_zero<<<1, 256, 0, stream>>>(s.count, 256);
      k_pt_zero<<<1, 256, 0, stream>>>(s.taken, 64);
      if (count_visits) k_pt_zero<<<1, 256, 0, stream>>>((uint32_t*)visits, 4);
      for (int b = 0; b <= D; ++b) {
        ScopedTimer tt(timed ? ctx : nullptr, "pathtrace_bounce");
        // the camera-ray bounce apart from the incoherent diffuse bounces
        ScopedTimer tk(timed ? ctx : nullptr, b == 0 ? "pathtrace_first_bounce" : "pathtrace_later_bounces");
        if (compat)
          launch_bounce<true>(b == 0, count_visits, fmode, pgrid, tb, ring, stream, g, s, cd, tab, hp, cfg->spp,
                              b, D, visits);
        else
          launch_bounce<false>(b == 0, count_visits, fmode, pgrid, tb, ring, stream, g, s, cd, tab, hp, cfg->spp,
                               b, D, visits);
        // that launch ran every remaining bounce
        if (!count_visits && ((fmode == 3 && b >= 1) || (fmode == 4 && b >= 2))) break;
      }
    };
    if (cfg->use_graph) {
      // every kernel argument of the frame is in the key: a cached graph is
      // replayed only for an identical launch sequence
      std::string key = "pathtrace:";
      auto put = [&key](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
      put(&g, sizeof g);
      put(&s, sizeof s);
      put(&cd, sizeof cd);
      put(&tab, sizeof tab);
      put(&hp, sizeof hp);
      const int32_t misc[9] = {cfg->spp, D, pgrid, tb, smode, fmode, compat ? 1 : 0, count_visits ? 1 : 0, ring};
      put(misc, sizeof misc);
      put(&visits, sizeof visits);
      ScopedTimer tm(ctx, "pathtrace_total");
      int st = graph_launch(ctx, key, [&](hipStream_t cs) { enqueue(cs, false); }, "path tracer");
      if (st) return st;
    } else {
      ScopedTimer tm(ctx, "pathtrace_total");
      enqueue(ctx->stream, true);
    }
    FLG_CHECK_LAUNCH(ctx);
    if (h_stats) {
      std::vector<uint32_t> cnt(D + 2);
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(cnt.data(), s.count, (D + 2) * 4, hipMemcpyDeviceToHost, ctx->stream));
      FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
      uint64_t tot = 0;
      for (int b = 0; b <= D; ++b) tot += cnt[b];
      h_stats[0] = tot;
      h_stats[1] = (uint64_t)npaths;
      if (count_visits) {
        unsigned long long v[2] = {0, 0};
        FLG_CHECK_HIP(ctx, hipMemcpyAsync(v, visits, 16, hipMemcpyDeviceToHost, ctx->stream));
        FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        h_stats[2] = v[0];
        h_stats[3] = v[1];
      }
    }
    return FLG_OK;
  });
}

int flg_autotune_trace(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays, flg_hit* d_hits,
                       int64_t n, flg_tuning* out) {
  return host_guard(ctx, "flg_autotune_trace", [&]() -> int {
    if (!ctx || !bvh || n <= 0 || !d_rays || !d_hits || !out)
      return set_error(ctx, FLG_ERR_INVALID, "flg_autotune_trace: bad arguments");
    const int blocks[4] = {64, 128, 256, 512};
    const int rings[5] = {0, 4, 8, 16, 32};  // 0: private stack
    int best_b = 256, best_s = 1, best_r = 0, ncand = 0;
    float best = 1e30f, def = 0.0f;
    for (int rg : rings)
      for (int b : blocks) {
        if (!valid_ring(b, rg)) continue;
        flg_trace_cfg c{};
        c.block_size = b;
        c.stack_mode = rg ? 2 : 1;
        c.ring_depth = rg;
        float ms = 0.0f;
        int st = time_call(ctx, [&] { return flg_trace_closest(ctx, bvh, d_rays, d_hits, n, &c); }, 3, &ms);
        if (st) return st;
        ++ncand;
        if (b == 256 && rg == 0) def = ms;
        if (ms < best) {
          best = ms;
          best_b = b;
          best_s = c.stack_mode;
          best_r = rg;
        }
      }
    ctx->tuned_trace_block = best_b;
    ctx->tuned_trace_stack = best_s;
    ctx->tuned_trace_ring = best_r;
    *out = flg_tuning{};
    out->block_size = best_b;
    out->stack_mode = best_s;
    out->ring_depth = best_r;
    out->ms = best;
    out->ms_default = def;
    out->candidates = ncand;
    return FLG_OK;
  });
}

int flg_autotune_pathtrace(flg_ctx* ctx, const flg_bvh* bvh, const flg_camera* cam,
                           const flg_pt_cfg* cfg, flg_tuning* out) {
  return host_guard(ctx, "flg_autotune_pathtrace", [&]() -> int {
    if (!ctx || !bvh || !cam || !cfg || !out)
      return set_error(ctx, FLG_ERR_INVALID, "flg_autotune_pathtrace: bad arguments");
    // two stages: ray fetch x stack (private, or an LDS ring of 4..32
    // entries) at two launch shapes, then every shape for the best of those:
    // resident threads per CU 1024 .. 8192, in blocks of 64 .. 512
    struct Cand {
      int b, bpc, sm, rg, fm;
    };
    int n = 0;
    float best = 1e30f, def = 0.0f;
    Cand bc{256, 8, 1, 0, 1};
    auto run = [&](const Cand& k, float* ms) -> int {
      flg_pt_cfg c = *cfg;
      c.trace_block = k.b;
      c.blocks_per_cu = k.bpc;
      c.stack_mode = k.sm;
      c.ring_depth = k.rg;
      c.ray_fetch = k.fm;
      c.count_visits = 0;
      c.use_graph = 0;
      return time_call(
          ctx, [&] { return flg_pathtrace_diffuse(ctx, bvh, cam, &c, nullptr, nullptr, nullptr, nullptr); }, 2, ms);
    };
    auto consider = [&](const Cand& k) -> int {
      if (!valid_ring(k.b, k.rg) || k.bpc < 1 || k.bpc > 64) return FLG_OK;
      float ms = 0.0f;
      int st = run(k, &ms);
      if (st) return st;
      ++n;
      if (k.b == 256 && k.bpc == 8 && k.sm == 1 && k.fm == 1) def = ms;
      if (ms < best) {
        best = ms;
        bc = k;
      }
      return FLG_OK;
    };
    const int rings[5] = {0, 4, 8, 16, 32};
    const int shapes[2][2] = {{256, 8}, {128, 64}};
    for (int fm = 1; fm <= 4; ++fm)
      for (int rg : rings)
        for (const auto& sh : shapes) {
          int st = consider(Cand{sh[0], sh[1], rg ? 2 : 1, rg, fm});
          if (st) return st;
        }
    const Cand stage1 = bc;
    const int blocks[4] = {64, 128, 256, 512};
    const int threads_per_cu[4] = {1024, 2048, 4096, 8192};
    for (int b : blocks)
      for (int t : threads_per_cu) {
        if ((b == 256 && t / b == 8) || (b == 128 && t / b == 64)) continue;  // timed in stage one
        int st = consider(Cand{b, t / b, stage1.sm, stage1.rg, stage1.fm});
        if (st) return st;
      }
    if (def == 0.0f) {  // the untuned shape, for reference
      int st = run(Cand{256, 8, 1, 0, 1}, &def);
      if (st) return st;
    }
    ctx->tuned_pt_block = bc.b;
    ctx->tuned_pt_bpc = bc.bpc;
    ctx->tuned_pt_stack = bc.sm;
    ctx->tuned_pt_ring = bc.rg;
    ctx->tuned_pt_fetch = bc.fm;
    *out = flg_tuning{};
    out->block_size = bc.b;
    out->blocks_per_cu = bc.bpc;
    out->stack_mode = bc.sm;
    out->ring_depth = bc.rg;
    out->ray_fetch = bc.fm;
    out->ms = best;
    out->ms_default = def;
    out->candidates = n;
    return FLG_OK;
  });
}

}  // extern "C"
