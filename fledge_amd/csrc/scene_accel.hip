// Closest hit over a scene of several meshes: the scene's accel over its
// MeshPrimitives (Scene::init, src/scene.cpp:46: a NaiveBVHAccel over
// m_primitives, src/accel.cpp:35-133), with one LBVH per mesh underneath in
// place of each MeshPrimitive's own accel (src/primitive.cpp:54-90).
//
// The reference's answer.  NaiveBVHAccel::intersect tests its box (the union
// of its primitives' bounds merged into a default box at the origin,
// AABB::intersect_pbrt against the ray's tMax) and, in a leaf, asks its
// primitives in order, keeping a hit only when it is strictly closer than the
// best so far (accel.cpp:98-110); inner nodes ask the left child, then the
// right.  A mesh asked with the best so far reports the hit it reports alone
// whenever that hit is no farther: its traversal prunes only boxes entered
// beyond the best, so it visits the same triangles at its closest distance in
// the same order.  The scene's hit is therefore the closest over the meshes
// and, among meshes at that distance, the one the reference asks first: its
// rank in NaiveBVHAccel's leaf order (the scene's order for up to 64 meshes;
// beyond, the order its median splits leave, accel.cpp:43-73).
//
// The top level here is a binary tree over the mesh boxes, built on the host
// (median split on the longest axis of the box centres, one mesh per leaf) and
// visited nearest child first on the GPU, pruning only boxes entered beyond
// the best distance; the per-ray state keeps (t, rank) lexicographically
// smallest, so the visiting order does not change the answer.  The
// reference's root box test is kept as is.
#include <algorithm>
#include <climits>
#include <numeric>
#include <vector>

#include "traverse.hpp"
#include "volume.hpp"

namespace flg {

// One mesh as the top level sees it (64 bytes).
struct alignas(16) MeshRef {
  const BNode* nodes;
  const BTri* tris;
  int32_t n;
  int32_t rank;  // position in the reference's leaf order
  float root[6];
  int32_t pad[2];
};
static_assert(sizeof(MeshRef) == 64, "mesh ref is 64 bytes");

}  // namespace flg

struct flg_scene_accel {
  flg_ctx* ctx = nullptr;
  int32_t n = 0;
  bool compat = false;
  uint32_t root_ref = 0;   // node 0, or a leaf (kLeafBit | mesh) when n == 1
  float ref_box[6];        // the reference's root box: lo.xyz, hi.xyz
  flg::MeshRef* meshes = nullptr;  // device, n
  flg::BNode* nodes = nullptr;     // device, n - 1
  std::vector<int32_t> rank;       // host copy
};

namespace flg {

namespace {

constexpr int kTopStack = 64;  // > depth of a median-split tree over < 2^31 meshes
// Launch shape of the trace: 256-lane blocks, the 8 most recent entries of
// each lane's mesh traversal stack in LDS (the path tracer's ring stack), 8
// waves per SIMD.
constexpr int kSceneBlock = 256;
constexpr int kSceneRing = 8;
constexpr int kSceneWaves = 8;

struct Box {
  float lo[3], hi[3];
};

Box box_union(const Box& a, const Box& b) {
  Box r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = std::min(a.lo[k], b.lo[k]);
    r.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
  return r;
}

// NaiveBVHAccel's ctor order (accel.cpp:35-73): a node of more than 64
// primitives at depth <= 20 sorts them by the centre of their bounds along
// axis depth % 3 (std::sort, as the reference: its order among equal centres
// is the library's) and gives the first half to its left child.
void reference_order(std::vector<int32_t>& idx, size_t lo, size_t hi, int depth, const std::vector<Box>& b,
                     std::vector<int32_t>& out) {
  const size_t n = hi - lo;
  if (n <= 64 || depth > 20) {
    for (size_t i = lo; i < hi; ++i) out.push_back(idx[i]);
    return;
  }
  const int dim = depth % 3;
  std::sort(idx.begin() + lo, idx.begin() + hi, [&](int32_t x, int32_t y) {
    return (b[x].lo[dim] + b[x].hi[dim]) / 2.0f < (b[y].lo[dim] + b[y].hi[dim]) / 2.0f;
  });
  const size_t mid = lo + n / 2;
  reference_order(idx, lo, mid, depth + 1, b, out);
  reference_order(idx, mid, hi, depth + 1, b, out);
}

// The tree visited here: median split on the longest axis of the box centres.
// Returns the ref of the subtree over idx[lo, hi) and its box.
uint32_t build_top(std::vector<int32_t>& idx, size_t lo, size_t hi, const std::vector<Box>& b,
                   std::vector<BNode>& nodes, Box* box) {
  if (hi - lo == 1) {
    *box = b[idx[lo]];
    return kLeafBit | (uint32_t)idx[lo];
  }
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = lo; i < hi; ++i)
    for (int k = 0; k < 3; ++k) {
      const float c = 0.5f * (b[idx[i]].lo[k] + b[idx[i]].hi[k]);
      clo[k] = std::min(clo[k], c);
      chi[k] = std::max(chi[k], c);
    }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
  std::stable_sort(idx.begin() + lo, idx.begin() + hi, [&](int32_t x, int32_t y) {
    return b[x].lo[axis] + b[x].hi[axis] < b[y].lo[axis] + b[y].hi[axis];
  });
  const size_t mid = lo + (hi - lo) / 2;
  const uint32_t me = (uint32_t)nodes.size();
  nodes.push_back(BNode{});
  Box bl, br;
  const uint32_t l = build_top(idx, lo, mid, b, nodes, &bl);
  const uint32_t r = build_top(idx, mid, hi, b, nodes, &br);
  BNode& nd = nodes[me];
  nd.a = make_float4(bl.lo[0], bl.lo[1], bl.lo[2], bl.hi[0]);
  nd.b = make_float4(bl.hi[1], bl.hi[2], br.lo[0], br.lo[1]);
  nd.c = make_float4(br.lo[2], br.hi[0], br.hi[1], br.hi[2]);
  nd.d = make_int4((int32_t)l, (int32_t)r, 0, 0);
  *box = box_union(bl, br);
  return me;
}

struct SceneDev {
  const MeshRef* meshes;
  const BNode* nodes;
  uint32_t root_ref;
  float ref_box[6];
};

template <bool COMPAT>
__global__ __launch_bounds__(kSceneBlock, kSceneWaves) void k_scene_trace(SceneDev S, const flg_ray* rays,
                                                                          flg_scene_hit* hits, int64_t n) {
  __shared__ uint2 ring[kSceneBlock * kSceneRing];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* r = reinterpret_cast<const float4*>(rays + i);
  const float4 a = r[0], b = r[1];
  const V3 o = v3(a.x, a.y, a.z), d = v3(b.x, b.y, b.z);
  float best = a.w;
  int32_t best_rank = INT_MAX, best_mesh = -1, best_id = -1;
  float t0, t1;
  // the reference's root: its box against [0, tMax]
  if (box_pbrt(v3(S.ref_box[0], S.ref_box[1], S.ref_box[2]), v3(S.ref_box[3], S.ref_box[4], S.ref_box[5]), o, d,
               a.w, &t0, &t1)) {
    const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t stack[kTopStack];
    int sp = 0;
    stack[sp++] = S.root_ref;
    while (sp > 0) {
      const uint32_t ref = stack[--sp];
      if (ref & kLeafBit) {
        const int32_t m = (int32_t)(ref & ~kLeafBit);
        const MeshRef M = S.meshes[m];
        TraceGeom g;
        g.nodes = M.nodes;
        g.tris = M.tris;
        g.n = M.n;
        for (int k = 0; k < 6; ++k) g.root[k] = M.root[k];
        float tb = best;
        Counters c{0, 0};
        const int32_t p = trace_lane<COMPAT, false, kSceneBlock, kSceneRing>(g, o, d, tb, c, ring);
        if (p >= 0 && (tb < best || M.rank < best_rank)) {
          best = tb;
          best_rank = M.rank;
          best_mesh = m;
          best_id = (int32_t)f2u(M.tris[p].v0.w);
        }
        continue;
      }
      const BNode nd = S.nodes[ref];
      float tl, tr;
      const bool hl = slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, &tl) && !(tl > best);
      const bool hr = slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, &tr) && !(tr > best);
      const uint32_t rl = (uint32_t)nd.d.x, rr = (uint32_t)nd.d.y;
      if (hl && hr) {  // nearer child on top
        stack[sp++] = tl <= tr ? rr : rl;
        stack[sp++] = tl <= tr ? rl : rr;
      } else if (hl) {
        stack[sp++] = rl;
      } else if (hr) {
        stack[sp++] = rr;
      }
    }
  }
  flg_scene_hit h;
  h.mesh = best_mesh;
  h.id = best_id;
  h.t = best_mesh >= 0 ? best : a.w;
  h.pad = 0;
  hits[i] = h;
}

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_scene_accel_build(flg_ctx* ctx, flg_bvh* const* h_meshes, int32_t n, flg_scene_accel** out) {
  return host_guard(ctx, "flg_scene_accel_build", [&]() -> int {
    if (!ctx || !h_meshes || n <= 0 || !out) return set_error(ctx, FLG_ERR_INVALID, "flg_scene_accel_build: bad arguments");
    *out = nullptr;
    std::vector<Box> boxes(n);
    const bool compat = (h_meshes[0] && (h_meshes[0]->flags & FLG_NORMAL_COMPAT));
    for (int32_t m = 0; m < n; ++m) {
      const flg_bvh* b = h_meshes[m];
      if (!b || b->n <= 0 || b->ctx != ctx)
        return set_error(ctx, FLG_ERR_INVALID, "flg_scene_accel_build: a mesh is missing, empty or on another context");
      if (((b->flags & FLG_NORMAL_COMPAT) != 0) != compat)
        return set_error(ctx, FLG_ERR_INVALID, "flg_scene_accel_build: meshes differ in FLG_NORMAL_COMPAT");
      for (int k = 0; k < 3; ++k) {
        boxes[m].lo[k] = b->bound[k];
        boxes[m].hi[k] = b->bound[3 + k];
      }
    }
    flg_scene_accel* s = new flg_scene_accel;
    s->ctx = ctx;
    s->n = n;
    s->compat = compat;
    // the reference's leaf order and root box (merged into the default box at
    // the origin, accel.cpp:41-42 with aabb.h's AABB())
    std::vector<int32_t> idx(n), order;
    std::iota(idx.begin(), idx.end(), 0);
    order.reserve(n);
    reference_order(idx, 0, (size_t)n, 0, boxes, order);
    s->rank.assign(n, 0);
    for (int32_t k = 0; k < n; ++k) s->rank[order[k]] = k;
    Box rb{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    for (int32_t m = 0; m < n; ++m) rb = box_union(rb, boxes[m]);
    for (int k = 0; k < 3; ++k) {
      s->ref_box[k] = rb.lo[k];
      s->ref_box[3 + k] = rb.hi[k];
    }
    std::vector<BNode> nodes;
    nodes.reserve(n > 1 ? n - 1 : 0);
    std::iota(idx.begin(), idx.end(), 0);
    Box all;
    s->root_ref = build_top(idx, 0, (size_t)n, boxes, nodes, &all);
    std::vector<MeshRef> refs(n);
    for (int32_t m = 0; m < n; ++m) {
      const flg_bvh* b = h_meshes[m];
      refs[m].nodes = b->nodes;
      refs[m].tris = b->tris;
      refs[m].n = b->n;
      refs[m].rank = s->rank[m];
      for (int k = 0; k < 6; ++k) refs[m].root[k] = b->bound[k];
      refs[m].pad[0] = refs[m].pad[1] = 0;
    }
    (void)hipSetDevice(ctx->device);
    if (hipMalloc((void**)&s->meshes, sizeof(MeshRef) * (size_t)n) != hipSuccess ||
        (!nodes.empty() && hipMalloc((void**)&s->nodes, sizeof(BNode) * nodes.size()) != hipSuccess)) {
      flg_scene_accel_destroy(s);
      return set_error(ctx, FLG_ERR_OOM, "flg_scene_accel_build: device allocation failed");
    }
    if (hipMemcpyAsync(s->meshes, refs.data(), sizeof(MeshRef) * (size_t)n, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        (!nodes.empty() && hipMemcpyAsync(s->nodes, nodes.data(), sizeof(BNode) * nodes.size(),
                                          hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      flg_scene_accel_destroy(s);
      return set_error(ctx, FLG_ERR_HIP, "flg_scene_accel_build: upload failed");
    }
    *out = s;
    return FLG_OK;
  });
}

int flg_scene_accel_destroy(flg_scene_accel* s) {
  return host_guard(nullptr, "flg_scene_accel_destroy", [&]() -> int {
    if (!s) return FLG_ERR_INVALID;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->meshes) (void)hipFree(s->meshes);
    if (s->nodes) (void)hipFree(s->nodes);
    delete s;
    return FLG_OK;
  });
}

int flg_scene_accel_rank(const flg_scene_accel* s, int32_t* h_rank) {
  return host_guard(nullptr, "flg_scene_accel_rank", [&]() -> int {
    if (!s || !h_rank) return FLG_ERR_INVALID;
    std::copy(s->rank.begin(), s->rank.end(), h_rank);
    return FLG_OK;
  });
}

int flg_scene_trace_closest(flg_ctx* ctx, const flg_scene_accel* s, const flg_ray* d_rays, flg_scene_hit* d_hits,
                            int64_t n) {
  return host_guard(ctx, "flg_scene_trace_closest", [&]() -> int {
    if (!ctx || !s || s->ctx != ctx || (n > 0 && (!d_rays || !d_hits)) || n < 0)
      return set_error(ctx, FLG_ERR_INVALID, "flg_scene_trace_closest: bad arguments");
    if (n == 0) return FLG_OK;
    SceneDev S;
    S.meshes = s->meshes;
    S.nodes = s->nodes;
    S.root_ref = s->root_ref;
    for (int k = 0; k < 6; ++k) S.ref_box[k] = s->ref_box[k];
    ScopedTimer tm(ctx, "scene_trace");
    if (s->compat)
      k_scene_trace<true><<<div_up(n, kSceneBlock), kSceneBlock, 0, ctx->stream>>>(S, d_rays, d_hits, n);
    else
      k_scene_trace<false><<<div_up(n, kSceneBlock), kSceneBlock, 0, ctx->stream>>>(S, d_rays, d_hits, n);
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

}  // extern "C"
