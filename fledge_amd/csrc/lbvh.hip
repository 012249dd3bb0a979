// Karras radix-tree LBVH build on gfx950, for one mesh or a batch of meshes
// in the same launches (a mesh is a "segment").  Stages, each one launch for
// all segments:
//   1. per-block triangle-bound partials, then one block per segment reduces
//      them to the global bound (RadixBVHBuilder::prestage,
//      src/experimental/radix_bvh.cpp:49-60);
//   2. 30-bit Morton code of each triangle-AABB centroid (radix_bvh.cpp:62-73,
//      experimental_utils.hpp:33-45);
//   3. stable LSD radix sort of (code, triangle index), four 8-bit digit
//      passes, tile histograms in LDS, one exclusive scan over
//      [segment][digit][tile] (which keeps segments apart), wave64
//      ballot-match ranking for the stable scatter (replaces the unstable
//      tbb::parallel_sort, radix_bvh.cpp:77-81: ties keep index order);
//   4. Karras internal-node emission (radix_bvh.cpp:118-161);
//   5. bottom-up refit, one thread per leaf, agent-scope acq_rel counters
//      (radix_bvh.cpp:163-199), child boxes written into the parent's 64-byte
//      node as traversal reads them;
//   6. the triangle soup in Morton order for traversal.
#include <algorithm>
#include <numeric>

#include "bvh.hpp"

namespace flg {

namespace {

constexpr int kBlock = 256;      // threads per block (4 wave64)
constexpr int kTile = 2048;      // keys per sort tile (8 per thread)
constexpr int kItems = kTile / kBlock;

struct SegDesc {
  const float* verts;
  const int32_t* ind;
  int32_t ntris;
  int32_t tri_base;   // offset of the segment's triangles in the concatenated arrays
  int32_t node_base;  // offset of its internal nodes
  int32_t tile_base;  // offset of its sort tiles
  int32_t blk_base;   // offset of its 256-element blocks
  int32_t nblk;
};

struct BlockMap {
  const int32_t* seg;    // per block: segment
  const int32_t* start;  // per block: first local element
};

struct Bounds6 {
  float v[6];
};

FLG_D V3 load_vert(const float* verts, int32_t i) {
  const float* p = verts + 3 * (int64_t)i;
  return v3(p[0], p[1], p[2]);
}

// Triangle::getBound (src/experimental/base_bvh.hpp:90-102)
FLG_D void tri_bound(const SegDesc& s, int32_t tri, V3* lo, V3* hi) {
  const int32_t* v = s.ind + 3 * (int64_t)tri;
  V3 p0 = load_vert(s.verts, v[0]), p1 = load_vert(s.verts, v[1]), p2 = load_vert(s.verts, v[2]);
  *lo = vmin(vmin(p0, p1), p2);
  *hi = vmax(vmax(p0, p1), p2);
}

FLG_D float wave_min(float x) {
  for (int o = 32; o >= 1; o >>= 1) x = smin(x, __shfl_xor(x, o, 64));
  return x;
}
FLG_D float wave_max(float x) {
  for (int o = 32; o >= 1; o >>= 1) x = smax(x, __shfl_xor(x, o, 64));
  return x;
}

__global__ __launch_bounds__(kBlock) void k_bound_partial(const SegDesc* segs, BlockMap map,
                                                          Bounds6* partial) {
  __shared__ float red[6][kBlock / 64];
  const int b = blockIdx.x;
  const SegDesc s = segs[map.seg[b]];
  const int32_t i = map.start[b] + (int32_t)threadIdx.x;
  V3 lo = v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f);
  V3 hi = v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f);
  if (i < s.ntris) tri_bound(s, i, &lo, &hi);
  float v[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z),
                wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float r = red[k][0];
    for (int q = 1; q < kBlock / 64; ++q) r = k < 3 ? smin(r, red[k][q]) : smax(r, red[k][q]);
    partial[b].v[k] = r;
  }
}

// One block per segment: reduce the segment's block partials.
__global__ __launch_bounds__(kBlock) void k_bound_final(const SegDesc* segs,
                                                        const Bounds6* partial, float* gbound) {
  __shared__ float red[6][kBlock / 64];
  const SegDesc s = segs[blockIdx.x];
  V3 lo = v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f);
  V3 hi = v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f);
  for (int j = threadIdx.x; j < s.nblk; j += kBlock) {
    const Bounds6& p = partial[s.blk_base + j];
    lo = vmin(lo, v3(p.v[0], p.v[1], p.v[2]));
    hi = vmax(hi, v3(p.v[3], p.v[4], p.v[5]));
  }
  float v[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z),
                wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float r = red[k][0];
    for (int q = 1; q < kBlock / 64; ++q) r = k < 3 ? smin(r, red[k][q]) : smax(r, red[k][q]);
    gbound[6 * blockIdx.x + k] = r;
  }
}

// LeftShift3 / EncodeMorton3 (experimental_utils.hpp:33-45)
FLG_D uint32_t left_shift3(uint32_t x) {
  if (x == (1u << 10)) --x;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const SegDesc* segs, BlockMap map,
                                                   const float* gbound, uint32_t* codes,
                                                   uint32_t* keys, uint32_t* vals) {
  const int b = blockIdx.x;
  const int sid = map.seg[b];
  const SegDesc s = segs[sid];
  const int32_t i = map.start[b] + (int32_t)threadIdx.x;
  if (i >= s.ntris) return;
  const float* g = gbound + 6 * sid;
  V3 glo = v3(g[0], g[1], g[2]);
  V3 edges = v3(g[3], g[4], g[5]) - glo;
  // zero extent (planar input) is replaced by 1 (the reference asserts)
  if (edges.x == 0.0f) edges.x = 1.0f;
  if (edges.y == 0.0f) edges.y = 1.0f;
  if (edges.z == 0.0f) edges.z = 1.0f;
  V3 lo, hi;
  tri_bound(s, i, &lo, &hi);
  V3 c = (lo + hi) / 2.0f;
  V3 q = ((c - glo) / edges) * 1024.0f;
  uint32_t code = (left_shift3((uint32_t)q.z) << 2) | (left_shift3((uint32_t)q.y) << 1) |
                  left_shift3((uint32_t)q.x);
  codes[s.tri_base + i] = code;
  keys[s.tri_base + i] = code;
  vals[s.tri_base + i] = (uint32_t)i;
}

// ------------------------------------------------------------------ sort ---
__global__ __launch_bounds__(kBlock) void k_sort_hist(const SegDesc* segs, BlockMap tiles,
                                                      const uint32_t* keys, int shift,
                                                      uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int b = blockIdx.x;
  const SegDesc s = segs[tiles.seg[b]];
  const int32_t start = tiles.start[b];
  for (int r = 0; r < kItems; ++r) {
    int32_t i = start + r * kBlock + (int32_t)threadIdx.x;
    if (i < s.ntris) atomicAdd(&h[(keys[s.tri_base + i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  int ntiles = (s.ntris + kTile - 1) / kTile;
  int tl = start / kTile;
  hist[256 * (int64_t)s.tile_base + (int64_t)threadIdx.x * ntiles + tl] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(const SegDesc* segs, BlockMap tiles,
                                                         const uint32_t* keys_in,
                                                         const uint32_t* vals_in, int shift,
                                                         const uint32_t* offs, uint32_t* keys_out,
                                                         uint32_t* vals_out) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wcnt[kBlock / 64][256];
  const int b = blockIdx.x;
  const SegDesc s = segs[tiles.seg[b]];
  const int32_t start = tiles.start[b];
  const int ntiles = (s.ntris + kTile - 1) / kTile;
  const int tl = start / kTile;
  const int t = threadIdx.x;
  base[t] = offs[256 * (int64_t)s.tile_base + (int64_t)t * ntiles + tl];
  running[t] = 0;
  for (int w = 0; w < kBlock / 64; ++w) wcnt[w][t] = 0;
  __syncthreads();
  const int wave = t / 64, lane = t % 64;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kItems; ++r) {
    int32_t i = start + r * kBlock + t;
    bool valid = i < s.ntris;
    uint32_t key = valid ? keys_in[s.tri_base + i] : 0u;
    uint32_t val = valid ? vals_in[s.tri_base + i] : 0u;
    uint32_t digit = (key >> shift) & 0xFFu;
    uint64_t m = __ballot(valid);
    for (int bit = 0; bit < 8; ++bit) {
      bool set = (digit >> bit) & 1u;
      uint64_t bal = __ballot(valid && set);
      m &= set ? bal : ~bal;
    }
    uint32_t rank = (uint32_t)__popcll(m & lt_mask);
    bool leader = valid && (m >> lane) == 1ull;  // highest lane of its digit group
    if (leader) wcnt[wave][digit] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[digit] + running[digit] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      add += wcnt[w][t];
      wcnt[w][t] = 0;
    }
    running[t] += add;
    __syncthreads();
  }
}

// Exclusive scan of n uint32 values: per-block scans, one block scanning
// the block totals, then the add.
__global__ __launch_bounds__(kBlock) void k_scan_block(const uint32_t* in, uint32_t* out,
                                                       uint32_t* totals, int64_t n) {
  __shared__ uint32_t sums[kBlock];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t acc = 0;
  for (int k = 0; k < kItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    acc += v[k];
  }
  sums[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    uint32_t x = threadIdx.x >= (unsigned)off ? sums[threadIdx.x - off] : 0u;
    __syncthreads();
    sums[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = sums[threadIdx.x] - acc;
  for (int k = 0; k < kItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = sums[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_totals(uint32_t* totals, int nb) {
  __shared__ uint32_t sums[kBlock];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nb; c0 += kBlock) {
    int i = c0 + threadIdx.x;
    uint32_t v = i < nb ? totals[i] : 0u;
    sums[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      uint32_t x = threadIdx.x >= (unsigned)off ? sums[threadIdx.x - off] : 0u;
      __syncthreads();
      sums[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < nb) totals[i] = carry + sums[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += sums[kBlock - 1];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, const uint32_t* totals,
                                                     int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const uint32_t add = totals[blockIdx.x];
  for (int k = 0; k < kItems; ++k) {
    int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
    if (i < n) out[i] += add;
  }
}

// -------------------------------------------------------------- Karras ----
// RadixBVHBuilder::delta (radix_bvh.hpp:68-74)
FLG_D int delta(const uint32_t* code, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = code[i], c = code[j];
  if (a == c) return __clz((int)((uint32_t)i ^ (uint32_t)j)) + 32;
  return __clz((int)(a ^ c));
}

FLG_D int next2pow(int n) {  // Next2Pow experimental_utils.hpp:27-31
  if (n <= 1) return 1;
  return 1 << (32 - __clz(n - 1));
}

struct BuildOut {
  uint32_t* morton_sorted;
  uint32_t* order;
  BNode* nodes;
  BTri* tris;
  int32_t* split;
  int32_t* rlo;
  int32_t* rhi;
  int32_t* parent;
  uint8_t* leafflags;
  int32_t* leaf_parent;
  uint8_t* level;      // per global node: delta of its key range
  int32_t* pslot;      // per global node: 2 * global parent + side, -1 for a root
  int32_t* leaf_slot;  // per global leaf: 2 * global parent + side
};

__global__ __launch_bounds__(kBlock) void k_emit(const SegDesc* segs, BlockMap map, BuildOut o) {
  const int b = blockIdx.x;
  const SegDesc s = segs[map.seg[b]];
  const int n = s.ntris;
  const int i = map.start[b] + (int)threadIdx.x;
  // a single triangle has no internal node: its leaf's parent field is 0, the
  // zero-initialised value the other builds leave in unused fields
  if (n == 1 && i == 0) o.leaf_parent[s.tri_base] = 0;
  if (i >= n - 1) return;
  const uint32_t* code = o.morton_sorted + s.tri_base;
  int dl = delta(code, n, i, i + 1) - delta(code, n, i, i - 1);
  int d = dl == 0 ? 0 : (dl > 0 ? 1 : -1);
  int dmin = delta(code, n, i, i - d);
  int lmax = 2;
  while (delta(code, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t >>= 1)
    if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(code, n, i, j);
  int sp = 0;
  for (int t = next2pow(l); t >= 1; t >>= 1)
    if (delta(code, n, i, i + (sp + t) * d) > dnode) sp += t;
  int gamma = i + sp * d + (d < 0 ? d : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  bool lleaf = lo == gamma, rleaf = hi == gamma + 1;
  const int nb = s.node_base, tb = s.tri_base;
  o.split[nb + i] = gamma;
  o.rlo[nb + i] = lo;
  o.rhi[nb + i] = hi;
  o.leafflags[nb + i] = (uint8_t)((lleaf ? 1 : 0) | (rleaf ? 2 : 0));
  o.level[nb + i] = (uint8_t)dnode;  // common prefix length of the node's keys
  if (i == 0) {
    o.parent[nb] = 0;
    o.pslot[nb] = -1;
  }
  const int slot0 = 2 * (nb + i);
  if (lleaf) {
    o.leaf_parent[tb + gamma] = i;
    o.leaf_slot[tb + gamma] = slot0;
  } else {
    o.parent[nb + gamma] = i;
    o.pslot[nb + gamma] = slot0;
  }
  if (rleaf) {
    o.leaf_parent[tb + gamma + 1] = i;
    o.leaf_slot[tb + gamma + 1] = slot0 + 1;
  } else {
    o.parent[nb + gamma + 1] = i;
    o.pslot[nb + gamma + 1] = slot0 + 1;
  }
  int4 refs;
  refs.x = lleaf ? (int)(kLeafBit | (uint32_t)gamma) : gamma;
  refs.y = rleaf ? (int)(kLeafBit | (uint32_t)(gamma + 1)) : gamma + 1;
  refs.z = 0;
  refs.w = 0;
  o.nodes[nb + i].d = refs;
}

FLG_D void store_slot(BNode* node, int side, V3 lo, V3 hi) {
  float* f = reinterpret_cast<float*>(node) + (side ? 6 : 0);
  f[0] = lo.x; f[1] = lo.y; f[2] = lo.z;
  f[3] = hi.x; f[4] = hi.y; f[5] = hi.z;
}

// Bottom-up refit (the reference climbs leaf-to-root with fetch-and-add flags,
// radix_bvh.cpp:163-199; min/max are exact, so any children-before-parent
// order gives the same bounds).  Here the order comes from the node's common
// prefix length delta, which is strictly larger in a child than in its
// parent: leaves first, then the nodes of each delta level from 63 down to
// 0, one launch per level, no atomics on the tree.
__global__ __launch_bounds__(kBlock) void k_refit_leaves(const SegDesc* segs, BlockMap map,
                                                         BuildOut o) {
  const int b = blockIdx.x;
  const SegDesc s = segs[map.seg[b]];
  const int p = map.start[b] + (int)threadIdx.x;
  if (p >= s.ntris || s.ntris < 2) return;
  V3 lo, hi;
  tri_bound(s, (int32_t)o.order[s.tri_base + p], &lo, &hi);
  const int slot = o.leaf_slot[s.tri_base + p];
  store_slot(o.nodes + (slot >> 1), slot & 1, lo, hi);
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Counting sort of the internal nodes by delta level (64 buckets).
__global__ __launch_bounds__(kBlock) void k_level_hist(const uint8_t* level, int64_t nnodes,
                                                       uint32_t* hist) {
  __shared__ uint32_t h[64];
  if (threadIdx.x < 64) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nnodes; i += (int64_t)gridDim.x * kBlock)
    atomicAdd(&h[level[i] & 63], 1u);
  __syncthreads();
  if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_level_scan(const uint32_t* hist, uint32_t* offs, uint32_t* fill) {
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int d = 0; d < 64; ++d) {
      offs[d] = run;
      fill[d] = run;
      run += hist[d];
    }
    offs[64] = run;
  }
}

__global__ __launch_bounds__(kBlock) void k_level_scatter(const uint8_t* level, int64_t nnodes,
                                                          uint32_t* fill, uint32_t* list) {
  __shared__ uint32_t h[64], base[64];
  for (int64_t c0 = (int64_t)blockIdx.x * kBlock; c0 < nnodes; c0 += (int64_t)gridDim.x * kBlock) {
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    int64_t i = c0 + threadIdx.x;
    uint32_t d = 0, r = 0;
    if (i < nnodes) {
      d = level[i] & 63;
      r = atomicAdd(&h[d], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&fill[threadIdx.x], h[threadIdx.x]);
    __syncthreads();
    if (i < nnodes) list[base[d] + r] = (uint32_t)i;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_refit_level(BuildOut o, const uint32_t* offs,
                                                        const uint32_t* list, int d) {
  const uint32_t b0 = offs[d], b1 = offs[d + 1];
  for (uint32_t k = b0 + blockIdx.x * kBlock + threadIdx.x; k < b1; k += gridDim.x * kBlock) {
    const uint32_t x = list[k];
    const int slot = o.pslot[x];
    if (slot < 0) continue;  // a segment root
    const float* f = reinterpret_cast<const float*>(o.nodes + x);
    V3 alo = v3(f[0], f[1], f[2]), ahi = v3(f[3], f[4], f[5]);
    V3 clo = v3(f[6], f[7], f[8]), chi = v3(f[9], f[10], f[11]);
    // BVHBound{} merged with the other child, then with this one (radix_bvh.cpp:190-192)
    V3 lo = vmin(vmin(v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f), clo), alo);
    V3 hi = vmax(vmax(v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f), chi), ahi);
    store_slot(o.nodes + (slot >> 1), slot & 1, lo, hi);
  }
}

__global__ __launch_bounds__(kBlock) void k_soup(const SegDesc* segs, BlockMap map, BuildOut o) {
  const int b = blockIdx.x;
  const SegDesc s = segs[map.seg[b]];
  const int p = map.start[b] + (int)threadIdx.x;
  if (p >= s.ntris) return;
  const uint32_t tri = o.order[s.tri_base + p];
  const int32_t* v = s.ind + 3 * (int64_t)tri;
  V3 a = load_vert(s.verts, v[0]), bb = load_vert(s.verts, v[1]), c = load_vert(s.verts, v[2]);
  BTri t;
  t.v0 = make_float4(a.x, a.y, a.z, u2f(tri));
  t.v1 = make_float4(bb.x, bb.y, bb.z, 0.0f);
  t.v2 = make_float4(c.x, c.y, c.z, 0.0f);
  o.tris[s.tri_base + p] = t;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Builds every segment; out[i] receives the handle of mesh i.
static int build_segments(flg_ctx* ctx, const flg_mesh_desc* meshes, int nseg, uint32_t flags,
                          flg_bvh** out) {
  if (!ctx || !meshes || nseg <= 0 || !out) return set_error(ctx, FLG_ERR_INVALID, "lbvh build: bad arguments");
  if (flags > FLG_NORMAL_COMPAT) return set_error(ctx, FLG_ERR_INVALID, "lbvh build: unknown flags");
  FLG_CHECK_HIP(ctx, hipSetDevice(ctx->device));
  // ---- host-side segment table and block maps
  std::vector<SegDesc> segs(nseg);
  std::vector<int32_t> bseg, bstart, tseg, tstart;
  int64_t tri_total = 0, node_total = 0, tile_total = 0, blk_total = 0;
  for (int s = 0; s < nseg; ++s) {
    const flg_mesh_desc& m = meshes[s];
    if (m.ntris < 0 || m.nverts < 0 || m.ntris >= (1 << 30))
      return set_error(ctx, FLG_ERR_RANGE, "lbvh build: ntris out of range [0, 2^30)");
    if (m.ntris > 0 && (!m.d_verts || !m.d_ind))
      return set_error(ctx, FLG_ERR_INVALID, "lbvh build: null mesh pointer");
    SegDesc& d = segs[s];
    d.verts = m.d_verts;
    d.ind = m.d_ind;
    d.ntris = m.ntris;
    d.tri_base = (int32_t)tri_total;
    d.node_base = (int32_t)node_total;
    d.tile_base = (int32_t)tile_total;
    d.blk_base = (int32_t)blk_total;
    int nblk = (m.ntris + kBlock - 1) / kBlock;
    int ntile = (m.ntris + kTile - 1) / kTile;
    d.nblk = nblk;
    for (int k = 0; k < nblk; ++k) { bseg.push_back(s); bstart.push_back(k * kBlock); }
    for (int k = 0; k < ntile; ++k) { tseg.push_back(s); tstart.push_back(k * kTile); }
    tri_total += m.ntris;
    node_total += m.ntris > 0 ? m.ntris - 1 : 0;
    tile_total += ntile;
    blk_total += nblk;
  }
  if (tri_total >= (int64_t)1 << 31) return set_error(ctx, FLG_ERR_RANGE, "lbvh build: batch too large");

  // ---- persistent output block (shared by the batch's handles)
  const size_t N = (size_t)tri_total, NI = (size_t)node_total;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
  size_t o_nodes = take(NI * sizeof(BNode)), o_tris = take(N * sizeof(BTri));
  size_t o_morton = take(N * 4), o_msorted = take(N * 4), o_order = take(N * 4);
  size_t o_split = take(NI * 4), o_rlo = take(NI * 4), o_rhi = take(NI * 4), o_parent = take(NI * 4);
  size_t o_lf = take(NI), o_lp = take(N * 4), o_gb = take((size_t)nseg * 6 * 4);
  size_t total_bytes = off;
  void* blockp = nullptr;
  FLG_CHECK_HIP(ctx, hipMalloc(&blockp, total_bytes ? total_bytes : 256));
  char* base = (char*)blockp;

  // ---- scratch
  int st = FLG_OK;
  size_t s_off = 0;
  auto stake = [&](size_t bytes) { size_t o = s_off; s_off = align256(s_off + bytes); return o; };
  size_t so_segs = stake(nseg * sizeof(SegDesc));
  size_t so_bseg = stake(blk_total * 4), so_bstart = stake(blk_total * 4);
  size_t so_tseg = stake(tile_total * 4), so_tstart = stake(tile_total * 4);
  size_t so_part = stake(blk_total * sizeof(Bounds6));
  size_t so_k0 = stake(N * 4), so_v0 = stake(N * 4), so_k1 = stake(N * 4), so_v1 = stake(N * 4);
  int64_t hist_n = 256 * tile_total;
  int64_t scan_blocks = (hist_n + kTile - 1) / kTile;
  size_t so_hist = stake(hist_n * 4), so_offs = stake(hist_n * 4), so_tot = stake(scan_blocks * 4 + 4);
  size_t so_level = stake(NI), so_pslot = stake(NI * 4), so_lslot = stake(N * 4);
  size_t so_lhist = stake(64 * 4), so_loffs = stake(65 * 4), so_lfill = stake(64 * 4);
  size_t so_llist = stake(NI * 4);
  char* sc = (char*)scratch(ctx, "lbvh_build", s_off ? s_off : 256, &st);
  if (!sc) {
    (void)hipFree(blockp);
    return st;
  }
  auto fail = [&](int code) { (void)hipFree(blockp); return code; };
  SegDesc* d_segs = (SegDesc*)(sc + so_segs);
  int32_t* d_bseg = (int32_t*)(sc + so_bseg);
  int32_t* d_bstart = (int32_t*)(sc + so_bstart);
  int32_t* d_tseg = (int32_t*)(sc + so_tseg);
  int32_t* d_tstart = (int32_t*)(sc + so_tstart);
  if (hipMemcpyAsync(d_segs, segs.data(), nseg * sizeof(SegDesc), hipMemcpyHostToDevice, ctx->stream) ||
      (blk_total && hipMemcpyAsync(d_bseg, bseg.data(), blk_total * 4, hipMemcpyHostToDevice, ctx->stream)) ||
      (blk_total && hipMemcpyAsync(d_bstart, bstart.data(), blk_total * 4, hipMemcpyHostToDevice, ctx->stream)) ||
      (tile_total && hipMemcpyAsync(d_tseg, tseg.data(), tile_total * 4, hipMemcpyHostToDevice, ctx->stream)) ||
      (tile_total && hipMemcpyAsync(d_tstart, tstart.data(), tile_total * 4, hipMemcpyHostToDevice, ctx->stream)))
    return fail(set_error(ctx, FLG_ERR_HIP, "lbvh build: upload of the segment table failed"));
  BlockMap bmap{d_bseg, d_bstart}, tmap{d_tseg, d_tstart};
  Bounds6* d_part = (Bounds6*)(sc + so_part);
  uint32_t* k0 = (uint32_t*)(sc + so_k0);
  uint32_t* v0 = (uint32_t*)(sc + so_v0);
  uint32_t* k1 = (uint32_t*)(sc + so_k1);
  uint32_t* v1 = (uint32_t*)(sc + so_v1);
  uint32_t* hist = (uint32_t*)(sc + so_hist);
  uint32_t* offs = (uint32_t*)(sc + so_offs);
  uint32_t* tot = (uint32_t*)(sc + so_tot);

  BuildOut o;
  o.nodes = (BNode*)(base + o_nodes);
  o.tris = (BTri*)(base + o_tris);
  uint32_t* morton = (uint32_t*)(base + o_morton);
  o.morton_sorted = (uint32_t*)(base + o_msorted);
  o.order = (uint32_t*)(base + o_order);
  o.split = (int32_t*)(base + o_split);
  o.rlo = (int32_t*)(base + o_rlo);
  o.rhi = (int32_t*)(base + o_rhi);
  o.parent = (int32_t*)(base + o_parent);
  o.leafflags = (uint8_t*)(base + o_lf);
  o.leaf_parent = (int32_t*)(base + o_lp);
  o.level = (uint8_t*)(sc + so_level);
  o.pslot = (int32_t*)(sc + so_pslot);
  o.leaf_slot = (int32_t*)(sc + so_lslot);
  uint32_t* lhist = (uint32_t*)(sc + so_lhist);
  uint32_t* loffs = (uint32_t*)(sc + so_loffs);
  uint32_t* lfill = (uint32_t*)(sc + so_lfill);
  uint32_t* llist = (uint32_t*)(sc + so_llist);
  float* gb = (float*)(base + o_gb);

  if (blk_total > 0) {
    {
      ScopedTimer tm(ctx, "lbvh_bound");
      k_bound_partial<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, bmap, d_part);
      k_bound_final<<<nseg, kBlock, 0, ctx->stream>>>(d_segs, d_part, gb);
    }
    {
      ScopedTimer tm(ctx, "lbvh_morton");
      k_morton<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, bmap, gb, morton, k0, v0);
    }
    {
      ScopedTimer tm(ctx, "lbvh_sort");
      uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
      for (int pass = 0; pass < 4; ++pass) {
        int shift = 8 * pass;
        k_sort_hist<<<(int)tile_total, kBlock, 0, ctx->stream>>>(d_segs, tmap, kin, shift, hist);
        k_scan_block<<<(int)scan_blocks, kBlock, 0, ctx->stream>>>(hist, offs, tot, hist_n);
        k_scan_totals<<<1, kBlock, 0, ctx->stream>>>(tot, (int)scan_blocks);
        k_scan_add<<<(int)scan_blocks, kBlock, 0, ctx->stream>>>(offs, tot, hist_n);
        if (pass == 3) { kout = o.morton_sorted; vout = o.order; }
        k_sort_scatter<<<(int)tile_total, kBlock, 0, ctx->stream>>>(d_segs, tmap, kin, vin, shift,
                                                                     offs, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
      }
    }
    {
      ScopedTimer tm(ctx, "lbvh_emit");
      k_emit<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, bmap, o);
    }
    {
      ScopedTimer tm(ctx, "lbvh_refit");
      k_refit_leaves<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, bmap, o);
      if (NI > 0) {
        const int g = std::max(1, std::min(div_up((int64_t)NI, kBlock), ctx->num_cus * 4));
        k_zero_u32<<<1, 64, 0, ctx->stream>>>(lhist, 64);
        k_level_hist<<<g, kBlock, 0, ctx->stream>>>(o.level, (int64_t)NI, lhist);
        k_level_scan<<<1, 64, 0, ctx->stream>>>(lhist, loffs, lfill);
        k_level_scatter<<<g, kBlock, 0, ctx->stream>>>(o.level, (int64_t)NI, lfill, llist);
        for (int d = 63; d >= 0; --d)
          k_refit_level<<<g, kBlock, 0, ctx->stream>>>(o, loffs, llist, d);
      }
    }
    {
      ScopedTimer tm(ctx, "lbvh_soup");
      k_soup<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, bmap, o);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(check_hip(ctx, e, "lbvh build kernels"));
  }
  std::vector<float> hgb((size_t)nseg * 6, 0.0f);
  if (hipMemcpyAsync(hgb.data(), gb, hgb.size() * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(set_error(ctx, FLG_ERR_HIP, "lbvh build: synchronisation failed"));

  SharedBlock* sb = new SharedBlock;
  sb->ptr = blockp;
  sb->refs = nseg;
  for (int s = 0; s < nseg; ++s) {
    const SegDesc& d = segs[s];
    flg_bvh* h = new flg_bvh;
    h->ctx = ctx;
    h->n = d.ntris;
    h->flags = flags;
    for (int k = 0; k < 6; ++k) h->bound[k] = d.ntris > 0 ? hgb[6 * s + k] : 0.0f;
    h->nodes = o.nodes + d.node_base;
    h->tris = o.tris + d.tri_base;
    h->morton = morton + d.tri_base;
    h->morton_sorted = o.morton_sorted + d.tri_base;
    h->order = o.order + d.tri_base;
    h->split = o.split + d.node_base;
    h->rlo = o.rlo + d.node_base;
    h->rhi = o.rhi + d.node_base;
    h->parent = o.parent + d.node_base;
    h->leafflags = o.leafflags + d.node_base;
    h->leaf_parent = o.leaf_parent + d.tri_base;
    h->gbound = gb + 6 * s;
    h->bytes = (uint64_t)d.ntris * (sizeof(BTri) + 5 * 4) +
               (uint64_t)(d.ntris > 0 ? d.ntris - 1 : 0) * (sizeof(BNode) + 17);
    h->block = sb;
    out[s] = h;
  }
  return FLG_OK;
}

}  // namespace flg

using namespace flg;

extern "C" {

int flg_lbvh_build(flg_ctx* ctx, const float* d_verts, const int32_t* d_ind, int32_t ntris,
                   int32_t nverts, uint32_t flags, flg_bvh** out) {
  return host_guard(ctx, "flg_lbvh_build", [&]() -> int {
    if (!out) return set_error(ctx, FLG_ERR_INVALID, "flg_lbvh_build: null out");
    *out = nullptr;
    flg_mesh_desc m{d_verts, d_ind, ntris, nverts};
    return build_segments(ctx, &m, 1, flags, out);
  });
}

int flg_lbvh_build_batched(flg_ctx* ctx, const flg_mesh_desc* h_meshes, int32_t n, uint32_t flags,
                           flg_bvh** out) {
  return host_guard(ctx, "flg_lbvh_build_batched", [&]() -> int {
    if (!out || n <= 0) return set_error(ctx, FLG_ERR_INVALID, "flg_lbvh_build_batched: bad arguments");
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    return build_segments(ctx, h_meshes, n, flags, out);
  });
}

int flg_lbvh_destroy(flg_bvh* bvh) {
  return host_guard(nullptr, "flg_lbvh_destroy", [&]() -> int {
    if (!bvh) return FLG_ERR_INVALID;
    SharedBlock* sb = bvh->block;
    if (sb && --sb->refs == 0) {
      (void)hipSetDevice(bvh->ctx->device);
      (void)hipStreamSynchronize(bvh->ctx->stream);
      (void)hipFree(sb->ptr);
      delete sb;
    }
    delete bvh;
    return FLG_OK;
  });
}

int flg_lbvh_get_info(flg_bvh* bvh, flg_lbvh_info* out) {
  return host_guard(nullptr, "flg_lbvh_get_info", [&]() -> int {
    if (!bvh || !out) return FLG_ERR_INVALID;
    out->ntris = bvh->n;
    for (int k = 0; k < 6; ++k) out->bound[k] = bvh->bound[k];
    out->device_bytes = bvh->bytes;
    return FLG_OK;
  });
}

int flg_lbvh_export(flg_ctx* ctx, flg_bvh* bvh, const flg_lbvh_view* v) {
  return host_guard(ctx, "flg_lbvh_export", [&]() -> int {
    if (!ctx || !bvh || !v) return set_error(ctx, FLG_ERR_INVALID, "flg_lbvh_export: bad arguments");
    const size_t n = (size_t)bvh->n, ni = n > 0 ? n - 1 : 0;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
      if (!dst || !bytes) return FLG_OK;
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
      return FLG_OK;
    };
    int st;
    if ((st = cp(v->h_morton, bvh->morton, n * 4))) return st;
    if ((st = cp(v->h_morton_sorted, bvh->morton_sorted, n * 4))) return st;
    if ((st = cp(v->h_order, bvh->order, n * 4))) return st;
    if ((st = cp(v->h_split, bvh->split, ni * 4))) return st;
    if ((st = cp(v->h_range_lo, bvh->rlo, ni * 4))) return st;
    if ((st = cp(v->h_range_hi, bvh->rhi, ni * 4))) return st;
    if ((st = cp(v->h_parent, bvh->parent, ni * 4))) return st;
    if ((st = cp(v->h_leaf_parent, bvh->leaf_parent, n * 4))) return st;
    std::vector<uint8_t> lf(ni);
    std::vector<BNode> nodes(ni);
    if (ni) {
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(lf.data(), bvh->leafflags, ni, hipMemcpyDeviceToHost, ctx->stream));
      if (v->h_node_bound)
        FLG_CHECK_HIP(ctx, hipMemcpyAsync(nodes.data(), bvh->nodes, ni * sizeof(BNode),
                                          hipMemcpyDeviceToHost, ctx->stream));
    }
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < ni; ++i) {
      if (v->h_left_leaf) v->h_left_leaf[i] = lf[i] & 1;
      if (v->h_right_leaf) v->h_right_leaf[i] = (lf[i] >> 1) & 1;
      if (v->h_node_bound) {
        const float* f = reinterpret_cast<const float*>(&nodes[i]);
        // node bound = union of its two child boxes (merge order of the refit)
        float* q = v->h_node_bound + 6 * i;
        for (int k = 0; k < 3; ++k) {
          float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
          lo = smin(smin(lo, f[6 + k]), f[k]);
          hi = smax(smax(hi, f[9 + k]), f[3 + k]);
          q[k] = lo;
          q[3 + k] = hi;
        }
      }
    }
    return FLG_OK;
  });
}

int flg_lbvh_download(flg_ctx* ctx, const flg_bvh* bvh, void* h_nodes, size_t nodes_bytes,
                      void* h_tris, size_t tris_bytes) {
  return host_guard(ctx, "flg_lbvh_download", [&]() -> int {
    if (!ctx || !bvh) return set_error(ctx, FLG_ERR_INVALID, "flg_lbvh_download: bad arguments");
    const size_t n = (size_t)bvh->n, ni = n > 0 ? n - 1 : 0;
    if ((h_nodes && nodes_bytes != ni * sizeof(BNode)) || (h_tris && tris_bytes != n * sizeof(BTri)))
      return set_error(ctx, FLG_ERR_INVALID,
                       "flg_lbvh_download: nodes need " + std::to_string(ni * sizeof(BNode)) +
                           " bytes and tris " + std::to_string(n * sizeof(BTri)) + " bytes");
    if (h_nodes && ni)
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(h_nodes, bvh->nodes, nodes_bytes, hipMemcpyDeviceToHost,
                                        ctx->stream));
    if (h_tris && n)
      FLG_CHECK_HIP(ctx, hipMemcpyAsync(h_tris, bvh->tris, tris_bytes, hipMemcpyDeviceToHost,
                                        ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FLG_OK;
  });
}

}  // extern "C"
