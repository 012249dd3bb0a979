// Karras radix-tree LBVH build on gfx950, for one mesh or a batch of meshes
// in the same launches (a mesh is a "segment").  Stages, each one launch for
// all segments:
//   1. per-block triangle-bound partials, then one block per segment reduces
//      them to the global bound (RadixBVHBuilder::prestage,
//      src/experimental/radix_bvh.cpp:49-60);
//   2. 30-bit Morton code of each triangle-AABB centroid (radix_bvh.cpp:62-73,
//      experimental_utils.hpp:33-45);
//   3. stable LSD radix sort of (code, triangle index) in three passes of
//      10-bit digits (replaces the unstable tbb::parallel_sort,
//      radix_bvh.cpp:77-81: ties keep index order).  One launch counts every
//      pass's digits per segment, one scans them; then each pass is a single
//      onesweep launch: a 4096-key tile ranks its keys in LDS (wave64
//      ballot-match, stable), finds its global offset per digit by decoupled
//      look-back over the segment's earlier tiles, stages the keys in digit
//      order in LDS and writes them out in runs;
//   4. Karras internal-node emission (radix_bvh.cpp:118-161);
//   5. bottom-up refit (radix_bvh.cpp:163-199), child boxes written into the
//      parent's 64-byte node as traversal reads them, in two launches.  One
//      block per 1024 consecutive leaves gathers their triangles (writing the
//      soup in Morton order), then climbs from its leaves through every node
//      whose key range lies inside its chunk with the child boxes and arrival
//      flags in LDS (the second arrival merges), and also leaves prefix and
//      suffix unions of its leaf boxes where the next launch needs them, and
//      the chunk's union.  The other launch gives each node whose range
//      crosses a chunk boundary the boxes of its children that cross one
//      too, each as (suffix union in its first chunk) + (unions of the whole
//      chunks between) + (prefix union in its last chunk): no dependence
//      between nodes, no counter shared across workgroups, no agent-scope
//      fence.  Min and max are exact, so every grouping gives the reference's
//      bounds bit for bit.
// The output block of a build comes from a per-context pool: a destroyed
// BVH's block is kept and reused by a later build that fits it, so a rebuild
// of the same size allocates nothing.
#include <algorithm>
#include <numeric>

#include "bvh.hpp"

namespace flg {

namespace {

constexpr int kBlock = 256;                       // threads of the per-triangle kernels
constexpr int kDigitBits = 10;                    // three passes over 30-bit codes
constexpr int kDigits = 1 << kDigitBits;
constexpr int kPasses = 3;
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kRounds = 16;                       // rounds of 64 keys per wave
constexpr int kWaveKeys = 64 * kRounds;           // 1024 consecutive keys per wave
constexpr int kTile = kSortWaves * kWaveKeys;     // 4096 keys per onesweep tile
constexpr int kHistBlock = 16384;                 // keys per digit-histogram block
constexpr int kChunk = 1024;                      // leaves per refit block
constexpr int kChunkThreads = 256;                // threads of a refit block
constexpr int kChunkLeaves = kChunk / kChunkThreads;  // consecutive leaves per thread
constexpr uint32_t kCross = 0x80000000u;          // node info: the node's range crosses a chunk boundary
constexpr uint32_t kNoParent = 0x7FFFFFFFu;       // node info: a segment's root
// look-back status word: flag in the top two bits, count below
constexpr uint32_t kFlagAgg = 1u << 30, kFlagIncl = 2u << 30, kValMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;         // bounded wait, then an error flag

struct SegDesc {
  const float* verts;
  const int32_t* ind;
  int32_t ntris;
  int32_t tri_base;    // first triangle (sorted position) of the segment in the batch's arrays
  int32_t node_base;   // first internal node
  int32_t blk_base;    // first 256-triangle block
  int32_t tile_base;   // first sort tile
  int32_t hblk_base;   // first histogram block
  int32_t chunk_base;  // first refit chunk
  int32_t pad;
};

// The segment of block b of one of the block kinds (F: its first-block
// field): the last segment whose first block is <= b.  Segments without
// blocks of that kind share the next segment's first block and lose the
// search to it.
template <int32_t SegDesc::*F>
FLG_D int seg_of(const SegDesc* segs, int nseg, int b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].*F <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The segment of every 256-triangle block and of every refit chunk, looked
// up once per build so the many-block kernels of a batch read one word
// instead of searching the segment table.
__global__ __launch_bounds__(kBlock) void k_seg_maps(const SegDesc* segs, int nseg, int64_t nblk, int64_t nchunk,
                                                     int32_t* blk_seg, int32_t* chunk_seg);

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

__global__ __launch_bounds__(kBlock) void k_seg_maps(const SegDesc* segs, int nseg, int64_t nblk, int64_t nchunk,
                                                     int32_t* blk_seg, int32_t* chunk_seg) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < nblk) blk_seg[k] = seg_of<&SegDesc::blk_base>(segs, nseg, (int)k);
  if (k < nchunk) chunk_seg[k] = seg_of<&SegDesc::chunk_base>(segs, nseg, (int)k);
}

__global__ __launch_bounds__(kBlock) void k_bound_partial(const SegDesc* segs, const int32_t* blk_seg,
                                                          Bounds6* partial) {
  __shared__ float red[6][kBlock / 64];
  const int b = blockIdx.x;
  const SegDesc s = segs[blk_seg[b]];
  const int32_t i = (b - s.blk_base) * kBlock + (int32_t)threadIdx.x;
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
__global__ __launch_bounds__(kBlock) void k_bound_final(const SegDesc* segs, const Bounds6* partial,
                                                        float* gbound) {
  __shared__ float red[6][kBlock / 64];
  const SegDesc s = segs[blockIdx.x];
  const int nblk = (s.ntris + kBlock - 1) / kBlock;
  V3 lo = v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f);
  V3 hi = v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f);
  for (int j = threadIdx.x; j < nblk; j += kBlock) {
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

__global__ __launch_bounds__(kBlock) void k_morton(const SegDesc* segs, const int32_t* blk_seg, const float* gbound,
                                                   uint32_t* codes, uint32_t* keys, uint32_t* vals) {
  const int b = blockIdx.x;
  const int sid = blk_seg[b];
  const SegDesc s = segs[sid];
  const int32_t i = (b - s.blk_base) * kBlock + (int32_t)threadIdx.x;
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
__global__ void k_fill_u32(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Digit counts of every pass, per segment: hist[(seg * 3 + pass) * 1024 + d].
__global__ __launch_bounds__(kBlock) void k_digit_hist(const SegDesc* segs, int nseg, const uint32_t* keys,
                                                       uint32_t* hist) {
  __shared__ uint32_t h[kPasses * kDigits];
  for (int j = threadIdx.x; j < kPasses * kDigits; j += kBlock) h[j] = 0u;
  __syncthreads();
  const int b = blockIdx.x;
  const int sid = seg_of<&SegDesc::hblk_base>(segs, nseg, b);
  const SegDesc s = segs[sid];
  const int32_t start = (b - s.hblk_base) * kHistBlock;
  const int32_t end = min(start + kHistBlock, s.ntris);
  for (int32_t i = start + (int32_t)threadIdx.x; i < end; i += kBlock) {
    const uint32_t key = keys[s.tri_base + i];
#pragma unroll
    for (int p = 0; p < kPasses; ++p) atomicAdd(&h[p * kDigits + ((key >> (kDigitBits * p)) & (kDigits - 1))], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kPasses * kDigits; j += kBlock)
    if (h[j]) atomicAdd(&hist[(int64_t)sid * kPasses * kDigits + j], h[j]);
}

// Exclusive scan of one (segment, pass) histogram over its 1024 digits, in
// place: the digit's first position in the segment's sorted order.
__global__ __launch_bounds__(kDigits) void k_digit_scan(uint32_t* hist) {
  __shared__ uint32_t wsum[kDigits / 64];
  uint32_t* h = hist + (int64_t)blockIdx.x * kDigits;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t v = h[t];
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int q = 0; q < w; ++q) pre += wsum[q];
  h[t] = pre + x - v;
}

// One onesweep pass.  Tiles take ids in launch order from *counter, so every
// tile a tile waits for has started; a tile publishes its own counts before
// it waits, so the look-back chain always completes.
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const SegDesc* segs, int nseg, int pass,
                                                           const uint32_t* keys_in, const uint32_t* vals_in,
                                                           uint32_t* keys_out, uint32_t* vals_out,
                                                           const uint32_t* dbase, uint32_t* status,
                                                           uint32_t* counter, uint32_t* err) {
  __shared__ uint32_t hw[kSortWaves][kDigits];  // per wave: digit counts, then their prefix over waves
  __shared__ uint32_t toff[kDigits];            // first position of each digit in the tile
  __shared__ uint32_t gpos[kDigits];            // first output position of each digit's keys of the tile
  __shared__ uint32_t skey[kTile], sval[kTile];  // the tile in digit order
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ int tile_sh;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  if (t == 0) tile_sh = (int)atomicAdd(counter, 1u);
  for (int j = t; j < kSortWaves * kDigits; j += kSortThreads) (&hw[0][0])[j] = 0u;
  __syncthreads();
  const int tile = tile_sh;
  const int sid = seg_of<&SegDesc::tile_base>(segs, nseg, tile);
  const SegDesc s = segs[sid];
  const int lt = tile - s.tile_base;
  const int32_t start = lt * kTile;
  const int32_t count = min(kTile, s.ntris - start);
  const int shift = kDigitBits * pass;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t* kin = keys_in + s.tri_base + start;
  const uint32_t* vin = vals_in + s.tri_base + start;

  // rank: wave w takes keys [w * 1024, (w + 1) * 1024) of the tile in rounds
  // of 64; a key's rank among the wave's keys of its digit = the wave's count
  // so far + the lanes before it with the same digit (ballot match)
  uint32_t mkey[kRounds], mval[kRounds], mrank[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int i = wave * kWaveKeys + r * 64 + lane;
    const bool valid = i < count;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & (kDigits - 1);
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kDigitBits; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t bal = __ballot(valid && set);
      m &= set ? bal : ~bal;
    }
    const uint32_t before = hw[wave][d];
    const bool leader = valid && (m >> lane) == 1ull;  // the group's highest lane
    if (leader) hw[wave][d] = before + (uint32_t)__popcll(m);
    mkey[r] = key;
    mval[r] = val;
    mrank[r] = before + (uint32_t)__popcll(m & lt_mask);
  }
  __syncthreads();

  // thread t owns digits 4t .. 4t+3: the tile's count, and the waves' counts
  // turned into their exclusive prefix over the waves
  uint32_t cnt[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int d = 4 * t + k;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = hw[w][d];
      hw[w][d] = run;
      run += c;
    }
    cnt[k] = run;
  }
  // publish the tile's counts (already inclusive for a segment's first tile)
  uint32_t* st = status + (int64_t)tile * kDigits;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    __hip_atomic_store(st + 4 * t + k, (lt == 0 ? kFlagIncl : kFlagAgg) | cnt[k], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // the tile's exclusive digit offsets (block scan over 4-digit groups)
  const uint32_t sum4 = cnt[0] + cnt[1] + cnt[2] + cnt[3];
  uint32_t x = sum4;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t ex = x - sum4;
  for (int w = 0; w < wave; ++w) ex += wsum[w];
  // decoupled look-back: the segment's earlier tiles' counts of each digit
  const uint32_t* db = dbase + ((int64_t)sid * kPasses + pass) * kDigits;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int d = 4 * t + k;
    toff[d] = ex;
    ex += cnt[k];
    uint32_t excl = 0;
    if (lt > 0) {
      int j = tile - 1;
      uint32_t spins = 0;
      while (true) {
        const uint32_t w = __hip_atomic_load(status + (int64_t)j * kDigits + d, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t f = w & ~kValMask;
        if (f == kFlagIncl) {
          excl += w & kValMask;
          break;
        }
        if (f == kFlagAgg) {
          excl += w & kValMask;
          --j;  // never below the segment's first tile, which is inclusive
          spins = 0;
          continue;
        }
        if (++spins > kSpinLimit) {  // not expected; reported, never a hang
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(st + d, kFlagIncl | ((excl + cnt[k]) & kValMask), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    gpos[d] = db[d] + excl;
  }
  __syncthreads();
  // stage the tile in digit order, then write it out: a digit's keys of the
  // tile go to consecutive positions
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int i = wave * kWaveKeys + r * 64 + lane;
    if (i < count) {
      const uint32_t d = (mkey[r] >> shift) & (kDigits - 1);
      const uint32_t li = toff[d] + hw[wave][d] + mrank[r];
      skey[li] = mkey[r];
      sval[li] = mval[r];
    }
  }
  __syncthreads();
  uint32_t* kout = keys_out + s.tri_base;
  uint32_t* vout = vals_out + s.tri_base;
  for (int i = t; i < count; i += kSortThreads) {
    const uint32_t key = skey[i];
    const uint32_t d = (key >> shift) & (kDigits - 1);
    const uint32_t pos = gpos[d] + (uint32_t)i - toff[d];
    kout[pos] = key;
    vout[pos] = sval[i];
  }
}

// -------------------------------------------------------------- Karras ----
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
  uint32_t* ninfo;     // per global node: 2 * local parent + side (kNoParent: a root), | kCross
  uint32_t* leaf_slot; // per global leaf: 2 * local parent + side
  uint8_t* need_pre;   // per global leaf: a crossing child range ends here (prefix union wanted)
  uint8_t* need_suf;   // per global leaf: a crossing child range starts here (suffix union wanted)
  float* pre;          // per global leaf, where wanted: union of its chunk's leaves up to it (6 floats)
  float* suf;          // per global leaf, where wanted: union from it to its chunk's end
  float* ctot;         // per chunk: union of its leaves
};

FLG_D bool crosses(int a, int b) { return a / kChunk != b / kChunk; }

// RadixBVHBuilder::delta (radix_bvh.hpp:68-74)
FLG_D int delta(const uint32_t* code, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = code[i], c = code[j];
  if (a == c) return __clz((int)((uint32_t)i ^ (uint32_t)j)) + 32;
  return __clz((int)(a ^ c));
}

__global__ __launch_bounds__(kBlock) void k_emit(const SegDesc* segs, const int32_t* blk_seg, BuildOut o) {
  const int b = blockIdx.x;
  const int sid = blk_seg[b];
  const SegDesc s = segs[sid];
  const int n = s.ntris;
  const int i = (b - s.blk_base) * kBlock + (int)threadIdx.x;
  // a single triangle has no internal node: its leaf's parent field is 0, the
  // zero-initialised value the other builds leave in unused fields
  if (n == 1 && i == 0) o.leaf_parent[s.tri_base] = 0;
  bool top = false;
  if (i < n - 1) {
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
    top = crosses(lo, hi);
    const uint32_t cross = top ? kCross : 0u;
    const uint32_t slot0 = 2u * (uint32_t)i;
    if (i == 0) {
      o.parent[nb] = 0;
      o.ninfo[nb] = kNoParent | cross;
    }
    if (lleaf) {
      o.leaf_parent[tb + gamma] = i;
      o.leaf_slot[tb + gamma] = slot0;
    } else {
      o.parent[nb + gamma] = i;
      o.ninfo[nb + gamma] = slot0 | (crosses(lo, gamma) ? kCross : 0u);
    }
    if (rleaf) {
      o.leaf_parent[tb + gamma + 1] = i;
      o.leaf_slot[tb + gamma + 1] = slot0 + 1;
    } else {
      o.parent[nb + gamma + 1] = i;
      o.ninfo[nb + gamma + 1] = (slot0 + 1) | (crosses(gamma + 1, hi) ? kCross : 0u);
    }
    int4 refs;
    refs.x = lleaf ? (int)(kLeafBit | (uint32_t)gamma) : gamma;
    refs.y = rleaf ? (int)(kLeafBit | (uint32_t)(gamma + 1)) : gamma + 1;
    refs.z = 0;
    refs.w = 0;
    o.nodes[nb + i].d = refs;
    // the crossing children of a crossing node take their boxes from the
    // prefix and suffix unions at their range's ends
    if (top) {
      if (!lleaf && crosses(lo, gamma)) {
        o.need_suf[tb + lo] = 1;
        o.need_pre[tb + gamma] = 1;
      }
      if (!rleaf && crosses(gamma + 1, hi)) {
        o.need_suf[tb + gamma + 1] = 1;
        o.need_pre[tb + hi] = 1;
      }
    }
  }
}

FLG_D void store_slot(BNode* node, int side, V3 lo, V3 hi) {
  float* f = reinterpret_cast<float*>(node) + (side ? 6 : 0);
  f[0] = lo.x; f[1] = lo.y; f[2] = lo.z;
  f[3] = hi.x; f[4] = hi.y; f[5] = hi.z;
}

struct Box {
  V3 lo, hi;
};
FLG_D Box empty_box() {
  return Box{v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f),
             v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f)};
}
FLG_D Box join(const Box& a, const Box& b) { return Box{vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
FLG_D Box shfl_box(const Box& b, int src) {
  return Box{v3(__shfl(b.lo.x, src, 64), __shfl(b.lo.y, src, 64), __shfl(b.lo.z, src, 64)),
             v3(__shfl(b.hi.x, src, 64), __shfl(b.hi.y, src, 64), __shfl(b.hi.z, src, 64))};
}
FLG_D void put_box(float* f, const Box& b) {
  f[0] = b.lo.x; f[1] = b.lo.y; f[2] = b.lo.z;
  f[3] = b.hi.x; f[4] = b.hi.y; f[5] = b.hi.z;
}
FLG_D Box get_box(const float* f) { return Box{v3(f[0], f[1], f[2]), v3(f[3], f[4], f[5])}; }

// Bottom-up refit inside one chunk of kChunk leaves (the reference climbs
// leaf-to-root with fetch-and-add flags, radix_bvh.cpp:163-199; min/max are
// exact, so any children-before-parent order gives the same bounds).
// Thread t takes leaves A + 4t .. A + 4t + 3 of the chunk.
__global__ __launch_bounds__(kChunkThreads) void k_refit_chunks(const SegDesc* segs, const int32_t* chunk_seg,
                                                                BuildOut o) {
  __shared__ float sbox[kChunk][12];            // child boxes (L, R) of the chunk's nodes, at x - A
  __shared__ uint32_t sinfo[kChunk];            // ninfo of the chunk's nodes
  __shared__ uint32_t arrive[kChunk / 4];       // one byte per node: children arrived
  __shared__ float wpre[kChunkThreads / 64][6], wsuf[kChunkThreads / 64][6];
  const int b = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const SegDesc s = segs[chunk_seg[b]];
  const int chunk = b - s.chunk_base;
  const int A = chunk * kChunk;
  const int end = min(A + kChunk, s.ntris);
  const int nb = s.node_base, tb = s.tri_base;
  const int nodes_here = min(A + kChunk, s.ntris - 1) - A;  // nodes A .. ; may be <= 0
  for (int j = t; j < kChunk; j += kChunkThreads) sinfo[j] = j < nodes_here ? o.ninfo[nb + A + j] : kCross;
  for (int j = t; j < kChunk / 4; j += kChunkThreads) arrive[j] = 0u;
  __syncthreads();
  // is local node x one of this chunk's inside nodes (range within the chunk)?
  auto inside = [&](uint32_t x) { return x - (uint32_t)A < (uint32_t)kChunk && !(sinfo[x - A] & kCross); };

  // leaves: every load first (the stores below could alias them, so they
  // would otherwise wait for each other), then the soup records and the
  // boxes into the parents' slots
  Box lb[kChunkLeaves];
  uint32_t lslot[kChunkLeaves];
  uint32_t tri[kChunkLeaves];
  V3 va[kChunkLeaves], vb[kChunkLeaves], vc[kChunkLeaves];
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    const int p = A + kChunkLeaves * t + k;
    tri[k] = p < end ? o.order[tb + p] : 0u;
    lslot[k] = p < end && s.ntris >= 2 ? o.leaf_slot[tb + p] : kNoParent;
  }
  int32_t vi[kChunkLeaves][3];
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    const int p = A + kChunkLeaves * t + k;
    const int32_t* ip = s.ind + 3 * (int64_t)tri[k];
    vi[k][0] = p < end ? ip[0] : 0;
    vi[k][1] = p < end ? ip[1] : 0;
    vi[k][2] = p < end ? ip[2] : 0;
  }
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    const int p = A + kChunkLeaves * t + k;
    if (p < end) {
      va[k] = load_vert(s.verts, vi[k][0]);
      vb[k] = load_vert(s.verts, vi[k][1]);
      vc[k] = load_vert(s.verts, vi[k][2]);
    } else {
      va[k] = vb[k] = vc[k] = v3(0.0f, 0.0f, 0.0f);
    }
  }
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    const int p = A + kChunkLeaves * t + k;
    lb[k] = empty_box();
    if (p >= end) continue;
    BTri rec;
    rec.v0 = make_float4(va[k].x, va[k].y, va[k].z, u2f(tri[k]));
    rec.v1 = make_float4(vb[k].x, vb[k].y, vb[k].z, 0.0f);
    rec.v2 = make_float4(vc[k].x, vc[k].y, vc[k].z, 0.0f);
    o.tris[tb + p] = rec;
    lb[k] = Box{vmin(vmin(va[k], vb[k]), vc[k]), vmax(vmax(va[k], vb[k]), vc[k])};
    if (lslot[k] == kNoParent) continue;
    const uint32_t x = lslot[k] >> 1;
    store_slot(o.nodes + nb + x, lslot[k] & 1, lb[k].lo, lb[k].hi);
    if (inside(x)) put_box(&sbox[x - A][(lslot[k] & 1) * 6], lb[k]);
  }

  // prefix and suffix unions of the chunk's leaf boxes (for the crossing
  // nodes), and the chunk's union
  Box tp = lb[0], ts = lb[kChunkLeaves - 1];
#pragma unroll
  for (int k = 1; k < kChunkLeaves; ++k) {
    tp = join(tp, lb[k]);
    ts = join(ts, lb[kChunkLeaves - 1 - k]);
  }
  Box ip = tp, is = ts;  // inclusive over the wave's lanes up to / from this one
  for (int off = 1; off < 64; off <<= 1) {
    const Box up = shfl_box(ip, lane >= off ? lane - off : lane);
    const Box dn = shfl_box(is, lane + off < 64 ? lane + off : lane);
    if (lane >= off) ip = join(ip, up);
    if (lane + off < 64) is = join(is, dn);
  }
  if (lane == 63) put_box(wpre[wave], ip);
  if (lane == 0) put_box(wsuf[wave], is);
  __syncthreads();
  Box ep = empty_box(), es = empty_box();  // exclusive: the leaves before / after this thread's
  for (int w = 0; w < wave; ++w) ep = join(ep, get_box(wpre[w]));
  for (int w = wave + 1; w < kChunkThreads / 64; ++w) es = join(es, get_box(wsuf[w]));
  {
    const Box xp = shfl_box(ip, lane > 0 ? lane - 1 : 0), xs = shfl_box(is, lane < 63 ? lane + 1 : 63);
    if (lane > 0) ep = join(ep, xp);
    if (lane < 63) es = join(es, xs);
  }
  Box run = ep;
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    const int p = A + kChunkLeaves * t + k;
    run = join(run, lb[k]);
    if (p < end && o.need_pre[tb + p]) put_box(o.pre + 6 * (int64_t)(tb + p), run);
  }
  run = es;
#pragma unroll
  for (int k = kChunkLeaves - 1; k >= 0; --k) {
    const int p = A + kChunkLeaves * t + k;
    run = join(run, lb[k]);
    if (p < end && o.need_suf[tb + p]) put_box(o.suf + 6 * (int64_t)(tb + p), run);
  }
  if (t == 0) put_box(o.ctot + 6 * (int64_t)b, join(ep, run));  // thread 0: nothing before it
  if (s.ntris < 2) return;

  // climb through the chunk's inside nodes: the second child to arrive
  // merges the node (children's boxes in LDS) and moves up
#pragma unroll
  for (int k = 0; k < kChunkLeaves; ++k) {
    if (lslot[k] == kNoParent) continue;
    uint32_t x = lslot[k] >> 1;
    while (inside(x)) {
      const uint32_t xl = x - (uint32_t)A;
      const uint32_t prev = __hip_atomic_fetch_add(&arrive[xl >> 2], 1u << (8 * (xl & 3)), __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
      if (((prev >> (8 * (xl & 3))) & 0xFFu) == 0) break;  // the other child comes later
      const uint32_t info = sinfo[xl];
      const uint32_t ps = info & ~kCross;
      if (ps == kNoParent) break;  // the segment's root
      // BVHBound{} merged with the right child, then the left (radix_bvh.cpp:190-192)
      const float* f = sbox[xl];
      const V3 lo = vmin(vmin(v3(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f), v3(f[6], f[7], f[8])),
                         v3(f[0], f[1], f[2]));
      const V3 hi = vmax(vmax(v3(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f),
                              v3(f[9], f[10], f[11])),
                         v3(f[3], f[4], f[5]));
      const uint32_t y = ps >> 1;
      store_slot(o.nodes + nb + y, ps & 1, lo, hi);
      if (inside(y)) put_box(&sbox[y - A][(ps & 1) * 6], Box{lo, hi});
      x = y;
    }
  }
}

// Crossing nodes: each child range that crosses a chunk boundary too gets its
// box from the suffix union at its start, the unions of the whole chunks in
// between and the prefix union at its end.  A wave reads the crossing flags of
// 64 consecutive nodes and serves its crossing nodes one after another.
__global__ __launch_bounds__(kBlock) void k_refit_top(const SegDesc* segs, int nseg, int64_t nnodes, BuildOut o) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t w0 = ((int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64) * 64; w0 < nnodes; w0 += nwaves * 64) {
    const int64_t mine = w0 + lane;
    for (uint64_t m = __ballot(mine < nnodes && (o.ninfo[mine] & kCross)); m; m &= m - 1) {
      const int x = (int)(w0 + __ffsll((unsigned long long)m) - 1);
      const SegDesc s = segs[seg_of<&SegDesc::node_base>(segs, nseg, x)];
      const int lo = o.rlo[x], hi = o.rhi[x], gamma = o.split[x];
      const uint8_t lf = o.leafflags[x];
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int a = side ? gamma + 1 : lo, z = side ? hi : gamma;
        if ((lf >> side) & 1 || !crosses(a, z)) continue;  // done by k_refit_chunks
        Box bx = empty_box();
        for (int c = a / kChunk + 1 + lane; c < z / kChunk; c += 64)
          bx = join(bx, get_box(o.ctot + 6 * (int64_t)(s.chunk_base + c)));
        for (int off = 32; off >= 1; off >>= 1) bx = join(bx, shfl_box(bx, lane ^ off));
        if (lane == 0) {
          bx = join(bx, get_box(o.suf + 6 * (int64_t)(s.tri_base + a)));
          bx = join(bx, get_box(o.pre + 6 * (int64_t)(s.tri_base + z)));
          store_slot(o.nodes + x, side, bx.lo, bx.hi);
        }
      }
    }
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ----------------------------------------------------------- output pool ---
// A build's output block comes from the context's pool when one fits
// (smallest first, at most 4x the request plus 64 MiB), else from hipMalloc
// with an eighth to spare; a destroyed BVH's block goes back to the pool.
void* pool_take(flg_ctx* ctx, size_t bytes, size_t* got, int* st) {
  auto& pool = ctx->bvh_pool;
  int best = -1;
  for (int k = 0; k < (int)pool.size(); ++k)
    if (pool[k].second >= bytes && pool[k].second <= 4 * bytes + ((size_t)64 << 20) &&
        (best < 0 || pool[k].second < pool[best].second))
      best = k;
  if (best >= 0) {
    void* p = pool[best].first;
    *got = pool[best].second;
    pool.erase(pool.begin() + best);
    *st = FLG_OK;
    return p;
  }
  const size_t cap = align256(bytes + bytes / 8 + 256);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cap);
  if (e == hipErrorOutOfMemory && !pool.empty()) {  // give the pooled blocks back and try again
    (void)hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : pool) (void)hipFree(b.first);
    pool.clear();
    e = hipMalloc(&p, cap);
  }
  if (e != hipSuccess) {
    *st = check_hip(ctx, e, "hipMalloc(lbvh output)");
    return nullptr;
  }
  *got = cap;
  *st = FLG_OK;
  return p;
}

}  // namespace

void bvh_pool_give(flg_ctx* ctx, void* p, size_t bytes) {
  auto& pool = ctx->bvh_pool;
  pool.emplace_back(p, bytes);
  if (pool.size() > 8) {  // keep the eight most recent
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(pool.front().first);
    pool.erase(pool.begin());
  }
}

void bvh_pool_release(flg_ctx* ctx) {
  for (auto& b : ctx->bvh_pool) (void)hipFree(b.first);
  ctx->bvh_pool.clear();
}

// Builds every segment; out[i] receives the handle of mesh i.
static int build_segments(flg_ctx* ctx, const flg_mesh_desc* meshes, int nseg, uint32_t flags,
                          flg_bvh** out) {
  if (!ctx || !meshes || nseg <= 0 || !out) return set_error(ctx, FLG_ERR_INVALID, "lbvh build: bad arguments");
  if (flags > FLG_NORMAL_COMPAT) return set_error(ctx, FLG_ERR_INVALID, "lbvh build: unknown flags");
  FLG_CHECK_HIP(ctx, hipSetDevice(ctx->device));
  // ---- host-side segment table
  std::vector<SegDesc> segs(nseg);
  int64_t tri_total = 0, node_total = 0, blk_total = 0, tile_total = 0, hblk_total = 0, chunk_total = 0;
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
    d.blk_base = (int32_t)blk_total;
    d.tile_base = (int32_t)tile_total;
    d.hblk_base = (int32_t)hblk_total;
    d.chunk_base = (int32_t)chunk_total;
    d.pad = 0;
    tri_total += m.ntris;
    node_total += m.ntris > 0 ? m.ntris - 1 : 0;
    blk_total += (m.ntris + kBlock - 1) / kBlock;
    tile_total += (m.ntris + kTile - 1) / kTile;
    hblk_total += (m.ntris + kHistBlock - 1) / kHistBlock;
    chunk_total += (m.ntris + kChunk - 1) / kChunk;
  }
  if (tri_total >= (int64_t)1 << 31) return set_error(ctx, FLG_ERR_RANGE, "lbvh build: batch too large");

  // ---- output block (shared by the batch's handles), from the pool
  const size_t N = (size_t)tri_total, NI = (size_t)node_total;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
  size_t o_nodes = take(NI * sizeof(BNode)), o_tris = take(N * sizeof(BTri));
  size_t o_morton = take(N * 4), o_msorted = take(N * 4), o_order = take(N * 4);
  size_t o_split = take(NI * 4), o_rlo = take(NI * 4), o_rhi = take(NI * 4), o_parent = take(NI * 4);
  size_t o_lf = take(NI), o_lp = take(N * 4), o_gb = take((size_t)nseg * 6 * 4);
  const size_t total_bytes = off ? off : 256;
  int st = FLG_OK;
  size_t block_bytes = 0;
  void* blockp = pool_take(ctx, total_bytes, &block_bytes, &st);
  if (!blockp) return st;
  char* base = (char*)blockp;
  auto fail = [&](int code) { bvh_pool_give(ctx, blockp, block_bytes); return code; };

  // ---- scratch: [segment table][bound partials][keys, values x2][node
  // info, leaf slots, prefix and suffix unions, chunk unions, crossing
  // nodes][zeroed: histograms, look-back status, tile counters, error flag,
  // crossing-node count, prefix/suffix request flags]
  size_t s_off = 0;
  auto stake = [&](size_t bytes) { size_t o = s_off; s_off = align256(s_off + bytes); return o; };
  const size_t so_segs = stake(nseg * sizeof(SegDesc));
  const size_t so_part = stake(blk_total * sizeof(Bounds6));
  const size_t so_k0 = stake(N * 4), so_v0 = stake(N * 4), so_k1 = stake(N * 4), so_v1 = stake(N * 4);
  const size_t so_ninfo = stake(NI * 4), so_lslot = stake(N * 4);
  const size_t so_pre = stake(N * 24), so_suf = stake(N * 24), so_ctot = stake(chunk_total * 24);
  const size_t so_blkseg = stake(blk_total * 4), so_chunkseg = stake(chunk_total * 4);
  const size_t hist_words = (size_t)nseg * kPasses * kDigits;
  const size_t status_words = (size_t)kPasses * tile_total * kDigits;
  const size_t flag_words = (2 * N + 3) / 4;  // two byte arrays
  const size_t zero_words = hist_words + status_words + kPasses + 1 + flag_words;
  const size_t so_zero = stake(zero_words * 4);
  char* sc = (char*)scratch(ctx, "lbvh_build", s_off ? s_off : 256, &st);
  if (!sc) return fail(st);
  SegDesc* d_segs = (SegDesc*)(sc + so_segs);
  if (hipMemcpyAsync(d_segs, segs.data(), nseg * sizeof(SegDesc), hipMemcpyHostToDevice, ctx->stream))
    return fail(set_error(ctx, FLG_ERR_HIP, "lbvh build: upload of the segment table failed"));
  Bounds6* d_part = (Bounds6*)(sc + so_part);
  uint32_t* k0 = (uint32_t*)(sc + so_k0);
  uint32_t* v0 = (uint32_t*)(sc + so_v0);
  uint32_t* k1 = (uint32_t*)(sc + so_k1);
  uint32_t* v1 = (uint32_t*)(sc + so_v1);
  uint32_t* zero = (uint32_t*)(sc + so_zero);
  int32_t* blk_seg = (int32_t*)(sc + so_blkseg);
  int32_t* chunk_seg = (int32_t*)(sc + so_chunkseg);
  uint32_t* hist = zero;
  uint32_t* status = hist + hist_words;
  uint32_t* counters = status + status_words;
  uint32_t* err = counters + kPasses;
  uint8_t* want = (uint8_t*)(err + 1);

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
  o.ninfo = (uint32_t*)(sc + so_ninfo);
  o.leaf_slot = (uint32_t*)(sc + so_lslot);
  o.need_pre = want;
  o.need_suf = want + N;
  o.pre = (float*)(sc + so_pre);
  o.suf = (float*)(sc + so_suf);
  o.ctot = (float*)(sc + so_ctot);
  float* gb = (float*)(base + o_gb);

  const int fill_grid = (int)std::max<int64_t>(1, std::min<int64_t>(div_up((int64_t)zero_words, 256),
                                                                     (int64_t)ctx->num_cus * 4));
  k_fill_u32<<<fill_grid, 256, 0, ctx->stream>>>(zero, (int64_t)zero_words);
  if (blk_total > 0) {
    k_seg_maps<<<div_up(std::max(blk_total, chunk_total), kBlock), kBlock, 0, ctx->stream>>>(
        d_segs, nseg, blk_total, chunk_total, blk_seg, chunk_seg);
    {
      ScopedTimer tm(ctx, "lbvh_bound");
      k_bound_partial<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, blk_seg, d_part);
      k_bound_final<<<nseg, kBlock, 0, ctx->stream>>>(d_segs, d_part, gb);
    }
    {
      ScopedTimer tm(ctx, "lbvh_morton");
      k_morton<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, blk_seg, gb, morton, k0, v0);
    }
    {
      ScopedTimer tm(ctx, "lbvh_sort");
      k_digit_hist<<<(int)hblk_total, kBlock, 0, ctx->stream>>>(d_segs, nseg, k0, hist);
      k_digit_scan<<<nseg * kPasses, kDigits, 0, ctx->stream>>>(hist);
      const uint32_t* kin[kPasses] = {k0, k1, k0};
      const uint32_t* vin[kPasses] = {v0, v1, v0};
      uint32_t* kout[kPasses] = {k1, k0, o.morton_sorted};
      uint32_t* vout[kPasses] = {v1, v0, o.order};
      for (int pass = 0; pass < kPasses; ++pass)
        k_onesweep<<<(int)tile_total, kSortThreads, 0, ctx->stream>>>(
            d_segs, nseg, pass, kin[pass], vin[pass], kout[pass], vout[pass], hist,
            status + (size_t)pass * tile_total * kDigits, counters + pass, err);
    }
    {
      ScopedTimer tm(ctx, "lbvh_emit");
      k_emit<<<(int)blk_total, kBlock, 0, ctx->stream>>>(d_segs, blk_seg, o);
    }
    {
      ScopedTimer tm(ctx, "lbvh_refit");
      k_refit_chunks<<<(int)chunk_total, kChunkThreads, 0, ctx->stream>>>(d_segs, chunk_seg, o);
      if (NI > 0) {
        const int tgrid = (int)std::max<int64_t>(1, std::min<int64_t>(div_up((int64_t)NI, kBlock),
                                                                       (int64_t)ctx->num_cus * 8));
        k_refit_top<<<tgrid, kBlock, 0, ctx->stream>>>(d_segs, nseg, (int64_t)NI, o);
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(check_hip(ctx, e, "lbvh build kernels"));
  }
  std::vector<float> hgb((size_t)nseg * 6, 0.0f);
  uint32_t herr = 0;
  if (hipMemcpyAsync(hgb.data(), gb, hgb.size() * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(set_error(ctx, FLG_ERR_HIP, "lbvh build: synchronisation failed"));
  if (herr) return fail(set_error(ctx, FLG_ERR_HIP, "lbvh build: a sort tile's look-back did not complete"));

  SharedBlock* sb = new SharedBlock;
  sb->ptr = blockp;
  sb->bytes = block_bytes;
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
      // back to the context's pool: later work on the context's stream is
      // ordered after every use of the block, so no synchronisation
      bvh_pool_give(bvh->ctx, sb->ptr, sb->bytes);
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
