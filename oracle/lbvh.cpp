// TEST INFRASTRUCTURE — CPU restatement of fledge's radix-tree LBVH build and
// closest-hit traversal (src/experimental/radix_bvh.{hpp,cpp},
// intersector.hpp, experimental_utils.hpp, base_bvh.hpp).  Used only as the
// parity oracle; see oracle.hpp.
#include <algorithm>
#include <numeric>

#include "oracle_impl.hpp"

namespace orc {

// BVHBound::merge base_bvh.hpp:56-60 ; Triangle::getBound base_bvh.hpp:90-102
static inline void tri_bound(const float* vp, const int32_t* ind, int32_t tri, V3& lo, V3& hi) {
  const int32_t* v = ind + 3 * (int64_t)tri;
  V3 p0 = v3(vp[3 * (int64_t)v[0]], vp[3 * (int64_t)v[0] + 1], vp[3 * (int64_t)v[0] + 2]);
  V3 p1 = v3(vp[3 * (int64_t)v[1]], vp[3 * (int64_t)v[1] + 1], vp[3 * (int64_t)v[1] + 2]);
  V3 p2 = v3(vp[3 * (int64_t)v[2]], vp[3 * (int64_t)v[2] + 1], vp[3 * (int64_t)v[2] + 2]);
  lo = vmin(vmin(p0, p1), p2);
  hi = vmax(vmax(p0, p1), p2);
}

// LeftShift3 experimental_utils.hpp:33-40 (value 1024 clamped to 1023)
uint32_t left_shift3(uint32_t x) {
  if (x == (1u << 10)) --x;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
// EncodeMorton3 experimental_utils.hpp:42-45: z at bit 2, y at bit 1, x at bit 0;
// the float -> uint32 conversion truncates.
uint32_t encode_morton3(V3 q) {
  return (left_shift3((uint32_t)q.z) << 2) | (left_shift3((uint32_t)q.y) << 1) |
         left_shift3((uint32_t)q.x);
}

// Next2Pow experimental_utils.hpp:27-31 (2^ceil(log2 n)), restated in integers
static inline int next2pow(int n) {
  if (n <= 1) return 1;
  return 1 << (32 - __builtin_clz((uint32_t)(n - 1)));
}

// RadixBVHBuilder::delta radix_bvh.hpp:68-74 (on the sorted codes)
static inline int delta(const uint32_t* code, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  if (code[i] == code[j]) return __builtin_clz((uint32_t)(i ^ j)) + 32;
  return __builtin_clz(code[i] ^ code[j]);
}

void lbvh_build(Lbvh& b, const float* verts, const int32_t* ind, int32_t ntris, int32_t nverts) {
  (void)nverts;
  b.n = ntris;
  b.verts = verts;
  b.ind = ind;
  const int n = ntris;
  b.morton_orig.assign(n, 0);
  b.morton.assign(n, 0);
  b.order.resize(n);
  if (n == 0) return;
  // prestage radix_bvh.cpp:49-60: global bound (min/max are exact, so any
  // reduction order gives the same result)
  V3 glo = v3(FLT_MAX, FLT_MAX, FLT_MAX), ghi = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
  for (int i = 0; i < n; ++i) {
    V3 lo, hi;
    tri_bound(verts, ind, i, lo, hi);
    glo = vmin(glo, lo);
    ghi = vmax(ghi, hi);
  }
  b.glo = glo;
  b.ghi = ghi;
  // radix_bvh.cpp:62-73: centroid of the triangle AABB, normalised by the
  // global extent, times 1024, truncated.  A zero extent (planar input) is
  // replaced by 1 (the reference divides by zero and asserts).
  V3 edges = ghi - glo;
  if (edges.x == 0.0f) edges.x = 1.0f;
  if (edges.y == 0.0f) edges.y = 1.0f;
  if (edges.z == 0.0f) edges.z = 1.0f;
  for (int i = 0; i < n; ++i) {
    V3 lo, hi;
    tri_bound(verts, ind, i, lo, hi);
    V3 c = (lo + hi) / 2.0f;
    V3 q = (c - glo) / edges;
    b.morton_orig[i] = encode_morton3(q * (float)(1 << 10));
  }
  // radix_bvh.cpp:77-81 sorts by code with an unstable sort; ties are
  // resolved here by original index (stable sort).
  std::iota(b.order.begin(), b.order.end(), 0u);
  std::stable_sort(b.order.begin(), b.order.end(), [&](uint32_t a, uint32_t c) {
    return b.morton_orig[a] < b.morton_orig[c];
  });
  for (int i = 0; i < n; ++i) b.morton[i] = b.morton_orig[b.order[i]];

  karras_emit(b);
  const int ni = n - 1;
  // Bottom-up refit radix_bvh.cpp:163-199.  Min/max are exact, so the bound of
  // every node equals the union of its children's bounds whatever the order in
  // which the climbing threads arrive; evaluate children before parents.
  std::vector<int> post;
  post.reserve(std::max(ni, 0));
  if (ni > 0) {
    std::vector<int> stk{0};
    while (!stk.empty()) {
      int x = stk.back();
      stk.pop_back();
      post.push_back(x);
      if (!b.lleaf[x]) stk.push_back(b.split[x]);
      if (!b.rleaf[x]) stk.push_back(b.split[x] + 1);
    }
  }
  for (auto it = post.rbegin(); it != post.rend(); ++it) {
    int x = *it;
    V3 alo, ahi, clo, chi;
    int g = b.split[x];
    if (b.lleaf[x]) tri_bound(verts, ind, (int32_t)b.order[g], alo, ahi);
    else { alo = b.nlo[g]; ahi = b.nhi[g]; }
    if (b.rleaf[x]) tri_bound(verts, ind, (int32_t)b.order[g + 1], clo, chi);
    else { clo = b.nlo[g + 1]; chi = b.nhi[g + 1]; }
    V3 lo = v3(FLT_MAX, FLT_MAX, FLT_MAX), hi = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    lo = vmin(lo, clo); hi = vmax(hi, chi);  // merge(other) then merge(subnode)
    lo = vmin(lo, alo); hi = vmax(hi, ahi);
    b.nlo[x] = lo;
    b.nhi[x] = hi;
  }
}

// parallelBuilder node emission radix_bvh.cpp:118-161 over b.morton (sorted).
void karras_emit(Lbvh& b) {
  const int n = b.n;
  const int ni = n - 1;
  b.split.assign(std::max(ni, 0), 0);
  b.dir.assign(std::max(ni, 0), 0);
  b.rlo.assign(std::max(ni, 0), 0);
  b.rhi.assign(std::max(ni, 0), 0);
  b.parent.assign(std::max(ni, 0), 0);  // zero-initialised: the root's parent reads 0
  b.lleaf.assign(std::max(ni, 0), 0);
  b.rleaf.assign(std::max(ni, 0), 0);
  b.nlo.assign(std::max(ni, 0), v3(FLT_MAX, FLT_MAX, FLT_MAX));
  b.nhi.assign(std::max(ni, 0), v3(-FLT_MAX, -FLT_MAX, -FLT_MAX));
  b.leaf_parent.assign(n, 0);
  const uint32_t* code = b.morton.data();
  // parallelBuilder node emission radix_bvh.cpp:118-161
  for (int i = 0; i < ni; ++i) {
    int dl = delta(code, n, i, i + 1) - delta(code, n, i, i - 1);
    int d = dl == 0 ? 0 : (dl > 0 ? 1 : -1);  // Sign math_utils.h:174-176
    b.dir[i] = d;
    int dmin = delta(code, n, i, i - d);
    int lmax = 2;
    while (delta(code, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t >>= 1)
      if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(code, n, i, j);
    int s = 0;
    for (int t = next2pow(l); t >= 1; t >>= 1)
      if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
    int gamma = i + s * d + std::min(d, 0);
    b.split[i] = gamma;
    b.rlo[i] = std::min(i, j);
    b.rhi[i] = std::max(i, j);
    b.lleaf[i] = (std::min(i, j) == gamma);
    b.rleaf[i] = (std::max(i, j) == gamma + 1);
    if (b.lleaf[i]) b.leaf_parent[gamma] = i; else b.parent[gamma] = i;
    if (b.rleaf[i]) b.leaf_parent[gamma + 1] = i; else b.parent[gamma + 1] = i;
  }
}

// BoundIntersect intersector.hpp:103-120 (t0 = 0, t1 = FLT_MAX)
bool bound_intersect(V3 lo, V3 hi, V3 o, V3 d, float& tnear) {
  float t0 = 0.0f, t1 = FLT_MAX;
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / comp(d, i);
    float tn = (comp(lo, i) - comp(o, i)) * inv;
    float tf = (comp(hi, i) - comp(o, i)) * inv;
    if (tn > tf) std::swap(tn, tf);
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  tnear = t0;
  return true;
}

// msub intersector.hpp:41-43, fused per the declared policy (embree fmsub)
static inline float msub(float a, float b, float c) { return std::fmaf(a, b, -c); }

// StableTriangleNormal intersector.hpp:45-60
static inline V3 stable_triangle_normal(V3 a, V3 b, V3 c) {
  const float ab_x = a.z * b.y, ab_y = a.x * b.z, ab_z = a.y * b.x;
  const float bc_x = b.z * c.y, bc_y = b.x * c.z, bc_z = b.y * c.x;
  V3 cab = v3(msub(a.y, b.z, ab_x), msub(a.z, b.x, ab_y), msub(a.x, b.y, ab_z));
  V3 cbc = v3(msub(b.y, c.z, bc_x), msub(b.z, c.x, bc_y), msub(b.x, c.y, bc_z));
  const bool sx = std::fabs(ab_x) < std::fabs(bc_x);
  const bool sy = std::fabs(ab_y) < std::fabs(bc_y);
  const bool sz = std::fabs(ab_z) < std::fabs(bc_z);
  return v3(sx ? cab.x : cbc.x, sy ? cab.y : cbc.y, sz ? cab.z : cbc.z);
}

// StableNormalize vector.h:495-504, in double then rounded to float.
// compat: the reference's squared norm x^2 + y^2 + z + z; fixed: x^2 + y^2 + z^2.
static inline V3 stable_normalize(V3 v, bool compat) {
  double x = v.x, y = v.y, z = v.z;
  double sq = compat ? (x * x + y * y + z + z) : (x * x + y * y + z * z);
  double r = std::sqrt(sq);
  return v3((float)(x / r), (float)(y / r), (float)(z / r));
}

// PlueckerTriangleIntersect intersector.hpp:63-101
bool pluecker(V3 a, V3 b, V3 c, V3 O, V3 D, bool compat, float* thit, V3* ng) {
  const float ulp = FLT_EPSILON;
  const V3 v0 = a - O, v1 = b - O, v2 = c - O;
  const V3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
  const float U = dot(cross(e0, v2 + v0), D);
  const float V = dot(cross(e1, v0 + v1), D);
  const float W = dot(cross(e2, v1 + v2), D);
  const float UVW = U + V + W;
  const float eps = ulp * std::fabs(UVW);
  bool valid = smin(U, smin(V, W)) >= -eps || smax(U, smax(V, W)) <= eps;
  if (!valid) return false;
  const V3 Ng = stable_normalize(stable_triangle_normal(e0, e1, e2), compat);
  const float dd = dot(Ng, D);
  const float den = dd + dd;  // Twice math_utils.h:164-167
  const float tt = dot(v0, Ng);
  const float T = tt + tt;
  const float t = T / den;
  if (!(0.0f <= t)) return false;
  if (!(den != 0.0f)) return false;
  if (ng) *ng = Ng;
  if (thit) *thit = t;
  return true;
}

// RadixBVHBuilder::recursiveIntersect radix_bvh.cpp:202-245, as an explicit
// stack in the same visit order: node box test and prune (best < tnear), then
// the node's leaf children (left then right; equal t replaces), then the left
// internal subtree, then the right one.
static void visit(const Lbvh& b, int node, V3 o, V3 d, bool compat, float& best, int32_t& hit,
                  V3& hng, TraceCounters* ctr) {
  float tnear;
  if (ctr) ctr->nodes++;
  if (!bound_intersect(b.nlo[node], b.nhi[node], o, d, tnear)) return;
  if (best < tnear) return;
  int g = b.split[node];
  for (int side = 0; side < 2; ++side) {
    bool leaf = side == 0 ? b.lleaf[node] : b.rleaf[node];
    if (!leaf) continue;
    uint32_t tri = b.order[g + side];
    const int32_t* v = b.ind + 3 * (int64_t)tri;
    const float* P = b.verts;
    V3 pa = v3(P[3 * (int64_t)v[0]], P[3 * (int64_t)v[0] + 1], P[3 * (int64_t)v[0] + 2]);
    V3 pb = v3(P[3 * (int64_t)v[1]], P[3 * (int64_t)v[1] + 1], P[3 * (int64_t)v[1] + 2]);
    V3 pc = v3(P[3 * (int64_t)v[2]], P[3 * (int64_t)v[2] + 1], P[3 * (int64_t)v[2] + 2]);
    float th;
    V3 ng;
    if (ctr) ctr->tris++;
    if (!pluecker(pa, pb, pc, o, d, compat, &th, &ng)) continue;
    if (th > best) continue;
    best = th;
    hit = (int32_t)tri;
    hng = ng;
  }
  if (!b.lleaf[node]) visit(b, g, o, d, compat, best, hit, hng, ctr);
  if (!b.rleaf[node]) visit(b, g + 1, o, d, compat, best, hit, hng, ctr);
}

int32_t lbvh_trace(const Lbvh& b, V3 o, V3 d, float tmax, bool compat, float* t, V3* ng,
                   TraceCounters* ctr) {
  float best = tmax;
  int32_t hit = -1;
  V3 hng = v3(0, 0, 0);
  if (b.n == 1) {
    // Single triangle: no internal node exists (the reference would read
    // m_internal_nodes[0]); test the global box, then the triangle.
    float tn;
    if (bound_intersect(b.glo, b.ghi, o, d, tn) && !(best < tn)) {
      const int32_t* v = b.ind;
      const float* P = b.verts;
      V3 pa = v3(P[3 * (int64_t)v[0]], P[3 * (int64_t)v[0] + 1], P[3 * (int64_t)v[0] + 2]);
      V3 pb = v3(P[3 * (int64_t)v[1]], P[3 * (int64_t)v[1] + 1], P[3 * (int64_t)v[1] + 2]);
      V3 pc = v3(P[3 * (int64_t)v[2]], P[3 * (int64_t)v[2] + 1], P[3 * (int64_t)v[2] + 2]);
      float th;
      V3 g;
      if (pluecker(pa, pb, pc, o, d, compat, &th, &g) && !(th > best)) {
        best = th;
        hit = 0;
        hng = g;
      }
    }
  } else if (b.n > 1) {
    visit(b, 0, o, d, compat, best, hit, hng, ctr);
  }
  if (t) *t = hit >= 0 ? best : tmax;
  if (ng) *ng = hng;
  return hit;
}

// Brute force over every triangle with the same test: the closest t wins;
// *nties counts the other triangles hit at exactly that t (tie order is a
// property of the traversal, so tests use brute force only where it is 0).
int32_t brute_trace(const float* verts, const int32_t* ind, int32_t ntris, V3 o, V3 d, float tmax,
                    bool compat, float* t, int* nties) {
  float best = tmax;
  int32_t hit = -1;
  int ties = 0;
  for (int32_t i = 0; i < ntris; ++i) {
    const int32_t* v = ind + 3 * (int64_t)i;
    V3 pa = v3(verts[3 * (int64_t)v[0]], verts[3 * (int64_t)v[0] + 1], verts[3 * (int64_t)v[0] + 2]);
    V3 pb = v3(verts[3 * (int64_t)v[1]], verts[3 * (int64_t)v[1] + 1], verts[3 * (int64_t)v[1] + 2]);
    V3 pc = v3(verts[3 * (int64_t)v[2]], verts[3 * (int64_t)v[2] + 1], verts[3 * (int64_t)v[2] + 2]);
    float th;
    if (!pluecker(pa, pb, pc, o, d, compat, &th, nullptr)) continue;
    if (th > best) continue;
    if (hit >= 0 && th == best) { ++ties; continue; }
    best = th;
    hit = i;
    ties = 0;
  }
  if (t) *t = hit >= 0 ? best : tmax;
  if (nties) *nties = ties;
  return hit;
}

}  // namespace orc
