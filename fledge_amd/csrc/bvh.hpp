// Device layout of a built LBVH and the device-side geometry helpers shared by
// the builder and the traversal kernels.
//
// HBM layout (per BVH of n triangles):
//   nodes  [n-1] x 64 B  Karras internal node i: both child boxes + child refs
//                        (a ref with bit 31 set is a leaf = sorted position)
//   tris   [n]   x 48 B  triangle soup in Morton order: v0.xyz + original id,
//                        v1.xyz, v2.xyz (three 16-byte loads per test)
//   parity arrays (codes, order, split, ranges, parents, flags) kept for export
#pragma once
#include "common.hpp"

namespace flg {

struct alignas(16) BNode {
  float4 a;  // L.lo.x L.lo.y L.lo.z L.hi.x
  float4 b;  // L.hi.y L.hi.z R.lo.x R.lo.y
  float4 c;  // R.lo.z R.hi.x R.hi.y R.hi.z
  int4 d;    // refL refR - -
};
static_assert(sizeof(BNode) == 64, "node is 64 bytes");

struct alignas(16) BTri {
  float4 v0;  // w = original triangle id (bit pattern)
  float4 v1;
  float4 v2;
};
static_assert(sizeof(BTri) == 48, "triangle is 48 bytes");

constexpr uint32_t kLeafBit = 0x80000000u;

struct SharedBlock {
  void* ptr = nullptr;
  size_t bytes = 0;
  int refs = 0;
};

// The context's pool of LBVH output blocks (lbvh.hip): a destroyed BVH's
// block is kept for a later build; released with the context.
void bvh_pool_give(flg_ctx* ctx, void* p, size_t bytes);
void bvh_pool_release(flg_ctx* ctx);

}  // namespace flg

struct flg_bvh {
  flg_ctx* ctx = nullptr;
  int32_t n = 0;
  uint32_t flags = 0;
  float bound[6] = {0, 0, 0, 0, 0, 0};  // host copy of the global bound
  flg::BNode* nodes = nullptr;
  flg::BTri* tris = nullptr;
  uint32_t* morton = nullptr;         // original order
  uint32_t* morton_sorted = nullptr;
  uint32_t* order = nullptr;          // sorted position -> original triangle
  int32_t* split = nullptr;
  int32_t* rlo = nullptr;
  int32_t* rhi = nullptr;
  int32_t* parent = nullptr;
  uint8_t* leafflags = nullptr;       // bit 0: left child is a leaf, bit 1: right
  int32_t* leaf_parent = nullptr;
  float* gbound = nullptr;            // device copy of the global bound (6 floats)
  uint64_t bytes = 0;
  flg::SharedBlock* block = nullptr;
};

namespace flg {

// Pluecker watertight ray/triangle test (src/experimental/intersector.hpp:
// msub :41-43, StableTriangleNormal :45-60, PlueckerTriangleIntersect :63-101,
// StableNormalize src/common/vector.h:495-504 in double).
FLG_HD float msub(float a, float b, float c) { return __builtin_fmaf(a, b, -c); }

FLG_HD V3 stable_triangle_normal(V3 a, V3 b, V3 c) {
  const float ab_x = a.z * b.y, ab_y = a.x * b.z, ab_z = a.y * b.x;
  const float bc_x = b.z * c.y, bc_y = b.x * c.z, bc_z = b.y * c.x;
  V3 cab = v3(msub(a.y, b.z, ab_x), msub(a.z, b.x, ab_y), msub(a.x, b.y, ab_z));
  V3 cbc = v3(msub(b.y, c.z, bc_x), msub(b.z, c.x, bc_y), msub(b.x, c.y, bc_z));
  const bool sx = fabsf(ab_x) < fabsf(bc_x);
  const bool sy = fabsf(ab_y) < fabsf(bc_y);
  const bool sz = fabsf(ab_z) < fabsf(bc_z);
  return v3(sx ? cab.x : cbc.x, sy ? cab.y : cbc.y, sz ? cab.z : cbc.z);
}

FLG_HD V3 stable_normalize(V3 v, bool compat) {
  double x = v.x, y = v.y, z = v.z;
  double sq = compat ? (x * x + y * y + z + z) : (x * x + y * y + z * z);
  double r = __builtin_sqrt(sq);
  return v3((float)(x / r), (float)(y / r), (float)(z / r));
}

// Edge test only: true if the ray passes the UVW sign test.
FLG_HD bool pluecker_uvw(V3 v0, V3 v1, V3 v2, V3 D) {
  const V3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
  const float U = dot(cross(e0, v2 + v0), D);
  const float V = dot(cross(e1, v0 + v1), D);
  const float W = dot(cross(e2, v1 + v2), D);
  const float UVW = U + V + W;
  const float eps = 1.1920928955078125e-07f * fabsf(UVW);
  return smin(U, smin(V, W)) >= -eps || smax(U, smax(V, W)) <= eps;
}

// The part of the test after the edge test: returns true and t / Ng on a hit
// (t >= 0, den != 0).
FLG_HD bool pluecker_t(V3 a, V3 b, V3 c, V3 O, V3 D, bool compat, float* thit, V3* ng) {
  const V3 v0 = a - O, v1 = b - O, v2 = c - O;
  const V3 e0 = v2 - v0, e1 = v0 - v1, e2 = v1 - v2;
  const V3 Ng = stable_normalize(stable_triangle_normal(e0, e1, e2), compat);
  const float dd = dot(Ng, D);
  const float den = dd + dd;
  const float tt = dot(v0, Ng);
  const float T = tt + tt;
  const float t = T / den;
  if (!(0.0f <= t)) return false;
  if (!(den != 0.0f)) return false;
  *thit = t;
  *ng = Ng;
  return true;
}

// Full test (PlueckerTriangleIntersect): the edge test, then t and Ng.
FLG_HD bool pluecker(V3 a, V3 b, V3 c, V3 O, V3 D, bool compat, float* thit, V3* ng) {
  if (!pluecker_uvw(a - O, b - O, c - O, D)) return false;
  return pluecker_t(a, b, c, O, D, compat, thit, ng);
}

// OffsetRayOrigin (src/common/ray.h:64-93)
FLG_HD V3 offset_ray_origin(V3 p, V3 n, V3 dir) {
  const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
  V3 offset = n;
  if (dot(n, dir) <= 0) offset = -n;
  int oi0 = (int)(offset.x * int_scale), oi1 = (int)(offset.y * int_scale),
      oi2 = (int)(offset.z * int_scale);
  float px = u2f((uint32_t)((int32_t)f2u(p.x) + ((p.x < 0) ? -oi0 : oi0)));
  float py = u2f((uint32_t)((int32_t)f2u(p.y) + ((p.y < 0) ? -oi1 : oi1)));
  float pz = u2f((uint32_t)((int32_t)f2u(p.z) + ((p.z < 0) ? -oi2 : oi2)));
  return v3(fabsf(p.x) < origin ? p.x + float_scale * offset.x : px,
            fabsf(p.y) < origin ? p.y + float_scale * offset.y : py,
            fabsf(p.z) < origin ? p.z + float_scale * offset.z : pz);
}

// BoundIntersect (intersector.hpp:103-120) with a precomputed 1/d (the
// reference computes the same IEEE quotient per call).  The reference returns
// as soon as t0 > t1 after an axis; t0 only grows and t1 only shrinks from
// one axis to the next (each update keeps the old value unless the new one
// compares past it, NaN included), so one test after the three axes gives the
// same answer and the same t0, without a branch per axis.  t0 and t1 are never
// NaN, so `tn > t0 ? tn : t0` is fmaxf(t0, tn) (a NaN operand yields the
// other) up to the sign of a zero result, which no comparison sees; the swap
// keeps the reference's comparison, under which a NaN bound is not swapped.
FLG_HD bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz, V3 o, V3 inv,
                 float* tnear) {
  float t0 = 0.0f, t1 = 3.402823466e+38f;
  float tn = (lox - o.x) * inv.x, tf = (hix - o.x) * inv.x;
  if (tn > tf) { float t = tn; tn = tf; tf = t; }
  t0 = fmaxf(t0, tn);
  t1 = fminf(t1, tf);
  tn = (loy - o.y) * inv.y;
  tf = (hiy - o.y) * inv.y;
  if (tn > tf) { float t = tn; tn = tf; tf = t; }
  t0 = fmaxf(t0, tn);
  t1 = fminf(t1, tf);
  tn = (loz - o.z) * inv.z;
  tf = (hiz - o.z) * inv.z;
  if (tn > tf) { float t = tn; tn = tf; tf = t; }
  t0 = fmaxf(t0, tn);
  t1 = fminf(t1, tf);
  *tnear = t0;
  return !(t0 > t1);
}

}  // namespace flg
