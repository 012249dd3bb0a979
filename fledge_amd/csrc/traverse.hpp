// Closest-hit traversal shared by the GPU kernels and the host mirror of the
// C++ drop-ins (RadixBVHBuilder::recursiveIntersect order and tie rule,
// src/experimental/radix_bvh.cpp:202-245; see trace.hip).
#pragma once
#include "bvh.hpp"

namespace flg {

constexpr int kStack = 64;  // > tree depth: delta strictly grows per level and is < 64

struct TraceGeom {
  const BNode* nodes;
  const BTri* tris;
  int32_t n;
  float root[6];
};

inline TraceGeom geom_of(const flg_bvh* b) {
  TraceGeom g;
  g.nodes = b->nodes;
  g.tris = b->tris;
  g.n = b->n;
  for (int k = 0; k < 6; ++k) g.root[k] = b->bound[k];
  return g;
}

struct Counters {
  uint32_t nodes, tris;
};

template <bool COMPAT, bool COUNT>
FLG_HD void test_tri(const TraceGeom& g, uint32_t p, V3 o, V3 d, float& best, int32_t& hit,
                    Counters& c) {
  const BTri t = g.tris[p];
  if (COUNT) ++c.tris;
  float th;
  V3 ng;
  if (!pluecker(v3(t.v0.x, t.v0.y, t.v0.z), v3(t.v1.x, t.v1.y, t.v1.z), v3(t.v2.x, t.v2.y, t.v2.z),
                o, d, COMPAT, &th, &ng))
    return;
  if (th > best) return;
  best = th;
  hit = (int32_t)p;  // sorted position; the caller maps it to the original id
}

// Returns the sorted position of the hit (-1 on a miss); best holds t.
template <bool COMPAT, bool COUNT>
FLG_HD int32_t trace_one(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c) {
  int32_t hit = -1;
  if (g.n <= 0) return -1;
  const V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  float tn;
  if (!slab(g.root[0], g.root[1], g.root[2], g.root[3], g.root[4], g.root[5], o, inv, &tn)) return -1;
  if (best < tn) return -1;
  if (g.n == 1) {
    test_tri<COMPAT, COUNT>(g, 0, o, d, best, hit, c);
    return hit;
  }
  int32_t stack[kStack];
  float stackt[kStack];
  int sp = 0;
  int32_t node = 0;
  // every node is visited at most once; the bound only guards a corrupt tree
  for (int64_t visits = 0; visits < (int64_t)g.n; ++visits) {
    if (COUNT) ++c.nodes;
    const BNode nd = g.nodes[node];
    const int32_t refL = nd.d.x, refR = nd.d.y;
    if (refL < 0) test_tri<COMPAT, COUNT>(g, (uint32_t)refL & 0x7fffffffu, o, d, best, hit, c);
    if (refR < 0) test_tri<COMPAT, COUNT>(g, (uint32_t)refR & 0x7fffffffu, o, d, best, hit, c);
    float tL = 0.0f, tR = 0.0f;
    bool goL = refL >= 0 && slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, &tL) &&
               !(best < tL);
    bool goR = refR >= 0 && slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, &tR) &&
               !(best < tR);
    if (goL) {
      if (goR && sp < kStack) {
        stack[sp] = refR;
        stackt[sp] = tR;
        ++sp;
      }
      node = refL;
      continue;
    }
    if (goR) {
      node = refR;
      continue;
    }
    bool found = false;
    while (sp > 0) {
      --sp;
      if (!(best < stackt[sp])) {
        node = stack[sp];
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  return hit;
}

}  // namespace flg
