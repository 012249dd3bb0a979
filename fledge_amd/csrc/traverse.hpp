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

// Traversal stacks of (node, tnear) entries.  The depth of the tree is below
// 64 (the common-prefix length grows strictly from a node to its children and
// is below 64), so a 64-entry stack never overflows.

// All entries in private memory (scratch on the device).
struct PrivStack {
  int32_t node[kStack];
  float t[kStack];
  int sp = 0;
  FLG_HD bool empty() const { return sp == 0; }
  FLG_HD void push(int32_t n, float tn) {
    if (sp < kStack) {
      node[sp] = n;
      t[sp] = tn;
      ++sp;
    }
  }
  FLG_HD void pop(int32_t* n, float* tn) {
    --sp;
    *n = node[sp];
    *tn = t[sp];
  }
};

// The K most recent entries in a fast per-lane ring (LDS on the device; entry
// k of this lane at fast[k * stride]), older ones moved to a private spill
// stack when the ring is full.  Pops take the ring first, then the spill
// stack, so the order is exactly that of PrivStack.
template <int K>
struct RingStack {
  uint2* fast;
  int stride;
  int nfast = 0, top = 0, nspill = 0;  // top: next ring slot to write
  int32_t node[kStack - K];
  float t[kStack - K];
  FLG_HD RingStack(uint2* f, int s) : fast(f), stride(s) {}
  FLG_HD bool empty() const { return nfast == 0 && nspill == 0; }
  FLG_HD void push(int32_t n, float tn) {
    if (nfast == K) {  // full: the oldest ring entry sits in the slot about to be written
      if (nspill >= kStack - K) return;  // unreachable (depth < 64), as PrivStack
      const uint2 e = fast[top * stride];
      node[nspill] = (int32_t)e.x;
      t[nspill] = u2f(e.y);
      ++nspill;
      --nfast;
    }
    fast[top * stride] = make_uint2((uint32_t)n, f2u(tn));
    top = top + 1 == K ? 0 : top + 1;
    ++nfast;
  }
  FLG_HD void pop(int32_t* n, float* tn) {
    if (nfast > 0) {
      top = top == 0 ? K - 1 : top - 1;
      const uint2 e = fast[top * stride];
      *n = (int32_t)e.x;
      *tn = u2f(e.y);
      --nfast;
    } else {
      --nspill;
      *n = node[nspill];
      *tn = t[nspill];
    }
  }
};

// Returns the sorted position of the hit (-1 on a miss); best holds t.
template <bool COMPAT, bool COUNT, class Stack>
FLG_HD int32_t trace_one_s(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c, Stack& st) {
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
      if (goR) st.push(refR, tR);
      node = refL;
      continue;
    }
    if (goR) {
      node = refR;
      continue;
    }
    bool found = false;
    while (!st.empty()) {
      int32_t n;
      float t;
      st.pop(&n, &t);
      if (!(best < t)) {
        node = n;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  return hit;
}

template <bool COMPAT, bool COUNT>
FLG_HD int32_t trace_one(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c) {
  PrivStack st;
  return trace_one_s<COMPAT, COUNT>(g, o, d, best, c, st);
}

}  // namespace flg
