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

// A leaf test in two parts, on a triangle already loaded.  tri_edges is the
// edge test (it does not depend on the closest hit so far); tri_finish
// computes t for a triangle that passed it and applies the closest-hit rule
// (t <= best replaces).  Running both parts back to back is the reference's
// test; a node's two leaves can run their edge tests first and finish in
// order, which gives the same result.
template <bool COUNT>
FLG_HD bool tri_edges(const BTri& t, V3 o, V3 d, Counters& c) {
  if (COUNT) ++c.tris;
  return pluecker_uvw(v3(t.v0.x, t.v0.y, t.v0.z) - o, v3(t.v1.x, t.v1.y, t.v1.z) - o,
                      v3(t.v2.x, t.v2.y, t.v2.z) - o, d);
}

template <bool COMPAT>
FLG_HD void tri_finish(const BTri& t, uint32_t p, V3 o, V3 d, float& best, int32_t& hit) {
  float th;
  V3 ng;
  if (!pluecker_t(v3(t.v0.x, t.v0.y, t.v0.z), v3(t.v1.x, t.v1.y, t.v1.z), v3(t.v2.x, t.v2.y, t.v2.z), o,
                  d, COMPAT, &th, &ng))
    return;
  if (th > best) return;
  best = th;
  hit = (int32_t)p;  // sorted position; the caller maps it to the original id
}

template <bool COMPAT, bool COUNT>
FLG_HD void test_tri(const TraceGeom& g, uint32_t p, V3 o, V3 d, float& best, int32_t& hit,
                    Counters& c) {
  const BTri t = g.tris[p];
  if (tri_edges<COUNT>(t, o, d, c)) tri_finish<COMPAT>(t, p, o, d, best, hit);
}

// Traversal stacks of (node, tnear) entries.  The depth of the tree is below
// 64 (the common-prefix length grows strictly from a node to its children and
// is below 64), so a 64-entry stack never overflows.

// All entries in private memory (scratch on the device).  The entries are an
// array owned by the caller, so that the stack pointer stays in a register
// (a counter stored beside a dynamically indexed array goes to scratch with it).
struct PrivStack {
  uint2* e;  // kStack entries: node, tnear bits
  int sp = 0;
  FLG_HD explicit PrivStack(uint2* entries) : e(entries) {}
  FLG_HD void reset() { sp = 0; }
  FLG_HD bool empty() const { return sp == 0; }
  FLG_HD void push(int32_t n, float tn) {
    if (sp < kStack) {
      e[sp] = make_uint2((uint32_t)n, f2u(tn));
      ++sp;
    }
  }
  FLG_HD void pop(int32_t* n, float* tn) {
    --sp;
    const uint2 x = e[sp];
    *n = (int32_t)x.x;
    *tn = u2f(x.y);
  }
};

// The K most recent entries in a fast per-lane ring (LDS on the device; entry
// k of this lane at fast[k * stride]), older ones moved to a private spill
// array (kStack - K entries, owned by the caller) when the ring is full.  When
// the ring runs empty, the most recent spilled entries are copied back into it
// before the pop, so pops always read the ring (no load that could come from
// either memory) and the order is exactly that of PrivStack.
template <int K>
struct RingStack {
  uint2* fast;
  int stride;
  uint2* spill;
  int nfast = 0, top = 0, nspill = 0;  // top: next ring slot to write
  FLG_HD RingStack(uint2* f, int s, uint2* sp) : fast(f), stride(s), spill(sp) {}
  FLG_HD void reset() { nfast = top = nspill = 0; }
  FLG_HD bool empty() const { return nfast == 0 && nspill == 0; }
  FLG_HD void push(int32_t n, float tn) {
    if (nfast == K) {  // full: the oldest ring entry sits in the slot about to be written
      if (nspill >= kStack - K) return;  // unreachable (depth < 64), as PrivStack
      spill[nspill] = fast[top * stride];
      ++nspill;
      --nfast;
    }
    fast[top * stride] = make_uint2((uint32_t)n, f2u(tn));
    top = top + 1 == K ? 0 : top + 1;
    ++nfast;
  }
  FLG_HD void pop(int32_t* n, float* tn) {
    if (nfast == 0) {  // refill from the spill array, oldest first
      const int m = nspill < K ? nspill : K;
      top = 0;
      for (int j = nspill - m; j < nspill; ++j) {
        fast[top * stride] = spill[j];
        ++top;
      }
      if (top == K) top = 0;
      nfast = m;
      nspill -= m;
    }
    top = top == 0 ? K - 1 : top - 1;
    const uint2 x = fast[top * stride];
    --nfast;
    *n = (int32_t)x.x;
    *tn = u2f(x.y);
  }
};

// Resumable traversal of one ray: trace_begin tests the root (and a
// one-triangle tree) and returns false when nothing is left to visit;
// trace_visit visits one node and returns false when the traversal has
// ended.  begin + visits until false is exactly the recursive order of the
// reference; persistent kernels interleave visits of different rays with it.
struct TraceRay {
  V3 o, d, inv;
  float best;
  int32_t hit;   // sorted position of the closest hit so far, -1 if none
  int32_t node;  // next node to visit
  int64_t visits;
};

template <bool COMPAT, bool COUNT>
FLG_HD bool trace_begin(const TraceGeom& g, V3 o, V3 d, float tmax, TraceRay& r, Counters& c) {
  r.o = o;
  r.d = d;
  r.best = tmax;
  r.hit = -1;
  r.node = 0;
  r.visits = 0;
  if (g.n <= 0) return false;
  r.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  float tn;
  if (!slab(g.root[0], g.root[1], g.root[2], g.root[3], g.root[4], g.root[5], o, r.inv, &tn)) return false;
  if (r.best < tn) return false;
  if (g.n == 1) {
    test_tri<COMPAT, COUNT>(g, 0, o, d, r.best, r.hit, c);
    return false;
  }
  return true;
}

// PAIR chooses how a node's leaf children are fetched; both give the same
// tests in the same order.  With PAIR the two triangles are loaded together
// (two loads in flight instead of one after the other), the inner children's
// slab tests run while they arrive, and only the comparison with the closest
// hit waits for the leaf tests, which may lower it: this holds both triangles
// and both boxes in registers at once, which pays in the mesh kernels (80
// registers: 233 against 240 ms on the combined scene) and spills in the
// kernels held to 64 registers (12.4 against 10.0 ms on the path tracer's
// frame).  Without PAIR one triangle is in registers at a time, loaded again
// for the t computation of a candidate.
template <bool COMPAT, bool COUNT, bool PAIR = false, class Stack>
FLG_HD bool trace_visit(const TraceGeom& g, TraceRay& r, Stack& st, Counters& c) {
  // every node is visited at most once; the bound only guards a corrupt tree
  if (r.visits >= (int64_t)g.n) return false;
  ++r.visits;
  if (COUNT) ++c.nodes;
  const V3 o = r.o, d = r.d, inv = r.inv;
  const BNode nd = g.nodes[r.node];
  const int32_t refL = nd.d.x, refR = nd.d.y;
  const uint32_t pL = (uint32_t)refL & 0x7fffffffu, pR = (uint32_t)refR & 0x7fffffffu;
  // leaf children, left then right: both edge tests, then t for the
  // candidates in that order (a wave runs the costly part once unless some
  // lane has two candidates); then the inner children's boxes against the
  // closest hit so far
  float tL = 0.0f, tR = 0.0f;
  bool goL, goR;
  if constexpr (PAIR) {
    BTri tl, tr;
    if (refL < 0) tl = g.tris[pL];
    if (refR < 0) tr = g.tris[pR];
    const bool boxL = refL >= 0 && slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, &tL);
    const bool boxR = refR >= 0 && slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, &tR);
    const bool candL = refL < 0 && tri_edges<COUNT>(tl, o, d, c);
    const bool candR = refR < 0 && tri_edges<COUNT>(tr, o, d, c);
    if (candL || candR) tri_finish<COMPAT>(candL ? tl : tr, candL ? pL : pR, o, d, r.best, r.hit);
    if (candL && candR) tri_finish<COMPAT>(tr, pR, o, d, r.best, r.hit);
    goL = boxL && !(r.best < tL);
    goR = boxR && !(r.best < tR);
  } else {
    const bool candL = refL < 0 && tri_edges<COUNT>(g.tris[pL], o, d, c);
    const bool candR = refR < 0 && tri_edges<COUNT>(g.tris[pR], o, d, c);
    if (candL || candR) {
      const uint32_t p = candL ? pL : pR;
      tri_finish<COMPAT>(g.tris[p], p, o, d, r.best, r.hit);
    }
    if (candL && candR) tri_finish<COMPAT>(g.tris[pR], pR, o, d, r.best, r.hit);
    goL = refL >= 0 && slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, &tL) && !(r.best < tL);
    goR = refR >= 0 && slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, &tR) && !(r.best < tR);
  }
  if (goL) {
    if (goR) st.push(refR, tR);
    r.node = refL;
    return true;
  }
  if (goR) {
    r.node = refR;
    return true;
  }
  while (!st.empty()) {
    int32_t n;
    float t;
    st.pop(&n, &t);
    if (!(r.best < t)) {
      r.node = n;
      return true;
    }
  }
  return false;
}

// Returns the sorted position of the hit (-1 on a miss); best holds t.
template <bool COMPAT, bool COUNT, bool PAIR = false, class Stack>
FLG_HD int32_t trace_one_s(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c, Stack& st) {
  TraceRay r;
  bool more = trace_begin<COMPAT, COUNT>(g, o, d, best, r, c);
  while (more) more = trace_visit<COMPAT, COUNT, PAIR>(g, r, st, c);
  best = r.best;
  return r.hit;
}

template <bool COMPAT, bool COUNT>
FLG_HD int32_t trace_one(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c) {
  uint2 entries[kStack];
  PrivStack st(entries);
  return trace_one_s<COMPAT, COUNT>(g, o, d, best, c, st);
}

// Traversal with the stack policy of RING: 0 keeps the whole stack in private
// memory, K > 0 keeps the K most recent entries of each lane in LDS; PAIR as
// in trace_visit.
template <bool COMPAT, bool COUNT, int BS, int RING, bool PAIR = false>
FLG_D int32_t trace_lane(const TraceGeom& g, V3 o, V3 d, float& best, Counters& c, uint2* ring) {
  if constexpr (RING > 0) {
    uint2 spill[kStack - RING];
    RingStack<RING> st(ring + threadIdx.x, BS, spill);
    return trace_one_s<COMPAT, COUNT, PAIR>(g, o, d, best, c, st);
  } else {
    (void)ring;
    return trace_one<COMPAT, COUNT>(g, o, d, best, c);
  }
}

}  // namespace flg
