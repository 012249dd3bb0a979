// TEST INFRASTRUCTURE — CPU restatement of the scene's accel over meshes:
// NaiveBVHAccel's ctor and intersect (src/accel.cpp:35-133) over
// MeshPrimitives (src/primitive.cpp:54-90) whose accel is each mesh's LBVH
// (the oracle's build and traversal, lbvh.cpp), as Scene::init builds it
// (src/scene.cpp:46) and Scene::intersect asks it (src/scene.cpp:57).
#include <algorithm>
#include <memory>
#include <thread>
#include <vector>

#include "oracle_impl.hpp"

namespace orc {

namespace {

struct Box {
  V3 lo{0, 0, 0}, hi{0, 0, 0};  // AABB's default: the point at the origin (aabb.h:94)
};

Box merge(const Box& a, V3 lo, V3 hi) {
  Box r;
  r.lo = v3(std::min(a.lo.x, lo.x), std::min(a.lo.y, lo.y), std::min(a.lo.z, lo.z));
  r.hi = v3(std::max(a.hi.x, hi.x), std::max(a.hi.y, hi.y), std::max(a.hi.z, hi.z));
  return r;
}

struct Node {
  Box box;
  std::unique_ptr<Node> left, right;
  std::vector<int> prims;  // a leaf's primitives, in order
};

struct SceneTree {
  std::vector<const Lbvh*> meshes;
  std::unique_ptr<Node> root;
};

V3 lo_of(const Lbvh& b) { return b.glo; }
V3 hi_of(const Lbvh& b) { return b.ghi; }

// NaiveBVHAccel's ctor (accel.cpp:35-86)
std::unique_ptr<Node> build(const SceneTree& sc, std::vector<int> p, int depth) {
  auto n = std::make_unique<Node>();
  for (int m : p) n->box = merge(n->box, lo_of(*sc.meshes[m]), hi_of(*sc.meshes[m]));
  if (p.size() <= 64 || depth > 20) {
    n->prims = p;
    return n;
  }
  const int dim = depth % 3;
  auto centre = [&](int m) {
    const V3 lo = lo_of(*sc.meshes[m]), hi = hi_of(*sc.meshes[m]);
    const float c[3] = {(lo.x + hi.x) / 2, (lo.y + hi.y) / 2, (lo.z + hi.z) / 2};
    return c[dim];
  };
  std::sort(p.begin(), p.end(), [&](int a, int b) { return centre(a) < centre(b); });
  const size_t mid = p.size() / 2;
  n->left = build(sc, std::vector<int>(p.begin(), p.begin() + mid), depth + 1);
  n->right = build(sc, std::vector<int>(p.begin() + mid, p.end()), depth + 1);
  return n;
}

// AABB::intersect_pbrt (aabb.h:38-59) against [0, tmax]
bool intersect_pbrt(const Box& b, V3 o, V3 d, float tmax, float* t_min) {
  float t0 = 0, t1 = tmax;
  const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
  const float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
  for (int i = 0; i < 3; ++i) {
    const float inv = 1 / ds[i];
    float tn = (lo[i] - os[i]) * inv, tf = (hi[i] - os[i]) * inv;
    if (tn > tf) std::swap(tn, tf);
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  *t_min = t0;
  return true;
}

struct Hit {
  int mesh = -1, id = -1;
  float t = 0;
};

// NaiveBVHAccel::intersect (accel.cpp:92-130); tmax is the ray's m_tMax
bool intersect(const SceneTree& sc, const Node& n, bool compat, V3 o, V3 d, float& tmax, Hit& isect) {
  float t_min;
  if (!intersect_pbrt(n.box, o, d, tmax, &t_min)) return false;
  if (tmax < t_min) return false;
  if (!n.prims.empty()) {
    const float o_tmax = tmax;
    float best = o_tmax;
    for (int m : n.prims) {
      float t = 0;
      V3 ng;
      // MeshPrimitive::intersect: the mesh's accel, which shortens m_tMax on a hit
      const int32_t id = lbvh_trace(*sc.meshes[m], o, d, tmax, compat, &t, &ng, nullptr);
      if (id >= 0) {
        tmax = t;
        if (tmax < best) {
          isect.mesh = m;
          isect.id = id;
          isect.t = t;
          best = tmax;
        }
      }
    }
    return best != o_tmax;
  }
  bool l = false, r = false;
  if (n.left) l = intersect(sc, *n.left, compat, o, d, tmax, isect);
  if (n.right) r = intersect(sc, *n.right, compat, o, d, tmax, isect);
  return l || r;
}

}  // namespace

void scene_trace(const std::vector<const Lbvh*>& meshes, bool compat, const float* rays, int64_t n, int32_t* mesh,
                 int32_t* ids, float* ts, int nthreads) {
  SceneTree sc;
  sc.meshes = meshes;
  std::vector<int> p(meshes.size());
  for (size_t i = 0; i < p.size(); ++i) p[i] = (int)i;
  sc.root = build(sc, p, 0);
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      for (int64_t i = t; i < n; i += nthreads) {
        const float* r = rays + 8 * i;
        float tmax = r[3];
        Hit h;
        const bool hit = intersect(sc, *sc.root, compat, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), tmax, h);
        mesh[i] = hit ? h.mesh : -1;
        ids[i] = hit ? h.id : -1;
        ts[i] = hit ? h.t : r[3];
      }
    });
  for (auto& x : th) x.join();
}

}  // namespace orc
