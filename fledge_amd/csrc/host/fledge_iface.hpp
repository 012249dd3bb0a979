// The subset of fledge's C++ interfaces that the GPU drop-ins implement, with
// the reference's names and signatures:
//   Ray, AABB                  src/common/ray.h:98-132, src/common/aabb.h:11-68
//   Interaction, S/VInteraction src/interaction.hpp:17-81
//   Primitive, Accel           src/primitive.hpp:26-72, src/accel.hpp:12-21
//   Volume                     src/volume.hpp:16-30
//   Sampler                    src/common/sampler.h:17-36
//   Integrator                 src/integrator.hpp:10-14
//   InternalTriangleMesh, BVHBuilderBase  src/experimental/base_bvh.hpp:26-125
//   TriangleMesh               src/shape.hpp:53-62
//   RenderBase                 src/render.hpp:17-26
//   EIntegratorType, EBackendType  src/fledge.h:87-96
// Inside the fledge tree these come from fledge's own headers: define
// FLEDGE_HAS_REFERENCE_HEADERS and include them before fledge_gpu.hpp.  This
// file lets the drop-ins and their tests build where fledge's dependencies
// (fmt, TBB, Embree, OpenVDB) are absent.
#pragma once
#ifndef FLEDGE_HAS_REFERENCE_HEADERS

#include <cmath>
#include <cstdint>
#include <limits>
#include <memory_resource>
#include <string>
#include <vector>

namespace fledge {

using Float = float;
constexpr Float INF = std::numeric_limits<Float>::infinity();

struct Vector3f {
  Float m_vec[3];
  Vector3f() : m_vec{0, 0, 0} {}
  explicit Vector3f(Float v) : m_vec{v, v, v} {}
  Vector3f(Float x, Float y, Float z) : m_vec{x, y, z} {}
  Float& operator[](int i) { return m_vec[i]; }
  const Float& operator[](int i) const { return m_vec[i]; }
  Float& x() { return m_vec[0]; }
  Float& y() { return m_vec[1]; }
  Float& z() { return m_vec[2]; }
  const Float& x() const { return m_vec[0]; }
  const Float& y() const { return m_vec[1]; }
  const Float& z() const { return m_vec[2]; }
  Vector3f operator-() const { return {-m_vec[0], -m_vec[1], -m_vec[2]}; }
  Vector3f operator+(const Vector3f& r) const { return {m_vec[0] + r[0], m_vec[1] + r[1], m_vec[2] + r[2]}; }
  Vector3f operator*(Float s) const { return {m_vec[0] * s, m_vec[1] * s, m_vec[2] * s}; }
};
struct Vector2f {
  Float m_vec[2];
};

class Volume;
class MaterialDispatcher;
class AreaLight;
class Primitive;

struct Ray {
  Ray() : m_tMax(INF), m_volume(nullptr) {}
  Ray(const Vector3f& o, const Vector3f& d, Float tMax = INF, Volume const* volume = nullptr)
      : m_o(o), m_d(d), m_tMax(tMax), m_volume(volume) {}
  Vector3f at(Float t) const { return m_o + m_d * t; }
  Vector3f operator()(Float t) const { return at(t); }
  Vector3f m_o, m_d;
  mutable Float m_tMax;
  mutable Volume const* m_volume{nullptr};
};

class AABB {
 public:
  AABB() = default;
  AABB(Vector3f t_min, Vector3f t_max) : m_min(t_min), m_max(t_max) {}
  Vector3f m_min{Vector3f(0.0f)}, m_max{Vector3f(0.0f)};
};

class Interaction {
 public:
  virtual ~Interaction() = default;
  Vector3f m_p, m_ns{Vector3f(0.0f)}, m_ng{Vector3f(0.0f)}, m_wo;
  Ray m_ray;
};
class SInteraction : public Interaction {
 public:
  const Primitive* m_primitive{nullptr};
};
class VInteraction : public Interaction {
 public:
  VInteraction() = default;
  VInteraction(const Vector3f& p, const Vector3f& wo, Float g) : m_g(g) {
    m_p = p;
    m_wo = wo;
  }
  Float m_g{0};
};

struct TriangleMesh {
  int nInd, nVert;
  int* ind;
  Vector3f *p, *n;
  Vector2f* uv;
};

class Primitive {
 public:
  virtual ~Primitive() = default;
  virtual bool intersect(const Ray& ray, SInteraction& isect) const = 0;
  virtual AABB getBound() const = 0;
  virtual MaterialDispatcher* getMaterial() const = 0;
  virtual AreaLight* getAreaLight() const { return nullptr; }
  virtual Volume* getVolume() const { return nullptr; }
};

class Accel : public Primitive {
 public:
  ~Accel() override = default;
  bool intersect(const Ray& ray, SInteraction& isect) const override = 0;
  AreaLight* getAreaLight() const override { return nullptr; }
  MaterialDispatcher* getMaterial() const override { return nullptr; }
};

class Sampler {
 public:
  virtual ~Sampler() = default;
  virtual Float get1D() = 0;
};

class Volume {
 public:
  virtual ~Volume() = default;
  virtual Vector3f tr(const Ray& ray, Sampler& rng) const = 0;
  virtual Vector3f sample(const Ray& ray, Sampler& rng, VInteraction& vi, bool& success) const = 0;
  virtual AABB getBound() const { return m_aabb; }
  virtual void setBound(const AABB& aabb) { m_aabb = aabb; }
  AABB m_aabb;
};

// EFilmBufferType src/fledge.h:76-85; Film::getBuffer src/film.hpp:92-120
enum class EFilmBufferType {
  EColor = (1),
  EAlbedo = (1 << 1),
  ENormal = (1 << 2),
  EOutput = (1 << 3),
};

struct Film {
  Film(int resX, int resY) : m_resX(resX), m_resY(resY), m_color(size_t(resX) * resY) {}
  int getBufferIdx(int x, int y) const { return y * m_resX + x; }
  Vector3f& getBuffer(int x, int y, EFilmBufferType) { return m_color[getBufferIdx(x, y)]; }
  // Film::saveBuffer (film.cpp:13-49) for ".exr" names, through
  // flg_film_save_exr; other formats need OpenImageIO and return false.
  bool saveBuffer(const std::string& name, EFilmBufferType) const;
  int m_resX, m_resY;
  std::vector<Vector3f> m_color;
};

enum class EIntegratorType {
  EPathIntegrator = 0,
  EVolPathIntegrator,
};

enum class EBackendType {
  ECPUBackend = (1),
  EOptiXBackend = (1 << 1),
  EOptiXWaveFrontBackend = (1 << 2),
};

// The scene-description fields of Scene (src/scene.hpp:85-101) that the
// integrators and the render read.
class Scene {
 public:
  int m_resX{0}, m_resY{0}, m_SPP{0}, m_maxDepth{0};
  Float m_FoV{0};
  Vector3f m_up{0}, m_origin{0}, m_target{0};
  EIntegratorType m_integrator_type{EIntegratorType::EPathIntegrator};
  std::vector<Primitive*> m_primitives{};
  Volume* m_volume{nullptr};
  Film* m_film{nullptr};
};

class RenderBase {
 public:
  RenderBase() = default;
  virtual ~RenderBase() = default;
  virtual void init() = 0;
  virtual bool preProcess() = 0;
  virtual bool saveImage(const std::string& name, bool denoise = false) = 0;
  virtual bool render() = 0;
  virtual EBackendType getBackends() = 0;
};

class Integrator {
 public:
  virtual ~Integrator() = default;
  virtual void render(const Scene& scene) = 0;
};

namespace experimental {
struct InternalTriangleMesh {
  int nInd, nVert;
  int* ind;          // nInd = 3 * triangles
  Vector3f *p, *n;   // nVert
  Vector2f* uv;
};

class BVHBuilderBase {
 public:
  struct alignas(16) BVHRayHit {
    Vector3f ray_o{0};
    Vector3f ray_d{0};
    float tnear{0};
    float tfar{std::numeric_limits<float>::max()};
    Vector3f hit_ng{0};
    Vector3f hit_ns{0};
    bool hit{false};
  };
  struct alignas(16) BVHBound {
    Vector3f lower{std::numeric_limits<Float>::max()}, upper{-std::numeric_limits<Float>::max()};
  };
  BVHBuilderBase(InternalTriangleMesh* mesh, std::pmr::memory_resource* mem_resource)
      : m_mesh(mesh), m_has_normal(mesh->n != nullptr), m_mem_resource(mem_resource) {}
  virtual ~BVHBuilderBase() = default;
  virtual void build() = 0;
  virtual BVHBound getBound() const = 0;
  virtual bool intersect(BVHRayHit& rayhit) = 0;
  virtual std::pmr::memory_resource* getMemoryResource() { return m_mem_resource; }

 protected:
  InternalTriangleMesh* m_mesh;
  bool m_has_normal;
  std::pmr::memory_resource* m_mem_resource{std::pmr::get_default_resource()};
};
}  // namespace experimental

}  // namespace fledge

#endif  // FLEDGE_HAS_REFERENCE_HEADERS
