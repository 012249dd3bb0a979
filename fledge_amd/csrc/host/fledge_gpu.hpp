// C++ drop-ins that put fledge's ray-tracing hot path on an MI355X behind the
// reference's own interfaces:
//   LBVHAccel : Accel                      for src/accel.cpp (NaiveBVHAccel)
//   experimental::LBVHBuilder : BVHBuilderBase  for src/experimental/radix_bvh.cpp
//   GridVolume : Volume                    for src/volume.cpp (OpenVDBVolume)
//   GPUVolIntegrator : Integrator          for src/integrator.cpp (SVolIntegrator)
//   gpu_memory_resource, managed_memory_resource, async_memory_resource
//                                          for src/gpu/gpu_memory.{hpp,cpp}
// Every GPU call goes through the C ABI of libfledge_hip.so (fledge_hip.h).
// A non-zero status becomes a gpu::Error exception carrying flg_last_error().
#pragma once
#include <cstdint>
#include <memory>
#include <memory_resource>
#include <stdexcept>
#include <string>
#include <vector>

#include "fledge_hip.h"
#include "fledge_iface.hpp"

namespace fledge {

// ------------------------------------------------- src/gpu memory resources
// std::pmr resources over HIP allocations (gpu_memory.hpp:10-40,
// gpu_memory.cpp:24-79).  Allocation failures throw std::bad_alloc.
class managed_memory_resource : public std::pmr::memory_resource {
 public:
  ~managed_memory_resource() override = default;
  void* do_allocate(std::size_t bytes, std::size_t alignment) override;
  void do_deallocate(void* p, std::size_t bytes, std::size_t alignment) override;
  bool do_is_equal(const std::pmr::memory_resource& other) const noexcept override;
};

class gpu_memory_resource : public std::pmr::memory_resource {
 public:
  ~gpu_memory_resource() override = default;
  void* do_allocate(std::size_t bytes, std::size_t alignment) override;
  void do_deallocate(void* p, std::size_t bytes, std::size_t alignment) override;
  bool do_is_equal(const std::pmr::memory_resource& other) const noexcept override;
};

// Stream-ordered allocations from the device's default memory pool
// (hipMallocAsync / hipFreeAsync on the given stream; NULL = default stream).
class async_memory_resource : public std::pmr::memory_resource {
 public:
  explicit async_memory_resource(void* stream = nullptr) : m_stream(stream) {}
  ~async_memory_resource() override = default;
  void* do_allocate(std::size_t bytes, std::size_t alignment) override;
  void do_deallocate(void* p, std::size_t bytes, std::size_t alignment) override;
  bool do_is_equal(const std::pmr::memory_resource& other) const noexcept override;

 private:
  void* m_stream;
};

std::pmr::memory_resource* GlobalManagedMemoryResource();

namespace gpu {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), m_status(status) {}
  int status() const { return m_status; }

 private:
  int m_status;
};

// One flg_ctx: a device and the stream every call of this context runs on.
class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  flg_ctx* get() const { return m_ctx; }
  int device() const { return m_device; }
  void check(int status, const char* what) const;
  void synchronize() const;
  // The context that drop-ins constructed with the reference's own
  // signatures use: device 0, default stream, created on first use.
  static Context& global();

 private:
  flg_ctx* m_ctx = nullptr;
  int m_device = 0;
};

// Device buffer owned through a context (flg_malloc / flg_free).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const Context& ctx, size_t bytes);
  DeviceBuffer(const Context& ctx, const void* host, size_t bytes);
  ~DeviceBuffer();
  DeviceBuffer(DeviceBuffer&& o) noexcept;
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;

  void* data() const { return m_ptr; }
  size_t size() const { return m_bytes; }
  void upload(const void* host, size_t bytes);
  void download(void* host, size_t bytes) const;

 private:
  const Context* m_ctx = nullptr;
  void* m_ptr = nullptr;
  size_t m_bytes = 0;
};

// Host copy of a build's traversal image (flg_lbvh_download), traversed on the
// CPU by the same code the GPU kernels run.
class HostMirror {
 public:
  HostMirror() = default;
  void load(const Context& ctx, flg_bvh* bvh);
  // Closest hit with t <= tmax.  Returns the original triangle id or -1;
  // t and the geometric normal (the Pluecker test's) are written on a hit.
  int32_t intersect(const Vector3f& o, const Vector3f& d, float tmax, float* t, Vector3f* ng) const;
  // The Pluecker test's normal of triangle tri_id for this ray.
  Vector3f normal(int32_t tri_id, const Vector3f& o, const Vector3f& d) const;
  void setCompat(bool compat);
  int32_t ntris() const { return m_n; }

 private:
  bool m_compat = false;
  int32_t m_n = 0;
  float m_root[6] = {0, 0, 0, 0, 0, 0};
  std::vector<uint8_t> m_nodes;  // 64 B per internal node
  std::vector<uint8_t> m_tris;   // 48 B per triangle, sorted order
  std::vector<int32_t> m_by_id;  // original id -> sorted position
};

}  // namespace gpu

// ------------------------------------------------------------- Accel ------
// Closest-hit LBVH built on the GPU.  intersect() is the reference's
// single-ray contract (src/primitive.hpp:31-42: on a hit it shortens
// ray.m_tMax and fills isect, otherwise touches neither); it runs on a host
// mirror of the device tree, is const and re-entrant.  intersectBatch() traces
// on the GPU.
class LBVHAccel : public Accel {
 public:
  LBVHAccel(const std::vector<Vector3f>& p, const std::vector<int>& ind,
            gpu::Context& ctx = gpu::Context::global(), uint32_t flags = FLG_NORMAL_FIXED);
  LBVHAccel(const experimental::InternalTriangleMesh& mesh,
            gpu::Context& ctx = gpu::Context::global(), uint32_t flags = FLG_NORMAL_FIXED);
  // A mesh primitive's accel (MeshPrimitive ctor, src/primitive.cpp:54-68).
  LBVHAccel(const TriangleMesh& mesh, gpu::Context& ctx = gpu::Context::global(),
            uint32_t flags = FLG_NORMAL_FIXED);
  ~LBVHAccel() override;

  bool intersect(const Ray& ray, SInteraction& isect) const override;
  AABB getBound() const override;

  // Same as intersect(); also reports the original triangle index.
  bool intersect(const Ray& ray, SInteraction& isect, int32_t* tri_id) const;
  // Device rays in, device hits out, on the context's stream (asynchronous).
  void intersectBatch(const flg_ray* d_rays, flg_hit* d_hits, int64_t n) const;
  // Host convenience: uploads, traces, downloads.
  std::vector<flg_hit> intersectBatch(const std::vector<Ray>& rays) const;

  flg_bvh* handle() const { return m_bvh; }
  gpu::Context& context() const { return *m_ctx; }

 private:
  void build(const float* verts, int32_t nverts, const int32_t* ind, int32_t ntris, uint32_t flags);
  gpu::Context* m_ctx;
  flg_bvh* m_bvh = nullptr;
  gpu::HostMirror m_mirror;
  AABB m_bound;
};

// MakeTriangleMesh (src/plymesh.cpp:19-86) through flg_ply_load.  The
// reference allocates the arrays from its Resource; here the PlyMesh owns
// them and `mesh` points into it.  Throws gpu::Error.
struct PlyMesh {
  TriangleMesh mesh{0, 0, nullptr, nullptr, nullptr, nullptr};
  std::vector<Vector3f> p, n;
  std::vector<Vector2f> uv;
  std::vector<int> ind;
};
std::unique_ptr<PlyMesh> MakeTriangleMesh(const std::string& path);

// Scene::init's accel over the scene's meshes (src/scene.cpp:46: NaiveBVHAccel
// over its MeshPrimitives, src/accel.cpp:35-133), each mesh's accel an
// LBVHAccel.  intersect() is the single-ray contract on the host mirrors:
// the closest of the meshes' hits, ties between meshes to the one that
// NaiveBVHAccel asks first.  intersectBatch() traces on the GPU
// (flg_scene_trace_closest), same answer.  The meshes must outlive it.
class SceneAccel : public Accel {
 public:
  explicit SceneAccel(const std::vector<const LBVHAccel*>& meshes,
                      gpu::Context& ctx = gpu::Context::global());
  ~SceneAccel() override;

  bool intersect(const Ray& ray, SInteraction& isect) const override;
  // Same as intersect(); also reports the mesh (scene index) and the original
  // triangle index within it.
  bool intersect(const Ray& ray, SInteraction& isect, int32_t* mesh, int32_t* tri_id) const;
  AABB getBound() const override;

  void intersectBatch(const flg_ray* d_rays, flg_scene_hit* d_hits, int64_t n) const;
  std::vector<flg_scene_hit> intersectBatch(const std::vector<Ray>& rays) const;

 private:
  gpu::Context* m_ctx;
  std::vector<const LBVHAccel*> m_meshes;
  std::vector<int32_t> m_by_rank;  // scene indices in the order NaiveBVHAccel asks them
  flg_scene_accel* m_accel = nullptr;
  AABB m_box;                      // NaiveBVHAccel's root box (holds the origin)
};

namespace experimental {

// BVHBuilderBase over the GPU build (src/experimental/base_bvh.hpp:105-118),
// so bvh_test-style harnesses run unchanged.  The pmr resource is kept for
// getMemoryResource(); device memory comes from the context.
class LBVHBuilder : public BVHBuilderBase {
 public:
  LBVHBuilder(InternalTriangleMesh* mesh, std::pmr::memory_resource* mem_resource,
              gpu::Context* ctx = nullptr, uint32_t flags = FLG_NORMAL_FIXED);
  ~LBVHBuilder() override;

  void build() override;
  BVHBound getBound() const override;
  bool intersect(BVHRayHit& rayhit) override;  // host mirror, tfar in/out

  // With the hit's original triangle index (-1 on a miss).
  bool intersect(BVHRayHit& rayhit, int32_t* hit_id);
  // GPU batch: rayhits in/out, hit_ids may be null.
  void intersectBatch(BVHRayHit* rayhits, int32_t* hit_ids, size_t n);

  flg_bvh* handle() const { return m_bvh; }

 private:
  gpu::Context* m_ctx;
  uint32_t m_flags;
  flg_bvh* m_bvh = nullptr;
  gpu::HostMirror m_mirror;
  BVHBound m_bound;
};

}  // namespace experimental

// ------------------------------------------------------------ Volume ------
// Dense (x fastest, nx*ny*nz floats) or sparse-occupancy density grid in a box, with the
// OpenVDBVolume medium model (volume.cpp:33-148): sigma_t = sigma_s + sigma_a
// (first channel; grey media), majorant sigma_t * max density.  tr() and
// sample() run the GPU's trackers on a host copy of the grid and make no GPU
// call; the device copy is uploaded on first GPU use, through ctx (NULL: the
// global context, resolved then).
class GridVolume : public Volume {
 public:
  GridVolume(const float* density, int nx, int ny, int nz, const AABB& bound,
             const Vector3f& sigma_s, const Vector3f& sigma_a, Float g,
             gpu::Context* ctx = nullptr);
  // Sparse-occupancy grid: the 8^3 leaf bricks of an OpenVDB tree, which the
  // reference reads through a grid accessor (volume.cpp:40-59,76-81).
  // leaf_origin: 3 ints per leaf (voxel x, y, z of its first voxel, multiples
  // of 8 inside the grid, no two alike); leaf_values: 512 floats per leaf, x
  // fastest.  Voxels of no leaf are 0.  Host and device keep the brick pool of
  // flg_volume_create_sparse; results equal the dense grid of the same values.
  GridVolume(int n_leaves, const int32_t* leaf_origin, const float* leaf_values, int nx, int ny, int nz,
             const AABB& bound, const Vector3f& sigma_s, const Vector3f& sigma_a, Float g,
             gpu::Context* ctx = nullptr);
  ~GridVolume() override;

  Vector3f tr(const Ray& ray, Sampler& rng) const override;
  Vector3f sample(const Ray& ray, Sampler& rng, VInteraction& vi, bool& success) const override;
  void setBound(const AABB& aabb) override;

  // Replaces the density values (same dimensions; a dense grid).
  void update(const float* density);
  // Replaces the leaf values (same leaves; a sparse grid).
  void updateLeaves(const float* leaf_values);
  bool sparse() const { return m_sparse; }
  // Trilinear density at world point p (0 outside the grid): the lookup
  // both trackers make.
  float density(const Vector3f& p) const;
  flg_volume* handle() const;  // uploads on first use
  float maxDensity() const { return m_max_density; }
  gpu::Context& context() const;

 private:
  void pad(const float* density);
  mutable gpu::Context* m_ctx;
  int m_nx, m_ny, m_nz;
  Vector3f m_sigma_s, m_sigma_a;
  Float m_g;
  float m_max_density = 0;
  void fill_pool(const float* leaf_values);
  std::vector<float> m_padded;  // (nx+2)(ny+2)(nz+2), zero border
  std::vector<float> m_density;
  // sparse grids: the leaves, the brick pool and its slot map
  bool m_sparse = false;
  bool m_covers_all = false;
  std::vector<int32_t> m_origin;
  std::vector<float> m_leaf_values;
  std::vector<int32_t> m_cell_leaf;
  std::vector<float> m_pool;
  std::vector<int32_t> m_bmap;
  mutable gpu::DeviceBuffer m_dev_origin;
  mutable flg_volume* m_vol = nullptr;
  mutable gpu::DeviceBuffer m_dev;
};

// -------------------------------------------------------- Integrator ------
// Lights of the volume integrators: a point light and a constant environment
// (InfiniteAreaLight with a constant texture of colour env_radiance,
// light.cpp:81-159), both RGB.
struct GPUVolLights {
  bool has_point = false;
  Vector3f point_pos{0.0f};
  Vector3f point_intensity{0.0f};
  bool has_env = false;
  Vector3f env_radiance{0.0f};
};

// SVolIntegrator on the GPU (integrator.cpp:328-405): render() reads the
// scene description (resolution, spp, max depth, camera, scene.m_volume,
// which must be a GridVolume) and writes the mean radiance per pixel into
// scene.m_film's colour buffer.
class GPUVolIntegrator : public Integrator {
 public:
  GPUVolIntegrator(int max_depth, const GPUVolLights& lights);
  ~GPUVolIntegrator() override = default;

  // The reference's render loop and film (ParallelIntegrator::render,
  // integrator.cpp:123-235): centre path, samples 1..SPP-1 splatted through the
  // film's Mitchell filter, commitSamples; writes the film's colour buffer.
  void render(const Scene& scene) override;
  // Direct form: rgb is height*width*3 floats, row-major from pixel (0, 0).
  // film 0: per-pixel mean of spp samples on the per-sample Halton schedule;
  // film 1: the reference's render loop and film, as render(const Scene&).
  void render(const GridVolume& volume, const Vector3f& origin, const Vector3f& target,
              const Vector3f& up, Float fov, int width, int height, int spp,
              std::vector<float>& rgb, int film = 0);
  uint64_t lastLookups() const { return m_stats[1]; }

 private:
  int m_maxDepth;
  GPUVolLights m_lights;
  uint64_t m_stats[2] = {0, 0};
};

// ----------------------------------------------------------- scene file ---
// The entries of a scene file that Scene::parseXML reads (src/scene.cpp:
// 165-326; Mitsuba 0.6 XML through boost's property tree in the reference,
// a reader of that XML subset here).  File names are resolved against the
// scene file's directory, as Scene::getPath does (absolute names are kept).
struct SceneShapeDesc {
  std::string type;      // "ply" or "sphere"
  std::string filename;  // ply: the resolved file name
  std::string bsdf;      // "dielectric", "diffuse" or "roughconductor"
  Float int_ior = 1.0f, ext_ior = 1.0f;  // dielectric
  Vector3f center{0.0f};                 // sphere
  Float radius = 0;
};
struct SceneFileDesc {
  std::string base_dir;
  std::vector<SceneShapeDesc> shapes;  // in file order (the reference's m_primitives order)
  std::vector<std::string> envmaps;    // resolved file names of the envmap emitters
};
// Fills scene's resolution, spp, max depth, fov, camera and integrator type
// and returns the shapes and emitters.  Throws gpu::Error (FLG_ERR_INVALID)
// where the reference stops: an unreadable or malformed file, a missing
// version, integrator or sensor, an integrator other than path / volpath, a
// sensor that is not perspective, a bsdf other than the three above, an
// emitter that is not an envmap, a lookat vector without three components.
SceneFileDesc ParseSceneXML(const std::string& filename, Scene& scene);

// -------------------------------------------------------------- Render ----
// The backend bit of this renderer; EBackendType names the reference's own.
constexpr EBackendType kHIPBackend = static_cast<EBackendType>(1 << 3);

// A GPU RenderBase (src/render.hpp:17-26) in the place of the OptiX stub
// (src/optix/optix_render.cpp:25-66, chosen in src/main/fledge.cpp:26-30):
// init() reads the scene and picks what renders it, render() renders into
// scene.m_film's colour buffer, saveImage() writes it (".exr").
//   * a volume-integrator scene with no primitives and a GridVolume: the
//     volume-only integrator with the reference's render loop and film
//     (GPUVolIntegrator);
//   * a volume-integrator scene whose one primitive is an LBVHAccel (a
//     closed interface mesh) and whose m_volume is a GridVolume (the medium
//     inside the mesh, on its bounding box): VolPathIntegrator over the mesh
//     (flg_volpath_mesh_render) with the reference's render loop and film.
// Any other scene makes init() throw gpu::Error (FLG_ERR_INVALID).
class HIPRender : public RenderBase {
 public:
  HIPRender(Scene* scene, const GPUVolLights& lights);
  void init() override;
  bool preProcess() override { return m_init; }
  // denoise is not available (the reference's OIDN pass): false
  bool saveImage(const std::string& name, bool denoise = false) override;
  bool render() override;
  EBackendType getBackends() override { return kHIPBackend; }

  uint64_t lastLookups() const { return m_lookups; }

 private:
  enum class Mode { kNone, kVolume, kMeshMedium };
  Scene* m_scene;
  GPUVolLights m_lights;
  Mode m_mode = Mode::kNone;
  bool m_init = false;
  const LBVHAccel* m_mesh = nullptr;
  const GridVolume* m_volume = nullptr;
  uint64_t m_lookups = 0;
};

}  // namespace fledge

