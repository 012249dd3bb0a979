// C++ drop-ins over the C ABI (declarations and contracts in fledge_gpu.hpp).
// Built with hipcc so that the host mirrors compile the traversal and the
// trackers from the same headers as the GPU kernels (traverse.hpp,
// volume.hpp), with the same floating-point flags.
#include "fledge_gpu.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>
#include <iterator>
#include <new>

#include "../traverse.hpp"
#include "../volume.hpp"

namespace fledge {

// ------------------------------------------------------ memory resources
namespace {
void* aligned_or_throw(void* p, std::size_t alignment) {
  if (p && alignment && (reinterpret_cast<uintptr_t>(p) % alignment) != 0) throw std::bad_alloc();
  return p;
}
}  // namespace

void* managed_memory_resource::do_allocate(std::size_t bytes, std::size_t alignment) {
  void* p = nullptr;
  if (hipMallocManaged(&p, bytes ? bytes : 1) != hipSuccess) throw std::bad_alloc();
  return aligned_or_throw(p, alignment);
}
void managed_memory_resource::do_deallocate(void* p, std::size_t, std::size_t) { (void)hipFree(p); }
bool managed_memory_resource::do_is_equal(const std::pmr::memory_resource& o) const noexcept {
  return this == &o;
}

void* gpu_memory_resource::do_allocate(std::size_t bytes, std::size_t alignment) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw std::bad_alloc();
  return aligned_or_throw(p, alignment);
}
void gpu_memory_resource::do_deallocate(void* p, std::size_t, std::size_t) { (void)hipFree(p); }
bool gpu_memory_resource::do_is_equal(const std::pmr::memory_resource& o) const noexcept {
  return this == &o;
}

void* async_memory_resource::do_allocate(std::size_t bytes, std::size_t alignment) {
  void* p = nullptr;
  if (hipMallocAsync(&p, bytes ? bytes : 1, static_cast<hipStream_t>(m_stream)) != hipSuccess)
    throw std::bad_alloc();
  return aligned_or_throw(p, alignment);
}
void async_memory_resource::do_deallocate(void* p, std::size_t, std::size_t) {
  (void)hipFreeAsync(p, static_cast<hipStream_t>(m_stream));
}
bool async_memory_resource::do_is_equal(const std::pmr::memory_resource& o) const noexcept {
  auto* a = dynamic_cast<const async_memory_resource*>(&o);
  return a && a->m_stream == m_stream;
}

std::pmr::memory_resource* GlobalManagedMemoryResource() {
  static managed_memory_resource resource;
  return &resource;
}

namespace gpu {

// ---------------------------------------------------------------- Context
Context::Context(int device, void* stream) : m_device(device) {
  int st = flg_ctx_create(device, stream, &m_ctx);
  if (st != FLG_OK)
    throw Error(st, std::string("flg_ctx_create: ") + flg_status_string(st));
}

Context::~Context() {
  if (m_ctx) flg_ctx_destroy(m_ctx);
}

void Context::check(int status, const char* what) const {
  if (status == FLG_OK) return;
  const char* msg = m_ctx ? flg_last_error(m_ctx) : "";
  throw Error(status, std::string(what) + ": " + flg_status_string(status) + ": " + (msg ? msg : ""));
}

void Context::synchronize() const { check(flg_ctx_synchronize(m_ctx), "flg_ctx_synchronize"); }

Context& Context::global() {
  // never destroyed: a static destructor would run after the HIP runtime's
  static Context* ctx = new Context(0, nullptr);
  return *ctx;
}

// ----------------------------------------------------------- DeviceBuffer
DeviceBuffer::DeviceBuffer(const Context& ctx, size_t bytes) : m_ctx(&ctx), m_bytes(bytes) {
  if (bytes) ctx.check(flg_malloc(ctx.get(), bytes, &m_ptr), "flg_malloc");
}

DeviceBuffer::DeviceBuffer(const Context& ctx, const void* host, size_t bytes) : DeviceBuffer(ctx, bytes) {
  upload(host, bytes);
}

DeviceBuffer::~DeviceBuffer() {
  if (m_ptr) flg_free(m_ctx->get(), m_ptr);
}

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : m_ctx(o.m_ctx), m_ptr(o.m_ptr), m_bytes(o.m_bytes) {
  o.m_ptr = nullptr;
  o.m_bytes = 0;
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
  if (this != &o) {
    if (m_ptr) flg_free(m_ctx->get(), m_ptr);
    m_ctx = o.m_ctx;
    m_ptr = o.m_ptr;
    m_bytes = o.m_bytes;
    o.m_ptr = nullptr;
    o.m_bytes = 0;
  }
  return *this;
}

void DeviceBuffer::upload(const void* host, size_t bytes) {
  if (bytes > m_bytes) throw Error(FLG_ERR_RANGE, "DeviceBuffer::upload: larger than the buffer");
  if (bytes) m_ctx->check(flg_memcpy_htod(m_ctx->get(), m_ptr, host, bytes), "flg_memcpy_htod");
}

void DeviceBuffer::download(void* host, size_t bytes) const {
  if (bytes > m_bytes) throw Error(FLG_ERR_RANGE, "DeviceBuffer::download: larger than the buffer");
  if (bytes) m_ctx->check(flg_memcpy_dtoh(m_ctx->get(), host, m_ptr, bytes), "flg_memcpy_dtoh");
}

// ------------------------------------------------------------- HostMirror
void HostMirror::load(const Context& ctx, flg_bvh* bvh) {
  flg_lbvh_info info;
  ctx.check(flg_lbvh_get_info(bvh, &info), "flg_lbvh_get_info");
  m_n = info.ntris;
  std::memcpy(m_root, info.bound, sizeof(m_root));
  const size_t ni = m_n > 0 ? size_t(m_n) - 1 : 0;
  m_nodes.assign(ni * sizeof(flg::BNode), 0);
  m_tris.assign(size_t(m_n) * sizeof(flg::BTri), 0);
  if (reinterpret_cast<uintptr_t>(m_nodes.data()) % 16 || reinterpret_cast<uintptr_t>(m_tris.data()) % 16)
    throw Error(FLG_ERR_INVALID, "HostMirror: host allocation is not 16-byte aligned");
  ctx.check(flg_lbvh_download(ctx.get(), bvh, ni ? m_nodes.data() : nullptr, m_nodes.size(),
                              m_n ? m_tris.data() : nullptr, m_tris.size()),
            "flg_lbvh_download");
  m_by_id.assign(size_t(m_n), -1);
  const flg::BTri* tris = reinterpret_cast<const flg::BTri*>(m_tris.data());
  for (int32_t p = 0; p < m_n; ++p) {
    const uint32_t id = flg::f2u(tris[p].v0.w);
    if (id >= uint32_t(m_n)) throw Error(FLG_ERR_INVALID, "HostMirror: corrupt triangle ids");
    m_by_id[id] = p;
  }
}

void HostMirror::setCompat(bool compat) { m_compat = compat; }

int32_t HostMirror::intersect(const Vector3f& o, const Vector3f& d, float tmax, float* t,
                              Vector3f* ng) const {
  flg::TraceGeom g;
  g.nodes = reinterpret_cast<const flg::BNode*>(m_nodes.data());
  g.tris = reinterpret_cast<const flg::BTri*>(m_tris.data());
  g.n = m_n;
  for (int k = 0; k < 6; ++k) g.root[k] = m_root[k];
  const flg::V3 O = flg::v3(o[0], o[1], o[2]), D = flg::v3(d[0], d[1], d[2]);
  float best = tmax;
  flg::Counters c{0, 0};
  const int32_t p = m_compat ? flg::trace_one<true, false>(g, O, D, best, c)
                             : flg::trace_one<false, false>(g, O, D, best, c);
  if (p < 0) return -1;
  const flg::BTri& tri = g.tris[p];
  // the winning test once more for its normal (same inputs, same bits)
  float th;
  flg::V3 n;
  flg::pluecker(flg::v3(tri.v0.x, tri.v0.y, tri.v0.z), flg::v3(tri.v1.x, tri.v1.y, tri.v1.z),
                flg::v3(tri.v2.x, tri.v2.y, tri.v2.z), O, D, m_compat, &th, &n);
  *t = best;
  if (ng) *ng = Vector3f(n.x, n.y, n.z);
  return (int32_t)flg::f2u(tri.v0.w);
}

Vector3f HostMirror::normal(int32_t tri_id, const Vector3f& o, const Vector3f& d) const {
  if (tri_id < 0 || tri_id >= m_n) return Vector3f(0.0f);
  const flg::BTri& tri = reinterpret_cast<const flg::BTri*>(m_tris.data())[m_by_id[tri_id]];
  float th;
  flg::V3 n{0, 0, 0};
  flg::pluecker(flg::v3(tri.v0.x, tri.v0.y, tri.v0.z), flg::v3(tri.v1.x, tri.v1.y, tri.v1.z),
                flg::v3(tri.v2.x, tri.v2.y, tri.v2.z), flg::v3(o[0], o[1], o[2]),
                flg::v3(d[0], d[1], d[2]), m_compat, &th, &n);
  return Vector3f(n.x, n.y, n.z);
}

}  // namespace gpu

// -------------------------------------------------------------- LBVHAccel
namespace {
std::vector<float> flatten(const Vector3f* p, size_t n) {
  std::vector<float> v(3 * n);
  for (size_t i = 0; i < n; ++i) {
    v[3 * i] = p[i][0];
    v[3 * i + 1] = p[i][1];
    v[3 * i + 2] = p[i][2];
  }
  return v;
}

flg_ray to_flg_ray(const Vector3f& o, const Vector3f& d, float tmax) {
  flg_ray r;
  r.ox = o[0];
  r.oy = o[1];
  r.oz = o[2];
  r.tmax = tmax;
  r.dx = d[0];
  r.dy = d[1];
  r.dz = d[2];
  r.pad = 0.0f;
  return r;
}

flg_bvh* build_bvh(gpu::Context& ctx, const float* verts, int32_t nverts, const int32_t* ind,
                   int32_t ntris, uint32_t flags) {
  gpu::DeviceBuffer dv(ctx, verts, size_t(nverts) * 12);
  gpu::DeviceBuffer di(ctx, ind, size_t(ntris) * 12);
  flg_bvh* bvh = nullptr;
  ctx.check(flg_lbvh_build(ctx.get(), static_cast<const float*>(dv.data()),
                           static_cast<const int32_t*>(di.data()), ntris, nverts, flags, &bvh),
            "flg_lbvh_build");
  return bvh;  // the build synchronises: the staging buffers may go
}
}  // namespace

LBVHAccel::LBVHAccel(const std::vector<Vector3f>& p, const std::vector<int>& ind, gpu::Context& ctx,
                     uint32_t flags)
    : m_ctx(&ctx) {
  if (ind.size() % 3) throw gpu::Error(FLG_ERR_INVALID, "LBVHAccel: index count is not a multiple of 3");
  std::vector<float> v = flatten(p.data(), p.size());
  build(v.data(), (int32_t)p.size(), reinterpret_cast<const int32_t*>(ind.data()),
        (int32_t)(ind.size() / 3), flags);
}

LBVHAccel::LBVHAccel(const experimental::InternalTriangleMesh& mesh, gpu::Context& ctx, uint32_t flags)
    : m_ctx(&ctx) {
  if (mesh.nInd % 3) throw gpu::Error(FLG_ERR_INVALID, "LBVHAccel: nInd is not a multiple of 3");
  std::vector<float> v = flatten(mesh.p, size_t(mesh.nVert));
  build(v.data(), mesh.nVert, reinterpret_cast<const int32_t*>(mesh.ind), mesh.nInd / 3, flags);
}

LBVHAccel::LBVHAccel(const TriangleMesh& mesh, gpu::Context& ctx, uint32_t flags) : m_ctx(&ctx) {
  if (mesh.nInd % 3) throw gpu::Error(FLG_ERR_INVALID, "LBVHAccel: nInd is not a multiple of 3");
  std::vector<float> v = flatten(mesh.p, size_t(mesh.nVert));
  build(v.data(), mesh.nVert, reinterpret_cast<const int32_t*>(mesh.ind), mesh.nInd / 3, flags);
}

void LBVHAccel::build(const float* verts, int32_t nverts, const int32_t* ind, int32_t ntris,
                      uint32_t flags) {
  m_bvh = build_bvh(*m_ctx, verts, nverts, ind, ntris, flags);
  m_mirror.setCompat((flags & FLG_NORMAL_COMPAT) != 0);
  m_mirror.load(*m_ctx, m_bvh);
  flg_lbvh_info info;
  m_ctx->check(flg_lbvh_get_info(m_bvh, &info), "flg_lbvh_get_info");
  m_bound = AABB(Vector3f(info.bound[0], info.bound[1], info.bound[2]),
                 Vector3f(info.bound[3], info.bound[4], info.bound[5]));
}

LBVHAccel::~LBVHAccel() {
  if (m_bvh) flg_lbvh_destroy(m_bvh);
}

bool LBVHAccel::intersect(const Ray& ray, SInteraction& isect) const {
  return intersect(ray, isect, nullptr);
}

bool LBVHAccel::intersect(const Ray& ray, SInteraction& isect, int32_t* tri_id) const {
  float t;
  Vector3f ng;
  const int32_t id = m_mirror.intersect(ray.m_o, ray.m_d, ray.m_tMax, &t, &ng);
  if (tri_id) *tri_id = id;
  if (id < 0) return false;
  ray.m_tMax = t;
  isect.m_p = ray(t);
  isect.m_ng = ng;
  isect.m_ns = ng;
  isect.m_wo = -ray.m_d;
  isect.m_ray = ray;
  isect.m_primitive = this;
  return true;
}

AABB LBVHAccel::getBound() const { return m_bound; }

void LBVHAccel::intersectBatch(const flg_ray* d_rays, flg_hit* d_hits, int64_t n) const {
  m_ctx->check(flg_trace_closest(m_ctx->get(), m_bvh, d_rays, d_hits, n, nullptr), "flg_trace_closest");
}

std::vector<flg_hit> LBVHAccel::intersectBatch(const std::vector<Ray>& rays) const {
  std::vector<flg_ray> r(rays.size());
  for (size_t i = 0; i < rays.size(); ++i) r[i] = to_flg_ray(rays[i].m_o, rays[i].m_d, rays[i].m_tMax);
  std::vector<flg_hit> h(rays.size());
  if (rays.empty()) return h;
  gpu::DeviceBuffer dr(*m_ctx, r.data(), r.size() * sizeof(flg_ray));
  gpu::DeviceBuffer dh(*m_ctx, h.size() * sizeof(flg_hit));
  intersectBatch(static_cast<const flg_ray*>(dr.data()), static_cast<flg_hit*>(dh.data()),
                 (int64_t)rays.size());
  dh.download(h.data(), h.size() * sizeof(flg_hit));
  return h;
}

// ------------------------------------------------------------ LBVHBuilder
namespace experimental {

LBVHBuilder::LBVHBuilder(InternalTriangleMesh* mesh, std::pmr::memory_resource* mem_resource,
                         gpu::Context* ctx, uint32_t flags)
    : BVHBuilderBase(mesh, mem_resource), m_ctx(ctx ? ctx : &gpu::Context::global()), m_flags(flags) {}

LBVHBuilder::~LBVHBuilder() {
  if (m_bvh) flg_lbvh_destroy(m_bvh);
}

void LBVHBuilder::build() {
  if (m_bvh) {
    flg_lbvh_destroy(m_bvh);
    m_bvh = nullptr;
  }
  if (m_mesh->nInd % 3) throw gpu::Error(FLG_ERR_INVALID, "LBVHBuilder: nInd is not a multiple of 3");
  std::vector<float> v = flatten(m_mesh->p, size_t(m_mesh->nVert));
  m_bvh = build_bvh(*m_ctx, v.data(), m_mesh->nVert, reinterpret_cast<const int32_t*>(m_mesh->ind),
                    m_mesh->nInd / 3, m_flags);
  m_mirror.setCompat((m_flags & FLG_NORMAL_COMPAT) != 0);
  m_mirror.load(*m_ctx, m_bvh);
  flg_lbvh_info info;
  m_ctx->check(flg_lbvh_get_info(m_bvh, &info), "flg_lbvh_get_info");
  m_bound.lower = Vector3f(info.bound[0], info.bound[1], info.bound[2]);
  m_bound.upper = Vector3f(info.bound[3], info.bound[4], info.bound[5]);
}

BVHBuilderBase::BVHBound LBVHBuilder::getBound() const { return m_bound; }

bool LBVHBuilder::intersect(BVHRayHit& rayhit) { return intersect(rayhit, nullptr); }

bool LBVHBuilder::intersect(BVHRayHit& rayhit, int32_t* hit_id) {
  if (!m_bvh) throw gpu::Error(FLG_ERR_INVALID, "LBVHBuilder::intersect before build()");
  float t;
  Vector3f ng;
  const int32_t id = m_mirror.intersect(rayhit.ray_o, rayhit.ray_d, rayhit.tfar, &t, &ng);
  if (hit_id) *hit_id = id;
  if (id < 0) return false;
  rayhit.hit = true;
  rayhit.tfar = t;
  rayhit.hit_ng = rayhit.hit_ns = ng;
  return true;
}

void LBVHBuilder::intersectBatch(BVHRayHit* rayhits, int32_t* hit_ids, size_t n) {
  if (!m_bvh) throw gpu::Error(FLG_ERR_INVALID, "LBVHBuilder::intersectBatch before build()");
  if (n == 0) return;
  std::vector<flg_ray> r(n);
  for (size_t i = 0; i < n; ++i) r[i] = to_flg_ray(rayhits[i].ray_o, rayhits[i].ray_d, rayhits[i].tfar);
  std::vector<flg_hit> h(n);
  gpu::DeviceBuffer dr(*m_ctx, r.data(), n * sizeof(flg_ray));
  gpu::DeviceBuffer dh(*m_ctx, n * sizeof(flg_hit));
  m_ctx->check(flg_trace_closest(m_ctx->get(), m_bvh, static_cast<const flg_ray*>(dr.data()),
                                 static_cast<flg_hit*>(dh.data()), (int64_t)n, nullptr),
               "flg_trace_closest");
  dh.download(h.data(), n * sizeof(flg_hit));
  for (size_t i = 0; i < n; ++i) {
    if (hit_ids) hit_ids[i] = h[i].id;
    if (h[i].id < 0) continue;
    rayhits[i].hit = true;
    rayhits[i].tfar = h[i].t;
    rayhits[i].hit_ng = rayhits[i].hit_ns = m_mirror.normal(h[i].id, rayhits[i].ray_o, rayhits[i].ray_d);
  }
}

}  // namespace experimental

// ------------------------------------------------------------- GridVolume
namespace {
struct SamplerDraw {  // the trackers' draw source over a fledge Sampler
  Sampler& s;
  float next() { return s.get1D(); }
};
}  // namespace

GridVolume::GridVolume(const float* density, int nx, int ny, int nz, const AABB& bound,
                       const Vector3f& sigma_s, const Vector3f& sigma_a, Float g, gpu::Context* ctx)
    : m_ctx(ctx), m_nx(nx), m_ny(ny), m_nz(nz), m_sigma_s(sigma_s), m_sigma_a(sigma_a), m_g(g) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !density)
    throw gpu::Error(FLG_ERR_INVALID, "GridVolume: bad grid");
  if ((int64_t)(nx + 2) * (ny + 2) * (nz + 2) >= ((int64_t)1 << 31))
    throw gpu::Error(FLG_ERR_RANGE, "GridVolume: grid too large");
  for (int k = 0; k < 3; ++k)
    if (!(bound.m_max[k] > bound.m_min[k])) throw gpu::Error(FLG_ERR_INVALID, "GridVolume: empty bounds");
  m_aabb = bound;
  pad(density);
}

GridVolume::GridVolume(int n_leaves, const int32_t* leaf_origin, const float* leaf_values, int nx, int ny, int nz,
                       const AABB& bound, const Vector3f& sigma_s, const Vector3f& sigma_a, Float g,
                       gpu::Context* ctx)
    : m_ctx(ctx), m_nx(nx), m_ny(ny), m_nz(nz), m_sigma_s(sigma_s), m_sigma_a(sigma_a), m_g(g), m_sparse(true) {
  using flg::kPoolBrick;
  if (nx <= 0 || ny <= 0 || nz <= 0 || n_leaves < 0 || (n_leaves > 0 && (!leaf_origin || !leaf_values)))
    throw gpu::Error(FLG_ERR_INVALID, "GridVolume: bad sparse grid");
  for (int k = 0; k < 3; ++k)
    if (!(bound.m_max[k] > bound.m_min[k])) throw gpu::Error(FLG_ERR_INVALID, "GridVolume: empty bounds");
  const int64_t nbx = (nx + kPoolBrick - 1) / kPoolBrick, nby = (ny + kPoolBrick - 1) / kPoolBrick,
                nbz = (nz + kPoolBrick - 1) / kPoolBrick;
  if ((nbx + 1) * (nby + 1) * (nbz + 1) * flg::kPoolClasses >= ((int64_t)1 << 31) ||
      (int64_t)n_leaves * flg::kPoolFloats >= ((int64_t)1 << 31))
    throw gpu::Error(FLG_ERR_RANGE, "GridVolume: sparse grid too large");
  m_aabb = bound;
  m_origin.assign(leaf_origin, leaf_origin + size_t(n_leaves) * 3);
  // brick cell -> leaf, as flg_volume_create_sparse checks and builds it
  m_cell_leaf.assign(size_t(nbx * nby * nbz), -1);
  uint64_t covered = 0;
  for (int l = 0; l < n_leaves; ++l) {
    const int32_t ox = m_origin[3 * l], oy = m_origin[3 * l + 1], oz = m_origin[3 * l + 2];
    if (ox < 0 || oy < 0 || oz < 0 || ox >= nx || oy >= ny || oz >= nz || ((ox | oy | oz) & (kPoolBrick - 1)))
      throw gpu::Error(FLG_ERR_INVALID, "GridVolume: a leaf origin is outside the grid or not a multiple of 8");
    int32_t& c = m_cell_leaf[size_t(((int64_t)(oz >> 3) * nby + (oy >> 3)) * nbx + (ox >> 3))];
    if (c >= 0) throw gpu::Error(FLG_ERR_INVALID, "GridVolume: two leaves share an origin");
    c = l;
    covered += uint64_t(std::min(kPoolBrick, nx - ox)) * std::min(kPoolBrick, ny - oy) * std::min(kPoolBrick, nz - oz);
  }
  m_covers_all = covered == uint64_t(nx) * ny * nz;
  // slot map: per cell (-1 .. nb-1 per axis) and border class, the first
  // stored brick among the cells the lookup's corners touch
  const int64_t bcx = nbx + 1, bcy = nby + 1, bcz = nbz + 1;
  m_bmap.assign(size_t(bcx * bcy * bcz * flg::kPoolClasses), -1);
  for (int64_t cell = 0; cell < bcx * bcy * bcz; ++cell) {
    const int cx = int(cell % bcx) - 1, cy = int((cell / bcx) % bcy) - 1, cz = int(cell / (bcx * bcy)) - 1;
    for (int c = 0; c < flg::kPoolClasses; ++c)
      for (int d = 0; d < flg::kPoolClasses; ++d) {
        if (d & ~c) continue;
        const int x = cx + (d & 1), y = cy + ((d >> 1) & 1), z = cz + ((d >> 2) & 1);
        if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
        const int32_t l = m_cell_leaf[size_t((z * nby + y) * nbx + x)];
        if (l >= 0) {
          m_bmap[size_t(cell * flg::kPoolClasses + c)] = (l << 3) | d;
          break;
        }
      }
  }
  fill_pool(leaf_values);
}

// Pool slots (the leaf and its one-voxel apron) and the maximum density of
// the grid, from the leaf values.
void GridVolume::fill_pool(const float* leaf_values) {
  using flg::kPoolBrick;
  using flg::kPoolEdge;
  const int n_leaves = int(m_origin.size() / 3);
  m_leaf_values.assign(leaf_values, leaf_values + size_t(n_leaves) * 512);
  const int nbx = (m_nx + kPoolBrick - 1) / kPoolBrick, nby = (m_ny + kPoolBrick - 1) / kPoolBrick;
  m_pool.assign(size_t(n_leaves > 0 ? n_leaves : 1) * flg::kPoolFloats, 0.0f);
  float mx = m_covers_all ? -3.402823466e+38f : 0.0f;
  for (int l = 0; l < n_leaves; ++l) {
    const int ox = m_origin[3 * l], oy = m_origin[3 * l + 1], oz = m_origin[3 * l + 2];
    for (int a = 0; a < flg::kPoolFloats; ++a) {
      const int x = ox - 1 + a % kPoolEdge, y = oy - 1 + (a / kPoolEdge) % kPoolEdge,
                z = oz - 1 + a / (kPoolEdge * kPoolEdge);
      float v = 0.0f;
      if (x >= 0 && y >= 0 && z >= 0 && x < m_nx && y < m_ny && z < m_nz) {
        const int32_t src = m_cell_leaf[size_t(((int64_t)(z >> 3) * nby + (y >> 3)) * nbx + (x >> 3))];
        if (src >= 0) v = m_leaf_values[(size_t(src) << 9) + size_t(((z & 7) << 6) | ((y & 7) << 3) | (x & 7))];
      }
      m_pool[size_t(l) * flg::kPoolFloats + a] = v;
    }
    for (int q = 0; q < 512; ++q) {
      const int x = ox + (q & 7), y = oy + ((q >> 3) & 7), z = oz + (q >> 6);
      if (x < m_nx && y < m_ny && z < m_nz) mx = std::max(mx, m_leaf_values[(size_t(l) << 9) + q]);
    }
  }
  m_max_density = mx;
}

void GridVolume::updateLeaves(const float* leaf_values) {
  if (!m_sparse) throw gpu::Error(FLG_ERR_INVALID, "GridVolume::updateLeaves: a dense grid takes update()");
  if (!leaf_values && !m_origin.empty()) throw gpu::Error(FLG_ERR_INVALID, "GridVolume::updateLeaves: no values");
  fill_pool(leaf_values);
  if (m_vol && !m_origin.empty()) {
    m_dev.upload(m_leaf_values.data(), m_leaf_values.size() * 4);
    context().check(flg_volume_update_sparse(context().get(), m_vol, static_cast<const float*>(m_dev.data())),
                    "flg_volume_update_sparse");
  }
}

gpu::Context& GridVolume::context() const {
  if (!m_ctx) m_ctx = &gpu::Context::global();
  return *m_ctx;
}

GridVolume::~GridVolume() {
  if (m_vol) flg_volume_destroy(m_vol);
}

void GridVolume::pad(const float* density) {
  const size_t n = size_t(m_nx) * m_ny * m_nz;
  m_density.assign(density, density + n);
  const int px = m_nx + 2, py = m_ny + 2, pz = m_nz + 2;
  m_padded.assign(size_t(px) * py * pz, 0.0f);
  float mx = 0.0f;
  for (int k = 0; k < m_nz; ++k)
    for (int j = 0; j < m_ny; ++j)
      for (int i = 0; i < m_nx; ++i) {
        const float v = density[(size_t(k) * m_ny + j) * m_nx + i];
        m_padded[(size_t(k + 1) * py + (j + 1)) * px + (i + 1)] = v;
        mx = std::max(mx, v);
      }
  m_max_density = mx;
}

void GridVolume::update(const float* density) {
  if (m_sparse) throw gpu::Error(FLG_ERR_INVALID, "GridVolume::update: a sparse grid takes updateLeaves()");
  pad(density);
  if (m_vol) {
    m_dev.upload(m_density.data(), m_density.size() * 4);
    context().check(flg_volume_update(context().get(), m_vol, static_cast<const float*>(m_dev.data())),
                 "flg_volume_update");
  }
}

void GridVolume::setBound(const AABB& aabb) {
  for (int k = 0; k < 3; ++k)
    if (!(aabb.m_max[k] > aabb.m_min[k])) throw gpu::Error(FLG_ERR_INVALID, "GridVolume: empty bounds");
  m_aabb = aabb;
  if (m_vol) {  // the device copy is rebuilt with the new box on next use
    flg_volume_destroy(m_vol);
    m_vol = nullptr;
  }
}

flg_volume* GridVolume::handle() const {
  if (!m_vol) {
    if (!m_sparse && !m_dev.data()) m_dev = gpu::DeviceBuffer(context(), m_density.data(), m_density.size() * 4);
    const float b[6] = {m_aabb.m_min[0], m_aabb.m_min[1], m_aabb.m_min[2],
                        m_aabb.m_max[0], m_aabb.m_max[1], m_aabb.m_max[2]};
    flg_medium med{m_sigma_s[0], m_sigma_a[0], m_g};
    if (m_sparse) {
      const int n_leaves = int(m_origin.size() / 3);
      if (n_leaves > 0 && !m_dev_origin.data()) {
        m_dev_origin = gpu::DeviceBuffer(context(), m_origin.data(), m_origin.size() * 4);
        m_dev = gpu::DeviceBuffer(context(), m_leaf_values.data(), m_leaf_values.size() * 4);
      }
      context().check(flg_volume_create_sparse(context().get(), n_leaves,
                                               static_cast<const int32_t*>(m_dev_origin.data()),
                                               static_cast<const float*>(m_dev.data()), m_nx, m_ny, m_nz, b, &med,
                                               &m_vol),
                      "flg_volume_create_sparse");
      return m_vol;
    }
    context().check(flg_volume_create(context().get(), static_cast<const float*>(m_dev.data()), m_nx, m_ny,
                                   m_nz, b, &med, &m_vol),
                 "flg_volume_create");
  }
  return m_vol;
}

namespace {
flg::VolDev host_voldev(const float* data, const int32_t* bmap, int nx, int ny, int nz, const AABB& box,
                        const Vector3f& ss, const Vector3f& sa, Float g, float max_density) {
  const float b[6] = {box.m_min[0], box.m_min[1], box.m_min[2], box.m_max[0], box.m_max[1], box.m_max[2]};
  return flg::voldev_of(data, bmap, nx, ny, nz, b, flg_medium{ss[0], sa[0], g}, max_density);
}
}  // namespace

float GridVolume::density(const Vector3f& p) const {
  const flg::VolDev v = m_sparse ? host_voldev(m_pool.data(), m_bmap.data(), m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density)
                                 : host_voldev(m_padded.data(), nullptr, m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density);
  return flg::lookup(v, flg::v3(p[0], p[1], p[2])).rho;
}

Vector3f GridVolume::tr(const Ray& ray, Sampler& rng) const {
  const flg::VolDev v = m_sparse ? host_voldev(m_pool.data(), m_bmap.data(), m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density)
                                 : host_voldev(m_padded.data(), nullptr, m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density);
  SamplerDraw draw{rng};
  uint32_t nlook = 0;
  const float t = flg::vol_tr<false>(v, draw, flg::v3(ray.m_o[0], ray.m_o[1], ray.m_o[2]),
                                     flg::v3(ray.m_d[0], ray.m_d[1], ray.m_d[2]), ray.m_tMax, nlook);
  return Vector3f(t);
}

Vector3f GridVolume::sample(const Ray& ray, Sampler& rng, VInteraction& vi, bool& success) const {
  const flg::VolDev v = m_sparse ? host_voldev(m_pool.data(), m_bmap.data(), m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density)
                                 : host_voldev(m_padded.data(), nullptr, m_nx, m_ny, m_nz, m_aabb, m_sigma_s,
                                               m_sigma_a, m_g, m_max_density);
  SamplerDraw draw{rng};
  uint32_t nlook = 0;
  bool ok = false;
  flg::V3 pos{0, 0, 0};
  const float f = flg::vol_sample<false>(v, draw, flg::v3(ray.m_o[0], ray.m_o[1], ray.m_o[2]),
                                         flg::v3(ray.m_d[0], ray.m_d[1], ray.m_d[2]), ray.m_tMax, &ok,
                                         &pos, nlook);
  success = ok;
  if (ok) vi = VInteraction(Vector3f(pos.x, pos.y, pos.z), -ray.m_d, m_g);
  return Vector3f(f);
}

// ------------------------------------------------------- GPUVolIntegrator
GPUVolIntegrator::GPUVolIntegrator(int max_depth, const GPUVolLights& lights)
    : m_maxDepth(max_depth), m_lights(lights) {}

void GPUVolIntegrator::render(const GridVolume& volume, const Vector3f& origin, const Vector3f& target,
                              const Vector3f& up, Float fov, int width, int height, int spp,
                              std::vector<float>& rgb, int film) {
  if (width <= 0 || height <= 0 || spp <= 0)
    throw gpu::Error(FLG_ERR_INVALID, "GPUVolIntegrator::render: bad resolution or spp");
  gpu::Context& ctx = volume.context();
  flg_vol_scene sc;
  std::memset(&sc, 0, sizeof(sc));
  for (int k = 0; k < 3; ++k) {
    sc.camera.pos[k] = origin[k];
    sc.camera.look_at[k] = target[k];
    sc.camera.up[k] = up[k];
    sc.point_pos[k] = m_lights.point_pos[k];
    sc.point_intensity[k] = m_lights.point_intensity[k];
    sc.env_radiance[k] = m_lights.env_radiance[k];
  }
  sc.camera.fov_deg = fov;
  sc.has_point = m_lights.has_point ? 1 : 0;
  sc.has_env = m_lights.has_env ? 1 : 0;
  flg_render_cfg cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.width = width;
  cfg.height = height;
  cfg.spp = spp;
  cfg.max_depth = m_maxDepth;
  cfg.film = film;
  const size_t n = size_t(width) * height * 3;
  gpu::DeviceBuffer d(ctx, n * sizeof(float));
  ctx.check(flg_volpath_render(ctx.get(), volume.handle(), &sc, &cfg, static_cast<float*>(d.data()), m_stats),
            "flg_volpath_render");
  rgb.resize(n);
  d.download(rgb.data(), n * sizeof(float));
}

#ifndef FLEDGE_HAS_REFERENCE_HEADERS
bool Film::saveBuffer(const std::string& name, EFilmBufferType) const {
  const size_t dot = name.rfind('.');
  if (dot == std::string::npos || name.substr(dot) != ".exr") return false;
  std::vector<float> rgb(size_t(m_resX) * m_resY * 3);
  for (size_t i = 0; i < m_color.size(); ++i)
    for (int c = 0; c < 3; ++c) rgb[3 * i + c] = m_color[i][c];
  return flg_film_save_exr(name.c_str(), m_resX, m_resY, rgb.data()) == FLG_OK;
}
#endif

void GPUVolIntegrator::render(const Scene& scene) {
  const GridVolume* vol = dynamic_cast<const GridVolume*>(scene.m_volume);
  if (!vol) throw gpu::Error(FLG_ERR_INVALID, "GPUVolIntegrator: scene.m_volume is not a GridVolume");
  if (!scene.m_film) throw gpu::Error(FLG_ERR_INVALID, "GPUVolIntegrator: scene has no film");
  std::vector<float> rgb;
  render(*vol, scene.m_origin, scene.m_target, scene.m_up, scene.m_FoV, scene.m_resX, scene.m_resY,
         scene.m_SPP, rgb, /*film=*/1);
  for (int y = 0; y < scene.m_resY; ++y)
    for (int x = 0; x < scene.m_resX; ++x) {
      const float* c = &rgb[(size_t(y) * scene.m_resX + x) * 3];
      scene.m_film->getBuffer(x, y, EFilmBufferType::EColor) = Vector3f(c[0], c[1], c[2]);
    }
}

// ------------------------------------------------------------------ PLY ---
std::unique_ptr<PlyMesh> MakeTriangleMesh(const std::string& path) {
  flg_ply_mesh m;
  const int st = flg_ply_load(path.c_str(), &m);
  if (st != FLG_OK) throw gpu::Error(st, "MakeTriangleMesh: cannot read " + path);
  auto out = std::make_unique<PlyMesh>();
  out->p.resize((size_t)m.nverts);
  for (int32_t i = 0; i < m.nverts; ++i) out->p[i] = Vector3f(m.verts[3 * i], m.verts[3 * i + 1], m.verts[3 * i + 2]);
  if (m.normals) {
    out->n.resize((size_t)m.nverts);
    for (int32_t i = 0; i < m.nverts; ++i)
      out->n[i] = Vector3f(m.normals[3 * i], m.normals[3 * i + 1], m.normals[3 * i + 2]);
  }
  if (m.uv) {
    out->uv.resize((size_t)m.nverts);
    for (int32_t i = 0; i < m.nverts; ++i) out->uv[i] = Vector2f{{m.uv[2 * i], m.uv[2 * i + 1]}};
  }
  out->ind.assign(m.ind, m.ind + (size_t)m.ntris * 3);
  flg_ply_free(&m);
  out->mesh.nVert = (int)out->p.size();
  out->mesh.nInd = (int)out->ind.size();
  out->mesh.p = out->p.data();
  out->mesh.n = out->n.empty() ? nullptr : out->n.data();
  out->mesh.uv = out->uv.empty() ? nullptr : out->uv.data();
  out->mesh.ind = out->ind.data();
  return out;
}

// ---------------------------------------------------------- SceneAccel ---
namespace {

// AABB::intersect_pbrt (src/common/aabb.h:38-59)
bool intersect_pbrt(const AABB& b, const Ray& ray, Float* t_min) {
  Float t0 = 0, t1 = ray.m_tMax;
  for (int i = 0; i < 3; ++i) {
    const Float inv = 1 / ray.m_d[i];
    Float tn = (b.m_min[i] - ray.m_o[i]) * inv, tf = (b.m_max[i] - ray.m_o[i]) * inv;
    if (tn > tf) std::swap(tn, tf);
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  *t_min = t0;
  return true;
}

}  // namespace

SceneAccel::SceneAccel(const std::vector<const LBVHAccel*>& meshes, gpu::Context& ctx)
    : m_ctx(&ctx), m_meshes(meshes) {
  if (meshes.empty()) throw gpu::Error(FLG_ERR_INVALID, "SceneAccel: no meshes");
  std::vector<flg_bvh*> h(meshes.size());
  for (size_t i = 0; i < meshes.size(); ++i) h[i] = meshes[i]->handle();
  m_ctx->check(flg_scene_accel_build(m_ctx->get(), h.data(), (int32_t)h.size(), &m_accel), "flg_scene_accel_build");
  std::vector<int32_t> rank(meshes.size());
  m_ctx->check(flg_scene_accel_rank(m_accel, rank.data()), "flg_scene_accel_rank");
  m_by_rank.resize(meshes.size());
  for (size_t i = 0; i < meshes.size(); ++i) m_by_rank[rank[i]] = (int32_t)i;
  // the root box: the default box at the origin merged with every bound
  Vector3f lo(0.0f), hi(0.0f);
  for (const LBVHAccel* m : meshes) {
    const AABB b = m->getBound();
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.m_min[k]);
      hi[k] = std::max(hi[k], b.m_max[k]);
    }
  }
  m_box = AABB(lo, hi);
}

SceneAccel::~SceneAccel() {
  if (m_accel) flg_scene_accel_destroy(m_accel);
}

bool SceneAccel::intersect(const Ray& ray, SInteraction& isect) const {
  return intersect(ray, isect, nullptr, nullptr);
}

bool SceneAccel::intersect(const Ray& ray, SInteraction& isect, int32_t* mesh, int32_t* tri_id) const {
  if (mesh) *mesh = -1;
  if (tri_id) *tri_id = -1;
  Float t_min;
  if (!intersect_pbrt(m_box, ray, &t_min) || ray.m_tMax < t_min) return false;
  const Float o_tmax = ray.m_tMax;
  Float best = o_tmax;
  SInteraction t_isect;
  for (int32_t m : m_by_rank) {
    int32_t id = -1;
    if (m_meshes[m]->intersect(ray, t_isect, &id) && ray.m_tMax < best) {
      isect = t_isect;
      best = ray.m_tMax;
      if (mesh) *mesh = m;
      if (tri_id) *tri_id = id;
    }
  }
  return best != o_tmax;
}

AABB SceneAccel::getBound() const { return m_box; }

void SceneAccel::intersectBatch(const flg_ray* d_rays, flg_scene_hit* d_hits, int64_t n) const {
  m_ctx->check(flg_scene_trace_closest(m_ctx->get(), m_accel, d_rays, d_hits, n), "flg_scene_trace_closest");
}

std::vector<flg_scene_hit> SceneAccel::intersectBatch(const std::vector<Ray>& rays) const {
  std::vector<flg_ray> r(rays.size());
  for (size_t i = 0; i < rays.size(); ++i) r[i] = to_flg_ray(rays[i].m_o, rays[i].m_d, rays[i].m_tMax);
  std::vector<flg_scene_hit> h(rays.size());
  if (rays.empty()) return h;
  gpu::DeviceBuffer dr(*m_ctx, r.data(), r.size() * sizeof(flg_ray));
  gpu::DeviceBuffer dh(*m_ctx, h.size() * sizeof(flg_scene_hit));
  intersectBatch(static_cast<const flg_ray*>(dr.data()), static_cast<flg_scene_hit*>(dh.data()),
                 (int64_t)rays.size());
  m_ctx->check(flg_memcpy_dtoh(m_ctx->get(), h.data(), dh.data(), h.size() * sizeof(flg_scene_hit)),
               "flg_memcpy_dtoh");
  return h;
}

// ----------------------------------------------------------- scene file ---
namespace {

// An XML element: name, attributes in order, child elements (text is not
// kept: the scene format carries its values in attributes).
struct XNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XNode> kids;
  const std::string* attr(const char* k) const {
    for (const auto& a : attrs)
      if (a.first == k) return &a.second;
    return nullptr;
  }
  const XNode* child(const char* k) const {
    for (const auto& c : kids)
      if (c.name == k) return &c;
    return nullptr;
  }
};

[[noreturn]] void scene_error(const std::string& what) { throw gpu::Error(FLG_ERR_INVALID, "scene file: " + what); }

// Elements, attributes, comments, processing instructions, a DOCTYPE line
// and the five predefined entities.
class XReader {
 public:
  explicit XReader(const std::string& text) : s_(text) {}
  XNode document() {
    skip_misc();
    XNode root = element(0);
    skip_misc();
    if (i_ != s_.size()) scene_error("content after the root element");
    return root;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  void skip_to(const char* end) {
    const size_t e = s_.find(end, i_);
    if (e == std::string::npos) scene_error("unterminated markup");
    i_ = e + std::strlen(end);
  }
  void skip_misc() {  // whitespace, comments, <?...?>, <!DOCTYPE ...>
    while (true) {
      ws();
      if (starts("<!--")) skip_to("-->");
      else if (starts("<?")) skip_to("?>");
      else if (starts("<!")) skip_to(">");
      else return;
    }
  }
  std::string name() {
    const size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == '-' ||
                              s_[i_] == '.' || s_[i_] == ':'))
      ++i_;
    if (i_ == b) scene_error("a name was expected");
    return s_.substr(b, i_ - b);
  }
  std::string unescape(const std::string& v) {
    std::string o;
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] != '&') {
        o += v[k];
        continue;
      }
      const size_t e = v.find(';', k);
      if (e == std::string::npos) scene_error("bad entity");
      const std::string ent = v.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else scene_error("unknown entity &" + ent + ";");
      k = e;
    }
    return o;
  }
  // Scene files nest a few levels (scene, shape, transform, lookat); the
  // bound keeps a hostile file from exhausting the stack.
  static constexpr int kMaxNesting = 256;
  XNode element(int depth) {
    if (depth >= kMaxNesting) scene_error("elements nested too deeply");
    if (!starts("<")) scene_error("an element was expected");
    ++i_;
    XNode n;
    n.name = name();
    while (true) {
      ws();
      if (starts("/>")) {
        i_ += 2;
        return n;
      }
      if (starts(">")) {
        ++i_;
        break;
      }
      std::string k = name();
      ws();
      if (!starts("=")) scene_error("'=' expected after attribute " + k);
      ++i_;
      ws();
      if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) scene_error("a quoted value was expected");
      const char q = s_[i_++];
      const size_t e = s_.find(q, i_);
      if (e == std::string::npos) scene_error("unterminated attribute value");
      n.attrs.emplace_back(k, unescape(s_.substr(i_, e - i_)));
      i_ = e + 1;
    }
    while (true) {  // content: child elements; text and comments are skipped
      const size_t lt = s_.find('<', i_);
      if (lt == std::string::npos) scene_error("unterminated element " + n.name);
      i_ = lt;
      if (starts("</")) {
        i_ += 2;
        if (name() != n.name) scene_error("mismatched end tag of " + n.name);
        ws();
        if (!starts(">")) scene_error("'>' expected");
        ++i_;
        return n;
      }
      if (starts("<!--")) skip_to("-->");
      else if (starts("<?")) skip_to("?>");
      else n.kids.push_back(element(depth + 1));
    }
  }
};

const std::string& need_attr(const XNode& n, const char* k) {
  const std::string* v = n.attr(k);
  if (!v) scene_error("<" + n.name + "> has no " + k);
  return *v;
}

int to_int(const std::string& v) {
  size_t used = 0;
  int x = 0;
  try {
    x = std::stoi(v, &used);
  } catch (const std::exception&) {
    scene_error("not an integer: " + v);
  }
  if (used != v.size()) scene_error("not an integer: " + v);
  return x;
}

Float to_float(const std::string& v) {
  size_t used = 0;
  Float x = 0;
  try {
    x = std::stof(v, &used);
  } catch (const std::exception&) {
    scene_error("not a number: " + v);
  }
  if (used != v.size()) scene_error("not a number: " + v);
  return x;
}

// parseVector3f (scene.cpp:70-80): three comma-separated numbers
Vector3f to_vec3(const std::string& v) {
  std::vector<std::string> tok;
  size_t b = 0;
  while (true) {
    const size_t e = v.find(',', b);
    tok.push_back(v.substr(b, e == std::string::npos ? std::string::npos : e - b));
    if (e == std::string::npos) break;
    b = e + 1;
  }
  if (tok.size() != 3) scene_error("a vector of three components was expected: " + v);
  Vector3f r;
  for (int k = 0; k < 3; ++k) {
    size_t used = 0;
    try {
      r[k] = std::stof(tok[k], &used);
    } catch (const std::exception&) {
      scene_error("not a number: " + tok[k]);
    }
  }
  return r;
}

std::string resolve(const std::string& base, const std::string& p) {
  if (!p.empty() && p[0] == '/') return p;
  if (base.empty()) return p;
  return base + "/" + p;
}

}  // namespace

SceneFileDesc ParseSceneXML(const std::string& filename, Scene& scene) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) scene_error("cannot read " + filename);
  const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const XNode root = XReader(text).document();
  if (root.name != "scene") scene_error("the root element is not <scene>");
  (void)need_attr(root, "version");
  SceneFileDesc out;
  const size_t slash = filename.rfind('/');
  out.base_dir = slash == std::string::npos ? std::string(".") : filename.substr(0, slash);
  // setIntegrator (scene.cpp:83-112)
  const XNode* integ = root.child("integrator");
  if (!integ) scene_error("no <integrator>");
  const std::string& itype = need_attr(*integ, "type");
  if (itype == "path") scene.m_integrator_type = EIntegratorType::EPathIntegrator;
  else if (itype == "volpath") scene.m_integrator_type = EIntegratorType::EVolPathIntegrator;
  else scene_error("integrator type " + itype);
  for (const XNode& c : integ->kids)
    if (c.name == "integer" && need_attr(c, "name") == "maxDepth") scene.m_maxDepth = to_int(need_attr(c, "value"));
  // setSensor (scene.cpp:114-163)
  const XNode* sensor = root.child("sensor");
  if (!sensor) scene_error("no <sensor>");
  if (need_attr(*sensor, "type") != "perspective") scene_error("sensor type " + need_attr(*sensor, "type"));
  for (const XNode& c : sensor->kids)
    if (c.name == "float" && need_attr(c, "name") == "fov") scene.m_FoV = to_float(need_attr(c, "value"));
  const XNode* tr = sensor->child("transform");
  const XNode* look = tr ? tr->child("lookat") : nullptr;
  if (!look) scene_error("no <transform><lookat>");
  scene.m_up = to_vec3(need_attr(*look, "up"));
  scene.m_origin = to_vec3(need_attr(*look, "origin"));
  scene.m_target = to_vec3(need_attr(*look, "target"));
  const XNode* sampler = sensor->child("sampler");
  const XNode* count = sampler ? sampler->child("integer") : nullptr;
  if (!count) scene_error("no <sampler><integer>");
  scene.m_SPP = to_int(need_attr(*count, "value"));
  const XNode* film = sensor->child("film");
  if (!film) scene_error("no <film>");
  for (const XNode& c : film->kids) {
    if (c.name != "integer") continue;
    const std::string& nm = need_attr(c, "name");
    if (nm == "height") scene.m_resY = to_int(need_attr(c, "value"));
    if (nm == "width") scene.m_resX = to_int(need_attr(c, "value"));
  }
  // shapes and emitters in file order (scene.cpp:165-303)
  for (const XNode& c : root.kids) {
    if (c.name == "shape") {
      SceneShapeDesc sh;
      sh.type = need_attr(c, "type");
      const XNode* bsdf = c.child("bsdf");
      if (!bsdf) scene_error("a shape without <bsdf>");
      sh.bsdf = need_attr(*bsdf, "type");
      if (sh.bsdf == "dielectric") {
        for (const XNode& b : bsdf->kids) {
          if (b.name != "float") continue;
          const std::string& nm = need_attr(b, "name");
          if (nm == "intIOR") sh.int_ior = to_float(need_attr(b, "value"));
          if (nm == "extIOR") sh.ext_ior = to_float(need_attr(b, "value"));
        }
      } else if (sh.bsdf != "diffuse" && sh.bsdf != "roughconductor") {
        scene_error("bsdf type " + sh.bsdf);
      }
      if (sh.type == "ply") {
        for (const XNode& s2 : c.kids)
          if (s2.name == "string" && need_attr(s2, "name") == "filename") {
            SceneShapeDesc one = sh;
            one.filename = resolve(out.base_dir, need_attr(s2, "value"));
            out.shapes.push_back(one);
          }
      } else if (sh.type == "sphere") {
        const XNode* pt = c.child("point");
        const XNode* rad = c.child("float");
        if (!pt || !rad) scene_error("a sphere without <point> or <float>");
        sh.center = Vector3f(to_float(need_attr(*pt, "x")), to_float(need_attr(*pt, "y")),
                             to_float(need_attr(*pt, "z")));
        sh.radius = to_float(need_attr(*rad, "value"));
        out.shapes.push_back(sh);
      }
    } else if (c.name == "emitter") {
      if (need_attr(c, "type") != "envmap") scene_error("emitter type " + need_attr(c, "type"));
      for (const XNode& s2 : c.kids)
        if (s2.name == "string" && need_attr(s2, "name") == "filename")
          out.envmaps.push_back(resolve(out.base_dir, need_attr(s2, "value")));
    }
  }
  return out;
}

// -------------------------------------------------------------- Render ----
HIPRender::HIPRender(Scene* scene, const GPUVolLights& lights) : m_scene(scene), m_lights(lights) {}

void HIPRender::init() {
  m_init = false;
  m_mode = Mode::kNone;
  if (!m_scene || !m_scene->m_film) throw gpu::Error(FLG_ERR_INVALID, "HIPRender: no scene or no film");
  if (m_scene->m_integrator_type != EIntegratorType::EVolPathIntegrator)
    throw gpu::Error(FLG_ERR_INVALID, "HIPRender: only the volume integrators are available");
  m_volume = dynamic_cast<const GridVolume*>(m_scene->m_volume);
  if (!m_volume) throw gpu::Error(FLG_ERR_INVALID, "HIPRender: scene.m_volume is not a GridVolume");
  if (m_scene->m_primitives.empty()) {
    m_mode = Mode::kVolume;
  } else if (m_scene->m_primitives.size() == 1 &&
             (m_mesh = dynamic_cast<const LBVHAccel*>(m_scene->m_primitives[0])) != nullptr) {
    m_mode = Mode::kMeshMedium;
  } else {
    throw gpu::Error(FLG_ERR_INVALID, "HIPRender: primitives other than one interface mesh");
  }
  m_init = true;
}

bool HIPRender::render() {
  if (!m_init) return false;
  Scene& sc = *m_scene;
  if (m_mode == Mode::kVolume) {
    GPUVolIntegrator integ(sc.m_maxDepth, m_lights);
    integ.render(sc);
    m_lookups = integ.lastLookups();
    return true;
  }
  gpu::Context& ctx = m_mesh->context();
  flg_vol_scene vs;
  std::memset(&vs, 0, sizeof(vs));
  for (int k = 0; k < 3; ++k) {
    vs.camera.pos[k] = sc.m_origin[k];
    vs.camera.look_at[k] = sc.m_target[k];
    vs.camera.up[k] = sc.m_up[k];
    vs.point_pos[k] = m_lights.point_pos[k];
    vs.point_intensity[k] = m_lights.point_intensity[k];
    vs.env_radiance[k] = m_lights.env_radiance[k];
  }
  vs.camera.fov_deg = sc.m_FoV;
  vs.has_point = m_lights.has_point ? 1 : 0;
  vs.has_env = m_lights.has_env ? 1 : 0;
  flg_render_cfg cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.width = sc.m_resX;
  cfg.height = sc.m_resY;
  cfg.spp = sc.m_SPP;
  cfg.max_depth = sc.m_maxDepth;
  cfg.film = 1;  // the reference's render loop and film, as CPURender's integrators use
  const size_t n = size_t(sc.m_resX) * sc.m_resY * 3;
  gpu::DeviceBuffer d(ctx, n * sizeof(float));
  uint64_t stats[5] = {0, 0, 0, 0, 0};
  ctx.check(flg_volpath_mesh_render(ctx.get(), m_mesh->handle(), m_volume->handle(), &vs, &cfg,
                                    static_cast<float*>(d.data()), stats),
            "flg_volpath_mesh_render");
  m_lookups = stats[1];
  std::vector<float> rgb(n);
  d.download(rgb.data(), n * sizeof(float));
  for (int y = 0; y < sc.m_resY; ++y)
    for (int x = 0; x < sc.m_resX; ++x) {
      const float* c = &rgb[(size_t(y) * sc.m_resX + x) * 3];
      sc.m_film->getBuffer(x, y, EFilmBufferType::EColor) = Vector3f(c[0], c[1], c[2]);
    }
  return true;
}

bool HIPRender::saveImage(const std::string& name, bool denoise) {
  if (!m_init || denoise) return false;
  return m_scene->m_film->saveBuffer(name, EFilmBufferType::EColor);
}

}  // namespace fledge

