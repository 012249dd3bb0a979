// TEST INFRASTRUCTURE — flat C entry points of the CPU oracle for the Python
// tests (ctypes) and bench.py's cpu_baseline leg.  Not part of the product.
#include <algorithm>
#include <atomic>
#include <thread>

#include "oracle_impl.hpp"

namespace orc {
int diffuse_path(const Lbvh& b, bool compat, Halton& hs, V3 o, V3 d, int max_depth, int32_t* ids,
                 float* ts, TraceCounters* ctr);
void diffuse_render(const Lbvh& b, bool compat, const Camera& cam, int W, int H, int spp,
                    int max_depth, int64_t p0, int64_t p1, int32_t* nseg, int32_t* ids, float* ts,
                    int nthreads, uint64_t* rays_out);
V3 offset_ray_origin(V3 p, V3 n, V3 dir);
void diffuse_sample(V3 ng, V3 wo, float u0, float u1, V3* wi, float* pdf);
void shading_frame(V3 n, V3* t, V3* b, V3* nn);
}  // namespace orc

using namespace orc;

namespace {
struct Handle {
  Lbvh b;
  std::vector<float> verts;
  std::vector<int32_t> ind;
};
template <class F>
void parallel_for(int64_t n, int nthreads, F f) {
  if (nthreads < 1) nthreads = 1;
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      while (true) {
        int64_t b0 = next.fetch_add(256);
        if (b0 >= n) break;
        int64_t e = std::min(n, b0 + 256);
        for (int64_t i = b0; i < e; ++i) f(i, t);
      }
    });
  for (auto& x : th) x.join();
}
Camera make_camera(const float* c) {  // pos3, look_at3, up3, fov
  return Camera(v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]), v3(c[6], c[7], c[8]), c[9]);
}
}  // namespace

// Grid and medium of a VolScene: OpenVDBVolume ctor volume.cpp:33-59
// (sigma_t = sigma_a + sigma_s, max density from the grid, inverse in double).
void orc::init_vol_grid(VolScene& sc, const float* grid, int32_t nx, int32_t ny, int32_t nz,
                   const float* bounds6, const float* medium3) {
  sc.grid.nx = nx; sc.grid.ny = ny; sc.grid.nz = nz;
  sc.grid.data = grid;
  sc.grid.bmin = v3(bounds6[0], bounds6[1], bounds6[2]);
  sc.grid.bmax = v3(bounds6[3], bounds6[4], bounds6[5]);
  sc.grid.scale = v3((float)nx / (sc.grid.bmax.x - sc.grid.bmin.x),
                     (float)ny / (sc.grid.bmax.y - sc.grid.bmin.y),
                     (float)nz / (sc.grid.bmax.z - sc.grid.bmin.z));
  sc.med.sigma_s = medium3[0];
  sc.med.sigma_a = medium3[1];
  sc.med.g = medium3[2];
  sc.med.sigma_t = sc.med.sigma_a + sc.med.sigma_s;
  float mx = -FLT_MAX;
  for (int64_t i = 0; i < (int64_t)nx * ny * nz; ++i) mx = smax(mx, grid[i]);
  sc.med.max_density = mx;
  sc.med.inv_max_density = (float)(1.0 / (double)mx);
  sc.med.max_mu = sc.med.sigma_t * mx;
}

extern "C" {

// ---------------------------------------------------------------- math ----
void orc_logf(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = flg_logf(x[i]);
}
void orc_sincosf(const float* x, float* s, float* c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) flg_sincosf(x[i], &s[i], &c[i]);
}
// draws of the ratio-tracking walks: x = (path, walk, n) triples
void orc_walk_draw(const uint32_t* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = walk_draw(walk_key(x[3 * i], x[3 * i + 1]), x[3 * i + 2]);
}
void orc_pcg32(uint64_t initstate, uint64_t initseq, int64_t n, uint32_t* out) {
  Pcg32 r;
  r.seed(initstate, initseq);
  for (int64_t i = 0; i < n; ++i) out[i] = r.next();
}
// Triangle soup: per triangle 3 centroid draws then 9 vertex offsets
// u * 0.02 - 0.01, from PCG32(initstate, initseq); vertex i*3+k, index = identity.
void orc_gen_soup(uint64_t initstate, uint64_t initseq, int32_t ntris, float* verts, int32_t* ind) {
  Pcg32 r;
  r.seed(initstate, initseq);
  for (int32_t t = 0; t < ntris; ++t) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = r.uniform();
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) verts[(3 * (int64_t)t + v) * 3 + k] = c[k] + (r.uniform() * 0.02f - 0.01f);
    for (int v = 0; v < 3; ++v) ind[3 * (int64_t)t + v] = 3 * t + v;
  }
}

// ---------------------------------------------------------------- Morton ---
uint32_t orc_left_shift3(uint32_t x) { return left_shift3(x); }
uint32_t orc_encode_morton3(float x, float y, float z) { return encode_morton3(v3(x, y, z)); }

// ------------------------------------------------------------------ LBVH ---
void* orc_lbvh_build(const float* verts, const int32_t* ind, int32_t ntris, int32_t nverts) {
  Handle* h = new Handle;
  h->verts.assign(verts, verts + 3 * (int64_t)nverts);
  h->ind.assign(ind, ind + 3 * (int64_t)ntris);
  lbvh_build(h->b, h->verts.data(), h->ind.data(), ntris, nverts);
  return h;
}
void orc_lbvh_free(void* p) { delete (Handle*)p; }
int32_t orc_lbvh_ntris(void* p) { return ((Handle*)p)->b.n; }

// Any output pointer may be null.  Node arrays have n-1 entries, leaf arrays n.
void orc_lbvh_export(void* p, float* gbound6, uint32_t* morton_orig, uint32_t* morton_sorted,
                     uint32_t* order, int32_t* split, int32_t* rlo, int32_t* rhi, int32_t* parent,
                     uint8_t* lleaf, uint8_t* rleaf, float* nbound6, int32_t* leaf_parent) {
  const Lbvh& b = ((Handle*)p)->b;
  const int n = b.n, ni = std::max(n - 1, 0);
  if (gbound6) {
    gbound6[0] = b.glo.x; gbound6[1] = b.glo.y; gbound6[2] = b.glo.z;
    gbound6[3] = b.ghi.x; gbound6[4] = b.ghi.y; gbound6[5] = b.ghi.z;
  }
  for (int i = 0; i < n; ++i) {
    if (morton_orig) morton_orig[i] = b.morton_orig[i];
    if (morton_sorted) morton_sorted[i] = b.morton[i];
    if (order) order[i] = b.order[i];
    if (leaf_parent) leaf_parent[i] = b.leaf_parent[i];
  }
  for (int i = 0; i < ni; ++i) {
    if (split) split[i] = b.split[i];
    if (rlo) rlo[i] = b.rlo[i];
    if (rhi) rhi[i] = b.rhi[i];
    if (parent) parent[i] = b.parent[i];
    if (lleaf) lleaf[i] = b.lleaf[i];
    if (rleaf) rleaf[i] = b.rleaf[i];
    if (nbound6) {
      float* q = nbound6 + 6 * (int64_t)i;
      q[0] = b.nlo[i].x; q[1] = b.nlo[i].y; q[2] = b.nlo[i].z;
      q[3] = b.nhi[i].x; q[4] = b.nhi[i].y; q[5] = b.nhi[i].z;
    }
  }
}

// rays: 8 floats per ray (ox, oy, oz, tmax, dx, dy, dz, unused)
void orc_trace(void* p, const float* rays, int64_t n, int32_t compat, int32_t* ids, float* ts,
               float* ngs, uint64_t* nodes, uint64_t* tris, int nthreads) {
  const Lbvh& b = ((Handle*)p)->b;
  std::vector<TraceCounters> ctr(std::max(nthreads, 1));
  parallel_for(n, nthreads, [&](int64_t i, int t) {
    const float* r = rays + 8 * i;
    float th;
    V3 ng;
    ids[i] = lbvh_trace(b, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], compat != 0, &th, &ng,
                        &ctr[t]);
    if (ts) ts[i] = th;
    if (ngs) { ngs[3 * i] = ng.x; ngs[3 * i + 1] = ng.y; ngs[3 * i + 2] = ng.z; }
  });
  uint64_t a = 0, c = 0;
  for (auto& x : ctr) { a += x.nodes; c += x.tris; }
  if (nodes) *nodes = a;
  if (tris) *tris = c;
}

void orc_brute_trace(const float* verts, const int32_t* ind, int32_t ntris, const float* rays,
                     int64_t n, int32_t compat, int32_t* ids, float* ts, int32_t* nties,
                     int nthreads) {
  parallel_for(n, nthreads, [&](int64_t i, int) {
    const float* r = rays + 8 * i;
    int nt = 0;
    ids[i] = brute_trace(verts, ind, ntris, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3],
                         compat != 0, ts ? &ts[i] : nullptr, &nt);
    if (nties) nties[i] = nt;
  });
}

// --------------------------------------------------------------- sampler ---
int32_t orc_num_primes() { return (int32_t)prime_list().size(); }
int32_t orc_prime(int32_t i) { return prime_list()[i]; }
float orc_radical_inverse(int32_t base, uint64_t a) { return radical_inverse_(base, a); }
uint64_t orc_halton_offset(int32_t resx, int32_t resy, int32_t px, int32_t py) {
  Halton h(1, resx, resy);
  return h.offset_for_pixel(px, py);
}
// n consecutive get1D() values of sample `sample` of pixel (px, py)
void orc_halton_get(int32_t resx, int32_t resy, int32_t spp, int32_t px, int32_t py,
                    int64_t sample, int32_t n, float* out) {
  Halton h((uint64_t)spp, resx, resy);
  h.start(px, py, (uint64_t)sample);
  for (int i = 0; i < n; ++i) out[i] = h.get1D();
}
// get1D() values of the reference's sequential sampler use (setPixel, then
// reset() after each pixel sample): dims per sample, samples per pixel.
void orc_halton_sequence(int32_t resx, int32_t resy, int32_t spp, int32_t px, int32_t py,
                         int32_t dims, float* out) {
  Halton h((uint64_t)spp, resx, resy);
  for (int s = 0; s < spp; ++s) {
    h.start(px, py, (uint64_t)s);
    for (int d = 0; d < dims; ++d) out[(int64_t)s * dims + d] = h.get1D();
  }
}

// ---------------------------------------------------------------- camera ---
// Primary rays (1 spp jittered by Halton dims 0,1) or sample s of spp.
void orc_camera_rays(const float* cam10, int32_t W, int32_t H, int32_t spp, int32_t sample,
                     float* rays, int nthreads) {
  Camera cam = make_camera(cam10);
  parallel_for((int64_t)W * H, nthreads, [&](int64_t pix, int) {
    Halton h((uint64_t)spp, W, H);
    int x = (int)(pix % W), y = (int)(pix / W);
    h.start(x, y, (uint64_t)sample);
    float ux = h.get1D();
    float uy = h.get1D();
    V3 o, d;
    cam.generate((float)x + ux, (float)y + uy, W, H, &o, &d);
    float* r = rays + 8 * pix;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = INFINITY;
    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 0.0f;
  });
}

// ---------------------------------------------------------------- volume ---
// medium: sigma_s, sigma_a, g ; lights11: point_present, px, py, pz,
// intensity (RGB), env_present, env_radiance (RGB) ; cam10 as above.
static void set_lights(VolScene& sc, const float* lights11) {
  sc.point.present = lights11[0] != 0.0f;
  sc.point.pos = v3(lights11[1], lights11[2], lights11[3]);
  sc.point.intensity = v3(lights11[4], lights11[5], lights11[6]);
  if (lights11[7] != 0.0f) sc.env.build(v3(lights11[8], lights11[9], lights11[10]), sc.grid.bmin, sc.grid.bmax);
}

void orc_volpath_render(const float* grid, int32_t nx, int32_t ny, int32_t nz, const float* bounds6,
                        const float* medium3, const float* lights11, const float* cam10, int32_t W,
                        int32_t H, int32_t spp, int32_t max_depth, float* rgb, int nthreads,
                        uint64_t* lookups, uint64_t* paths) {
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.cam = &cam;
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  svol_render(sc, rgb, nthreads, lookups, paths);
}

// orc_volpath_render over the pixels 0, stride, 2 stride, ... only (the CPU
// baseline's bounded sample); the other pixels of rgb are left untouched.
void orc_volpath_render_strided(const float* grid, int32_t nx, int32_t ny, int32_t nz, const float* bounds6,
                                const float* medium3, const float* lights11, const float* cam10, int32_t W,
                                int32_t H, int32_t spp, int32_t max_depth, int64_t stride, float* rgb,
                                int nthreads, uint64_t* lookups, uint64_t* paths) {
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.cam = &cam;
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  sc.pix_stride = stride < 1 ? 1 : stride;
  svol_render(sc, rgb, nthreads, lookups, paths);
}

// The reference's render loop and film (film.cpp); arguments as
// orc_volpath_render.  task_L / task_uv (may be NULL): per pixel the spp - 1
// splatted samples (RGB radiance; film position x, y).
void orc_volpath_render_film(const float* grid, int32_t nx, int32_t ny, int32_t nz, const float* bounds6,
                             const float* medium3, const float* lights11, const float* cam10, int32_t W,
                             int32_t H, int32_t spp, int32_t max_depth, float* rgb, float* task_L,
                             float* task_uv, uint64_t* lookups, uint64_t* paths) {
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.cam = &cam;
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  svol_render_film(sc, rgb, task_L, task_uv, lookups, paths);
}

// Per-sample RGB radiance and film position of the per-sample schedule
// (sample s of pixel p on Halton index offset(p) + s * stride, dimensions 0
// and 1 for the film position): L[3*W*H*spp], uv[2*W*H*spp], pixel-major.
void orc_volpath_samples(const float* grid, int32_t nx, int32_t ny, int32_t nz, const float* bounds6,
                         const float* medium3, const float* lights11, const float* cam10, int32_t W, int32_t H,
                         int32_t spp, int32_t max_depth, float* L, float* uv) {
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  Halton hs((uint64_t)spp, W, H);
  uint64_t lk = 0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      for (int s = 0; s < spp; ++s) {
        hs.start(x, y, (uint64_t)s);
        const float ux = (float)x + hs.get1D();
        const float uy = (float)y + hs.get1D();
        V3 o, d;
        cam.generate(ux, uy, W, H, &o, &d);
        const size_t k = ((size_t)y * W + x) * (size_t)spp + (size_t)s;
        const V3 Lk = svol_Li(sc, hs, o, d, &lk);
        L[3 * k] = Lk.x;
        L[3 * k + 1] = Lk.y;
        L[3 * k + 2] = Lk.z;
        uv[2 * k] = ux;
        uv[2 * k + 1] = uy;
      }
}

float orc_mitchell1d(float d) { return film_mitchell1d(d); }

// HaltonSampler driven as the reference's own sampler test does
// (tests/sampler_tests.cpp:24-45): setPixel(px, py), then per sample `ndraw`
// get1D() and reset().  out: spp * ndraw floats.
void orc_halton_stateful(int32_t spp, int32_t resx, int32_t resy, int32_t px, int32_t py, int32_t ndraw,
                         float* out) {
  Halton hs((uint64_t)spp, resx, resy);
  hs.set_pixel(px, py);
  for (int s = 0; s < spp; ++s) {
    for (int i = 0; i < ndraw; ++i) out[(size_t)s * ndraw + i] = hs.get1D();
    hs.reset();
  }
}

// VolPathIntegrator over the closed interface mesh of handle p with the
// density grid inside, on the mesh's bounding box; medium, lights and camera
// as orc_volpath_render.
// film = 1: the reference's render loop and film (single-threaded)
void orc_volpath_mesh_render(void* p, int32_t compat, const float* grid, int32_t nx, int32_t ny, int32_t nz,
                             const float* medium3, const float* lights11, const float* cam10, int32_t W, int32_t H,
                             int32_t spp, int32_t max_depth, float* rgb, int nthreads, uint64_t* lookups,
                             int32_t film) {
  const Lbvh& b = ((Handle*)p)->b;
  const float bounds6[6] = {b.glo.x, b.glo.y, b.glo.z, b.ghi.x, b.ghi.y, b.ghi.z};
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.cam = &cam;
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  if (film)
    volpath_mesh_render_film(b, compat != 0, sc, rgb, lookups);
  else
    volpath_mesh_render(b, compat != 0, sc, rgb, nthreads, lookups);
}

// orc_volpath_mesh_render (per-sample schedule) on every stride-th pixel
// only: a bounded sample of a frame for timing; paths_out = path samples run.
void orc_volpath_mesh_render_strided(void* p, int32_t compat, const float* grid, int32_t nx, int32_t ny,
                                     int32_t nz, const float* medium3, const float* lights11, const float* cam10,
                                     int32_t W, int32_t H, int32_t spp, int32_t max_depth, int64_t stride,
                                     float* rgb, int nthreads, uint64_t* lookups, uint64_t* paths_out) {
  const Lbvh& b = ((Handle*)p)->b;
  const float bounds6[6] = {b.glo.x, b.glo.y, b.glo.z, b.ghi.x, b.ghi.y, b.ghi.z};
  VolScene sc;
  init_vol_grid(sc, grid, nx, ny, nz, bounds6, medium3);
  set_lights(sc, lights11);
  Camera cam = make_camera(cam10);
  sc.cam = &cam;
  sc.width = W; sc.height = H; sc.spp = spp; sc.max_depth = max_depth;
  sc.pix_stride = stride < 1 ? 1 : stride;
  volpath_mesh_render(b, compat != 0, sc, rgb, nthreads, lookups);
  const int64_t npix = (int64_t)W * H;
  if (paths_out) *paths_out = (uint64_t)((npix + sc.pix_stride - 1) / sc.pix_stride) * (uint64_t)spp;
}

// Scene::intersect over the meshes of handles[0..n) in scene order.
void orc_scene_trace(void* const* handles, int32_t n, int32_t compat, const float* rays, int64_t nrays,
                     int32_t* mesh, int32_t* ids, float* ts, int nthreads) {
  std::vector<const Lbvh*> meshes(n);
  for (int32_t i = 0; i < n; ++i) meshes[i] = &((Handle*)handles[i])->b;
  scene_trace(meshes, compat != 0, rays, nrays, mesh, ids, ts, nthreads);
}

// Differentiable estimator over this mesh as the interface (arguments as
// orc_diff_render, the medium on the mesh's bounding box).
void orc_diff_mesh_render(void* p, int32_t compat, const double* theta, const float* theta_p, int32_t nx,
                          int32_t ny, int32_t nz, const float* params4, const float* lights11, const float* cam10,
                          int32_t W, int32_t H, int32_t spp, int32_t max_depth, double* lpath, double* img,
                          const double* dldpix, double* grad, double* galpha, int nthreads, uint64_t* lookups,
                          int32_t channel) {
  diff_mesh_render(((Handle*)p)->b, compat != 0, theta, theta_p, nx, ny, nz, params4, lights11, cam10, W, H, spp,
                   max_depth, lpath, img, dldpix, grad, galpha, nthreads, lookups, channel);
}

// ------------------------------------------------------------ path trace ---
void orc_diffuse_render(void* p, int32_t compat, const float* cam10, int32_t W, int32_t H,
                        int32_t spp, int32_t max_depth, int64_t p0, int64_t p1, int32_t* nseg,
                        int32_t* ids, float* ts, int nthreads, uint64_t* rays) {
  const Lbvh& b = ((Handle*)p)->b;
  Camera cam = make_camera(cam10);
  diffuse_render(b, compat != 0, cam, W, H, spp, std::min(max_depth, 63), p0, p1, nseg, ids, ts,
                 nthreads, rays);
}

void orc_offset_ray_origin(const float* p, const float* n, const float* d, float* out) {
  V3 r = offset_ray_origin(v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]), v3(d[0], d[1], d[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// Camera::generateRay (camera.h:21-31) at film position (x, y) of a W x H
// image: origin and unit direction.
void orc_camera_generate(const float* cam10, float x, float y, int32_t W, int32_t H, float* o, float* d) {
  Camera cam = make_camera(cam10);
  V3 ro, rd;
  cam.generate(x, y, W, H, &ro, &rd);
  o[0] = ro.x; o[1] = ro.y; o[2] = ro.z;
  d[0] = rd.x; d[1] = rd.y; d[2] = rd.z;
}

// The diffuse bounce of the path tracer for n samples: normal ng[3n], outgoing
// wo[3n], draws u[2n]; writes wi[3n] and pdf[n].
void orc_diffuse_sample(const float* ng, const float* wo, const float* u, int64_t n, float* wi, float* pdf) {
  for (int64_t i = 0; i < n; ++i) {
    V3 w;
    diffuse_sample(v3(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]), v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]),
                   u[2 * i], u[2 * i + 1], &w, &pdf[i]);
    wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
  }
}

// CoordinateTransition's axes for n normals: tangent, binormal, normal (9 floats each).
void orc_shading_frame(const float* n, int64_t count, float* out9) {
  for (int64_t i = 0; i < count; ++i) {
    V3 t, b, nn;
    shading_frame(v3(n[3 * i], n[3 * i + 1], n[3 * i + 2]), &t, &b, &nn);
    float* q = out9 + 9 * i;
    q[0] = t.x; q[1] = t.y; q[2] = t.z; q[3] = b.x; q[4] = b.y; q[5] = b.z; q[6] = nn.x; q[7] = nn.y; q[8] = nn.z;
  }
}

}  // extern "C"

// ------------------------------------------------------- small pieces ----
extern "C" {
// Karras node emission (radix_bvh.cpp:118-161) on caller-given sorted codes.
void orc_karras(const uint32_t* codes, int32_t n, int32_t* split, int32_t* rlo, int32_t* rhi,
                uint8_t* lleaf, uint8_t* rleaf) {
  // reuse the builder through a synthetic mesh is not possible (codes come
  // from geometry), so run the emission loop directly
  Lbvh b;
  b.n = n;
  b.morton.assign(codes, codes + n);
  b.order.resize(n);
  for (int i = 0; i < n; ++i) b.order[i] = (uint32_t)i;
  karras_emit(b);
  for (int i = 0; i + 1 < n; ++i) {
    split[i] = b.split[i];
    rlo[i] = b.rlo[i];
    rhi[i] = b.rhi[i];
    lleaf[i] = b.lleaf[i];
    rleaf[i] = b.rleaf[i];
  }
}
// Dist1D (distribution.hpp:12-62): builds from f[n] and samples u.
float orc_dist1d_sample(const float* f, int32_t n, float u, float* pdf, int32_t* off) {
  Dist1D d;
  d.init(f, n);
  int o = 0;
  float r = d.sample_c(u, pdf, &o);
  *off = o;
  return r;
}
// HGSampleP / HGP (math_utils.h:103-128): wi for wo, (u, v); returns HGP(wi, wo).
float orc_hg_sample(const float* wo, float u, float v, float g, float* wi) {
  V3 w = hg_sample_pub(v3(wo[0], wo[1], wo[2]), u, v, g);
  wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
  return hg_p_pub(w, v3(wo[0], wo[1], wo[2]), g);
}
float orc_hg_p(const float* wi, const float* wo, float g) {
  return hg_p_pub(v3(wi[0], wi[1], wi[2]), v3(wo[0], wo[1], wo[2]), g);
}
int32_t orc_pluecker(const float* a, const float* b, const float* c, const float* o, const float* d,
                     int32_t compat, float* t, float* ng) {
  V3 n;
  float th = 0;
  bool h = pluecker(v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2]),
                    v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), compat != 0, &th, &n);
  if (h) { *t = th; ng[0] = n.x; ng[1] = n.y; ng[2] = n.z; }
  return h ? 1 : 0;
}
}

// ------------------------------------------------ reference test rays ----
#include <random>
extern "C" {
// bvh_test.hpp:107-141 correctness1 rays: RandomCPU{} (mt19937 seed 0 with
// uniform_real_distribution<float>), origin (u,u,u)/2, direction
// Normalize(u,u,u); tnear 0, tfar FLT_MAX.  8 floats per ray.
void orc_bvh_test_rays(int64_t n, float* rays) {
  std::mt19937 gen(0u);
  std::uniform_real_distribution<float> dist(0.0f, 1.0f);
  for (int64_t i = 0; i < n; ++i) {
    float ox = dist(gen), oy = dist(gen), oz = dist(gen);
    float dx = dist(gen), dy = dist(gen), dz = dist(gen);
    V3 o = v3(ox, oy, oz) / 2.0f;
    V3 d = normalize(v3(dx, dy, dz));
    float* r = rays + 8 * i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = FLT_MAX;
    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 0.0f;
  }
}
// shape_tests.cpp:22-48 Watertight rays: RandomCPU(114514); origin
// 0.5 * UniformSampleSphere(u), direction UniformSampleSphere(u')
// (math_utils.h:95-101, evaluated in float).
void orc_watertight_rays(int64_t n, float* rays) {
  std::mt19937 gen(114514u);
  std::uniform_real_distribution<float> dist(0.0f, 1.0f);
  auto sphere = [](float ux, float uy) {
    float theta = std::acos(1 - 2 * uy);
    float phi = 2 * PI * ux;
    float st = std::sin(theta), ct = std::cos(theta), sp = std::sin(phi), cp = std::cos(phi);
    return v3(st * cp, st * sp, ct);
  };
  for (int64_t i = 0; i < n; ++i) {
    float a = dist(gen), b = dist(gen);
    V3 p = v3(0, 0, 0) + sphere(a, b) * 0.5f;
    float c = dist(gen), d = dist(gen);
    V3 dir = sphere(c, d);
    float* r = rays + 8 * i;
    r[0] = p.x; r[1] = p.y; r[2] = p.z; r[3] = INFINITY;
    r[4] = dir.x; r[5] = dir.y; r[6] = dir.z; r[7] = 0.0f;
  }
}
}
