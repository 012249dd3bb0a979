// TEST INFRASTRUCTURE — internal declarations of the CPU oracle (see oracle.hpp).
#pragma once
#include <functional>
#include <vector>

#include "oracle.hpp"

namespace orc {

// ------------------------------------------------------------------ LBVH ---
struct TraceCounters {
  uint64_t nodes = 0, tris = 0;
};

struct Lbvh {
  int32_t n = 0;
  const float* verts = nullptr;
  const int32_t* ind = nullptr;
  V3 glo{}, ghi{};
  std::vector<uint32_t> morton_orig, morton, order;
  std::vector<int32_t> split, dir, rlo, rhi, parent, leaf_parent;
  std::vector<uint8_t> lleaf, rleaf;
  std::vector<V3> nlo, nhi;
};

uint32_t left_shift3(uint32_t x);
uint32_t encode_morton3(V3 q);
void lbvh_build(Lbvh& b, const float* verts, const int32_t* ind, int32_t ntris, int32_t nverts);
void karras_emit(Lbvh& b);  // node emission over b.morton (sorted), b.n
bool bound_intersect(V3 lo, V3 hi, V3 o, V3 d, float& tnear);
bool pluecker(V3 a, V3 b, V3 c, V3 O, V3 D, bool compat, float* thit, V3* ng);
int32_t lbvh_trace(const Lbvh& b, V3 o, V3 d, float tmax, bool compat, float* t, V3* ng,
                   TraceCounters* ctr);
int32_t brute_trace(const float* verts, const int32_t* ind, int32_t ntris, V3 o, V3 d, float tmax,
                    bool compat, float* t, int* nties);

// ---------------------------------------------------------------- Halton ---
const std::vector<int>& prime_list();  // sampler.h:39-57
float radical_inverse_(int base, uint64_t a);  // sampler.h:59-72
uint64_t mul_mod_inverse(int64_t a, int64_t b);  // sampler.h:77-96

struct Halton {  // HaltonSampler sampler.h:143-252
  uint64_t spp = 1;
  int base_scales[2] = {1, 1}, base_exp[2] = {0, 0};
  uint64_t stride = 1;
  int64_t mul_inverse[2] = {0, 0};
  int px = 0, py = 0;
  uint64_t offset = 0;
  uint64_t index = 0;
  uint64_t cnt = 0;
  uint64_t psi = 0;  // m_pixel_sample_index
  bool overflow = false;
  Halton(uint64_t spp_, int resx, int resy);
  uint64_t offset_for_pixel(int x, int y) const;
  void start(int x, int y, uint64_t sample);  // setPixel + getIndexForSample(sample)
  void set_pixel(int x, int y);                // setPixel (sampler.h:164-169)
  bool reset();                                // reset (sampler.h:175-181)
  float sample_dimension(uint64_t idx, uint64_t dim);
  float get1D();
};

// ---------------------------------------------------------------- camera ---
struct Camera {  // src/common/camera.h:11-31
  V3 pos, forward, up, right, up2;
  float fov_ratio;
  Camera(V3 pos_, V3 look_at, V3 up_, float fov);
  void generate(float x, float y, int w, int h, V3* o, V3* d) const;
};

// ---------------------------------------------------------------- volume ---
struct Grid {
  int nx = 0, ny = 0, nz = 0;
  const float* data = nullptr;  // x fastest
  V3 bmin{}, bmax{};
  V3 scale{};  // res / extent
  float voxel(int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return 0.0f;
    return data[((int64_t)k * ny + j) * nx + i];
  }
};
struct Lookup {  // one trilinear lookup: base voxel and weights
  int i, j, k;
  float fx, fy, fz;
  float rho;
};
Lookup grid_lookup(const Grid& g, V3 p);

struct Medium {
  float sigma_s, sigma_a, sigma_t, g;
  float max_density, inv_max_density, max_mu;
};

struct Dist1D {  // src/distribution.hpp:12-62
  std::vector<float> f, cdf;
  float integral = 0;
  void init(const float* fv, int n);
  int inverse_cdf(float u) const;
  float sample_c(float u, float* pdf, int* off) const;
};
struct EnvLight {  // InfiniteAreaLight with a ConstTexture, light.cpp:93-159
  bool present = false;
  V3 radiance{};  // the texture's colour, RGB
  V3 center{};
  float radius = 0;
  std::vector<Dist1D> rows;
  Dist1D marg;
  void build(V3 radiance, V3 scene_lo, V3 scene_hi);
};
struct PointLight {
  bool present = false;
  V3 pos{};
  V3 intensity{};  // RGB radiant intensity: Li = I / d^2 per channel
};

struct VolScene {
  Grid grid;
  Medium med;
  EnvLight env;
  PointLight point;
  Camera* cam = nullptr;
  int width = 0, height = 0, spp = 1, max_depth = 1;
  // pixels rendered: 0, pix_stride, 2 pix_stride, ... (a bounded, evenly
  // spread sample of the image for timing; 1 = every pixel)
  int64_t pix_stride = 1;
};

bool box_intersect_pbrt(V3 lo, V3 hi, V3 o, V3 d, float tmax, float* t_min, float* t_max);
bool box_intersect(V3 lo, V3 hi, V3 o, V3 d, float* t_min, float* t_max);
float hg_p(V3 wi, V3 wo, float g);
V3 hg_sample(V3 wo, float u, float v, float g);
V3 env_sample_li(const EnvLight& env, V3 ref, float u0, float u1, V3* wi, float* pdf, V3* p_light);
V3 hg_sample_pub(V3 wo, float u, float v, float g);
float hg_p_pub(V3 wi, V3 wo, float g);
V3 svol_Li(const VolScene& sc, Halton& hs, V3 o, V3 d, uint64_t* lookups);
void init_vol_grid(VolScene& sc, const float* grid, int32_t nx, int32_t ny, int32_t nz,
                   const float* bounds6, const float* medium3);
// OpenVDBVolume::tr / ::sample for one ray, drawing from hs
float vol_tr_pub(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax);
float vol_sample_pub(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, bool* success,
                     V3* pos);
void svol_render(const VolScene& sc, float* rgb, int nthreads, uint64_t* lookups, uint64_t* paths);
// the trackers with the lookup count
float vol_tr_lk(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, uint64_t* lookups);
float vol_sample_lk(const VolScene& sc, Halton& hs, V3 o, V3 d, float ray_tmax, bool* success, V3* pos,
                    uint64_t* lookups);
V3 offset_ray_origin(V3 p, V3 n, V3 dir);  // OffsetRayOrigin (pathtrace.cpp)
// VolPathIntegrator over a closed interface mesh (volpath_mesh.cpp)
void volpath_mesh_render(const Lbvh& b, bool compat, const VolScene& sc, float* rgb, int nthreads,
                         uint64_t* lookups);
// The same integrator through the reference's render loop and film
// (film.cpp), single-threaded.
void volpath_mesh_render_film(const Lbvh& b, bool compat, const VolScene& sc, float* rgb, uint64_t* lookups);

// Closest hit over a scene of meshes (scene.cpp): NaiveBVHAccel over the
// meshes in scene order.  rays: n x 8 floats (o, tmax, d, pad).
void scene_trace(const std::vector<const Lbvh*>& meshes, bool compat, const float* rays, int64_t n, int32_t* mesh,
                 int32_t* ids, float* ts, int nthreads);

// The differentiable estimator over the interface mesh b (volume_diff.cpp):
// forward (lpath, img) or, with dldpix, the adjoint (grad, galpha), of one
// colour channel of the lights.
void diff_mesh_render(const Lbvh& b, bool compat, const double* theta, const float* theta_p, int32_t nx,
                      int32_t ny, int32_t nz, const float* params4, const float* lights11, const float* cam10,
                      int32_t W, int32_t H, int32_t spp, int32_t max_depth, double* lpath, double* img,
                      const double* dldpix, double* grad, double* galpha, int nthreads, uint64_t* lookups,
                      int32_t channel);
// the reference's render loop and film (film.cpp)
float film_mitchell1d(float d);
void film_render(const VolScene& sc, const std::function<V3(Halton&, V3, V3, uint64_t*)>& Li, float* rgb,
                 float* task_L, float* task_uv, uint64_t* lookups_out, uint64_t* paths_out);
void svol_render_film(const VolScene& sc, float* rgb, float* task_L, float* task_uv, uint64_t* lookups,
                      uint64_t* paths);

}  // namespace orc
