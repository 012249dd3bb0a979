// TEST INFRASTRUCTURE — CPU restatement of the reference's render loop and
// film, for the volume-only integrator and VolPathIntegrator: ParallelIntegrator::render and
// eval_pixel (src/integrator.cpp:123-235), HaltonSampler::setPixel /
// getPixelSample / reset (src/common/sampler.h:164-186), Film::addSample /
// commitSamples (src/film.hpp:143-174) and the film's filter, MitchellFilter
// constructed with radius 2 and B = C = 1 / 3 in integer arithmetic, i.e. 0
// (src/film.hpp:55, src/common/filter.h:35-57).  Single-threaded, in the
// reference's loop order (tiles of 16 x 16 with the tile column outer, pixels
// column-outer inside a tile, samples in order, splat loops x outer), so the
// film sums in the order one thread of the reference would use; the
// reference's unsynchronised addSample from several threads is not
// reproduced.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "oracle_impl.hpp"

namespace orc {

// MitchellFilter::mitchell1D (filter.h:44-53) at x = p * inv_radius, with
// B = C = 0; the operation order is the reference's.
float film_mitchell1d(float d) {
  const float B = 0.0f, C = 0.0f;
  float x = d * 0.5f;  // p.x() * m_inv_radius.x(), 1 / radius = 0.5
  x = std::fabs(2 * x);
  if (x > 1)
    return ((-B - 6 * C) * x * x * x + (6 * B + 30 * C) * x * x + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) *
           (1.f / 6.f);
  return ((12 - 9 * B - 6 * C) * x * x * x + (-18 + 12 * B + 6 * C) * x * x + (6 - 2 * B)) * (1.f / 6.f);
}

namespace {

struct FilmAcc {
  int W, H;
  std::vector<float> I;
  std::vector<V3> L;
  FilmAcc(int w, int h) : W(w), H(h), I((size_t)w * h, 0.0f), L((size_t)w * h, v3(0, 0, 0)) {}
  // Film::addSample (film.hpp:143-161)
  void add(float ux, float uy, V3 Lv) {
    const float px = ux - 0.5f, py = uy - 0.5f;
    int lx = (int)std::ceil(px - 2.0f), ly = (int)std::ceil(py - 2.0f);
    int hx = (int)std::floor(px + 2.0f) + 1, hy = (int)std::floor(py + 2.0f) + 1;
    lx = lx < 0 ? 0 : lx;
    ly = ly < 0 ? 0 : ly;
    hx = hx > W ? W : hx;
    hy = hy > H ? H : hy;
    for (int x = lx; x < hx; ++x)
      for (int y = ly; y < hy; ++y) {
        const float w = film_mitchell1d(px - (float)x) * film_mitchell1d(py - (float)y);
        const size_t i = (size_t)y * W + x;
        I[i] += w;
        L[i] = L[i] + Lv * w;
      }
  }
  // Film::commitSamples (film.hpp:163-174) into a zero-initialised buffer
  void commit(float* rgb) const {
    for (size_t i = 0; i < I.size(); ++i) {
      const V3 c = I[i] == 0.0f ? v3(0, 0, 0) : L[i] / I[i];
      rgb[3 * i] = c.x;
      rgb[3 * i + 1] = c.y;
      rgb[3 * i + 2] = c.z;
    }
  }
};

}  // namespace

// The render loop over an integrator's Li(sampler, o, d, lookups).  rgb:
// W*H*3.  task_L / task_uv (may be NULL): per pixel (x fastest) the spp - 1
// splatted samples in order, RGB radiance (3 floats) and film position.
void film_render(const VolScene& sc, const std::function<V3(Halton&, V3, V3, uint64_t*)>& Li, float* rgb,
                 float* task_L, float* task_uv, uint64_t* lookups_out, uint64_t* paths_out) {
  const int W = sc.width, H = sc.height, spp = sc.spp;
  const int T = spp - 1;
  FilmAcc film(W, H);
  uint64_t lk = 0;
  const int BS = 16;
  const int bx = (W - 1) / BS + 1, by = (H - 1) / BS + 1;
  for (int i = 0; i < bx; ++i)
    for (int j = 0; j < by; ++j) {
      const int x0 = i * BS, y0 = j * BS;
      const int xm = std::min(x0 + BS, W), ym = std::min(y0 + BS, H);
      Halton hs((uint64_t)spp, W, H);  // one sampler per tile
      for (int x = x0; x < xm; ++x)
        for (int y = y0; y < ym; ++y) {
          hs.set_pixel(x, y);
          V3 o, d;
          sc.cam->generate((float)x + 0.5f, (float)y + 0.5f, W, H, &o, &d);
          (void)Li(hs, o, d, &lk);  // the centre sample: not splatted
          for (int s = 1; s < spp; ++s) {
            const float ux = (float)x + hs.get1D();
            const float uy = (float)y + hs.get1D();
            sc.cam->generate(ux, uy, W, H, &o, &d);
            const V3 L = Li(hs, o, d, &lk);
            film.add(ux, uy, L);
            const size_t k = ((size_t)y * W + x) * (size_t)T + (size_t)(s - 1);
            if (task_L) {
              task_L[3 * k] = L.x;
              task_L[3 * k + 1] = L.y;
              task_L[3 * k + 2] = L.z;
            }
            if (task_uv) {
              task_uv[2 * k] = ux;
              task_uv[2 * k + 1] = uy;
            }
            hs.reset();
          }
        }
    }
  film.commit(rgb);
  if (lookups_out) *lookups_out = lk;
  if (paths_out) *paths_out = (uint64_t)W * H * (uint64_t)spp;
}

void svol_render_film(const VolScene& sc, float* rgb, float* task_L, float* task_uv, uint64_t* lookups_out,
                      uint64_t* paths_out) {
  film_render(
      sc, [&](Halton& hs, V3 o, V3 d, uint64_t* lk) { return svol_Li(sc, hs, o, d, lk); }, rgb, task_L, task_uv,
      lookups_out, paths_out);
}

}  // namespace orc
