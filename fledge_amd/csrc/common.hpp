// Shared device/host helpers of the fledge HIP library (gfx950).
//
// Floating point: the library is compiled with -ffp-contract=off, so every
// expression below rounds each operation exactly as written; the only fused
// operation on the ray path is the explicit __builtin_fmaf in msub
// (src/experimental/intersector.hpp:41-43, embree's fmsub).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/fledge_hip.h"

#define FLG_HD __host__ __device__ __forceinline__
#define FLG_D __device__ __forceinline__

namespace flg {

// ---------------------------------------------------------------- scalar ---
FLG_HD float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
FLG_HD float smax(float a, float b) { return (a < b) ? b : a; }  // std::max

struct V3 {
  float x, y, z;
};
FLG_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
FLG_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
FLG_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
FLG_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
FLG_HD V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
FLG_HD V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
FLG_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
FLG_HD V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// Dot = (x x' + y y') + z z' (Sum, src/common/vector.h:300-302)
FLG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Cross, src/common/vector.h:399-404
FLG_HD V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
FLG_HD float sqnorm(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
FLG_HD V3 normalize(V3 a) { return a / sqrtf(sqnorm(a)); }  // vector.h:490-493
FLG_HD V3 vmin(V3 a, V3 b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
FLG_HD V3 vmax(V3 a, V3 b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }

FLG_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
FLG_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// fledge constants, src/fledge.h:27-36
constexpr float PI = 3.14159265358979323846f;
constexpr float PI_OVER2 = PI / 2;
constexpr float PI_OVER4 = PI / 4;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float INV_4PI = 0.07957747154594766788f;
constexpr float SHADOW_EPS = 1e-4f;
constexpr float ONE_MINUS_EPSILON = 0.99999994f;  // src/common/sampler.h:15

// ------------------------------------------------------ transcendentals ---
// The polynomial coefficients and evaluation order of flg_logf, flg_sindf and
// flg_cosdf below come from fdlibm's e_logf.c and FreeBSD msun's k_sinf.c and
// k_cosf.c (float conversion by Ian Lance Taylor, Cygnus Support; kernels
// optimised by Bruce D. Evans), distributed under this notice:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
//
// Natural log of a float: fdlibm e_logf.c method (reduce to m in
// [sqrt(2)/2, sqrt(2)), s = f / (2 + f), degree-4 minimax polynomial in s^2),
// plain IEEE float operations in a fixed order; < 1 ulp.
FLG_HD float flg_logf(float x) {
  const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f;
  const float Lg1 = 0.66666662693f, Lg2 = 0.40000972152f, Lg3 = 0.28498786688f,
              Lg4 = 0.24279078841f;
  uint32_t ix = f2u(x);
  int k = 0;
  if (ix < 0x00800000u || (ix >> 31)) {
    if ((ix << 1) == 0) return -__builtin_inff();
    if (ix >> 31) return __builtin_nanf("");
    k -= 25;
    x *= 33554432.0f;
    ix = f2u(x);
  } else if (ix >= 0x7f800000u) {
    return x;
  }
  ix += 0x3f800000u - 0x3f3504f3u;
  k += (int)(ix >> 23) - 0x7f;
  ix = (ix & 0x007fffffu) + 0x3f3504f3u;
  float f = u2f(ix) - 1.0f;
  float s = f / (2.0f + f);
  float z = s * s;
  float w = z * z;
  float t1 = w * (Lg2 + w * Lg4);
  float t2 = z * (Lg1 + w * Lg3);
  float R = t2 + t1;
  float hfsq = 0.5f * f * f;
  float dk = (float)k;
  return s * (hfsq + R) + dk * ln2_lo - hfsq + f + dk * ln2_hi;
}

// Sine and cosine of a float: FreeBSD msun method — double reduction by pi/2
// with a two-part constant, __sindf / __cosdf minimax kernels on
// [-pi/4, pi/4], one rounding to float.  Domain on the path: |x| < 2^20.
FLG_HD double flg_sindf(double x) {
  const double S1 = -0.166666666416265235595, S2 = 0.0083333293858894631756,
               S3 = -0.000198393348360966317347, S4 = 0.0000027183114939898219064;
  double z = x * x, w = z * z, r = S3 + z * S4, s = z * x;
  return (x + s * (S1 + z * S2)) + s * w * r;
}
FLG_HD double flg_cosdf(double x) {
  const double C0 = -0.499999997251031003120, C1 = 0.0416666233237390631894,
               C2 = -0.00138867637746099294692, C3 = 0.0000243904487962774090654;
  double z = x * x, w = z * z, r = C2 + z * C3;
  return ((1.0 + z * C0) + w * C1) + (w * z) * r;
}
FLG_HD void flg_sincosf(float xf, float* s, float* c) {
  const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079631090164184570e+00,
               pio2_1t = 1.58932547735281966916e-08;
  double x = (double)xf;
  double fn = __builtin_rint(x * invpio2);
  int n = (int)fn;
  double y = (x - fn * pio2_1) - fn * pio2_1t;
  double sy = flg_sindf(y), cy = flg_cosdf(y);
  switch (n & 3) {
    case 0: *s = (float)sy; *c = (float)cy; break;
    case 1: *s = (float)cy; *c = (float)(-sy); break;
    case 2: *s = (float)(-sy); *c = (float)(-cy); break;
    default: *s = (float)(-cy); *c = (float)sy; break;
  }
}

// ------------------------------------------------- counter-based draws ---
// Integer hash "lowbias32" (C. Wellons, hash-prospector): xorshift-multiply
// rounds, a bijection on 32 bits.
FLG_HD uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
// Stream of one transmittance walk of the differentiable estimator: draw n of
// walk w of path p is hash(key(p, w) + n), so a walk neither consumes nor
// shifts the path's Halton dimensions and can stop at any step.
FLG_HD uint32_t walk_key(uint32_t path, uint32_t walk) { return lowbias32(lowbias32(path) + walk); }
// 1 - u for u uniform on {0, 2^-24, ..., 1 - 2^-24}: an exact float in (0, 1].
FLG_HD float walk_draw(uint32_t key, uint32_t n) {
  return (float)((lowbias32(key + n) >> 8) + 1u) * 0x1p-24f;
}

// ----------------------------------------------------------- host side ----
struct KernelTimer {
  double ms = 0;
  int64_t launches = 0;
};

struct PendingTimer {
  const char* name;
  hipEvent_t a, b;
};

}  // namespace flg

struct flg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  bool profiling = false;
  std::map<std::string, flg::KernelTimer> timers;
  std::vector<flg::PendingTimer> pending;
  // grow-only scratch buffers, keyed by purpose
  std::map<std::string, std::pair<void*, size_t>> scratch;
  int num_cus = 256;
  // launch shapes chosen by the autotuner (0: not tuned)
  int tuned_trace_block = 0, tuned_trace_stack = 0, tuned_trace_ring = 0, tuned_pt_block = 0, tuned_pt_bpc = 0,
      tuned_pt_stack = 0, tuned_pt_ring = 0, tuned_pt_fetch = 0;
  // captured bounce loops, keyed by the bytes of their launch arguments
  hipStream_t capture_stream = nullptr;
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
  // this device's address of the library's fixed scratch array (the scratch
  // of flg_volpath_render_noalloc), set by flg_ctx_create
  void* noalloc_scratch = nullptr;
  // output blocks of destroyed BVHs, kept for later builds (lbvh.hip)
  std::vector<std::pair<void*, size_t>> bvh_pool;
};

namespace flg {

int set_error(flg_ctx* ctx, int status, const std::string& msg);

// Runs an entry point's body so that no C++ exception crosses the C ABI: a
// failed host allocation (std::bad_alloc) becomes FLG_ERR_OOM and any other
// exception FLG_ERR_INVALID, with the entry point's name in the context's
// last error when there is a context.
template <class F>
int host_guard(flg_ctx* ctx, const char* what, F&& body) noexcept {
  int status = FLG_ERR_INVALID;
  const char* why = "unexpected host error";
  try {
    return body();
  } catch (const std::bad_alloc&) {
    status = FLG_ERR_OOM;
    why = "host allocation failed";
  } catch (...) {
  }
  try {
    return set_error(ctx, status, std::string(what) + ": " + why);
  } catch (...) {  // no memory left for the message either
    return status;
  }
}
void resolve_timers(flg_ctx* ctx);
int check_hip(flg_ctx* ctx, hipError_t e, const char* what);
// Runs enqueue's launches on the context's stream through a hipGraph cached
// in the context under `key` (every kernel argument of the sequence belongs
// in it): captured once on a private stream (the default stream cannot be
// captured), replayed afterwards.  At most 8 graphs are kept.
int graph_launch(flg_ctx* ctx, const std::string& key, const std::function<void(hipStream_t)>& enqueue,
                 const char* what);
void* scratch(flg_ctx* ctx, const char* key, size_t bytes, int* status);

// The fixed per-GPU scratch of flg_volpath_render_noalloc: 192 MiB of RGB
// path radiance (a 512 x 512 frame at 64 spp) plus a page of counters.
constexpr size_t kNoAllocScratchBytes = ((size_t)192 << 20) + 4096;
constexpr int kMaxNoAllocDevices = 64;
// The device address of that array on `device` (which must be current),
// looked up once per device; FLG_ERR_HIP if the lookup fails.
int noalloc_scratch_address(int device, void** out);

// Times a kernel launch with HIP events on the context stream when
// profiling is enabled.
struct ScopedTimer {
  flg_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  ScopedTimer(flg_ctx* c, const char* n);
  ~ScopedTimer();
};

#define FLG_CHECK_HIP(ctx, expr)                              \
  do {                                                        \
    hipError_t e__ = (expr);                                  \
    if (e__ != hipSuccess) return flg::check_hip(ctx, e__, #expr); \
  } while (0)

#define FLG_CHECK_LAUNCH(ctx) FLG_CHECK_HIP(ctx, hipGetLastError())

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Tile order of a w x h image: bands of TH rows cut into TW-wide tiles, left
// to right (the band's last tile narrower and the last band shorter when w or
// h is not a multiple), row-major inside a tile.  tile_pixel maps the k-th
// pixel of that order to (x, y), a bijection on the image.  Kernels that start
// paths in this order keep the paths in flight on a compact patch of the image
// instead of a few scanlines; a path's result depends only on its own index,
// so the order changes which paths run together and not what they compute.
// All in 32 bits: the band is (k / w) / TH, never k / (w * TH), whose divisor
// would wrap for large w (the callers' bound w * h * tasks < 2^31 allows
// w = 2^31 - 1 at h = tasks = 1); band * TH * w <= k after it.
template <uint32_t TW, uint32_t TH>
FLG_HD void tile_pixel(uint32_t k, uint32_t w, uint32_t h, uint32_t& x, uint32_t& y) {
  const uint32_t y0 = k / w / TH * TH;
  const uint32_t o = k - y0 * w;                    // position within the band
  const uint32_t rows = TH < h - y0 ? TH : h - y0;  // the band's height
  const uint32_t x0 = o / (TW * rows) * TW;         // the tile's first column
  const uint32_t in = o - x0 * rows;                // position within the tile
  const uint32_t tw = TW < w - x0 ? TW : w - x0;    // the tile's width
  x = x0 + in % tw;
  y = y0 + in / tw;
}

// Path index of the q-th path in tile order, the tasks of one pixel
// consecutive: a bijection on [0, w * h * tasks).
template <uint32_t TW = 16, uint32_t TH = 16>
FLG_HD uint32_t tile_order(uint32_t q, int32_t tasks, int32_t w, int32_t h) {
  const uint32_t pix = q / (uint32_t)tasks, t = q - pix * (uint32_t)tasks;
  uint32_t x, y;
  tile_pixel<TW, TH>(pix, (uint32_t)w, (uint32_t)h, x, y);
  return (y * (uint32_t)w + x) * (uint32_t)tasks + t;
}

}  // namespace flg
