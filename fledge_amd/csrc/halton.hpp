// Device Halton sampler: a pure function (pixel, sample, dimension) -> float
// with the results of fledge's HaltonSampler (src/common/sampler.h:143-252)
// under the per-sample schedule "sample s of pixel p uses global index
// offset(p) + s * stride and dimensions 0, 1, 2, ...".
//
// Tables (primes of the linear sieve below 262144, sampler.h:39-57, and their
// float reciprocals 1.0f / base) live in device memory, one copy per context.
// The digit loop of RadicalInverse_ (sampler.h:59-72) divides by the prime in
// double with a refined reciprocal and a small bias, exact for every 32-bit
// index (the host keeps offset + sample * stride below 2^32).
#pragma once
#include "common.hpp"

namespace flg {

struct HaltonTables {
  const int32_t* prime = nullptr;
  const float* inv = nullptr;
  int32_t nprimes = 0;
  int32_t* overflow = nullptr;  // set to 1 when a dimension beyond the table is asked for
  // optional copy of the first nlocal (prime, 1/prime bits) pairs in LDS: the
  // draws of a path hit it instead of the global table's latency
  const int2* local = nullptr;
  int32_t nlocal = 0;
};

// Entries of the tables a kernel stages in LDS (8 KB per block).
constexpr int kLocalPrimes = 1024;

// Per-resolution constants of HaltonSampler's ctor (sampler.h:145-163).
struct HaltonParams {
  int32_t bs0, bs1, exp0, exp1;
  uint32_t stride;
  uint32_t mulinv0, mulinv1;
};

int get_halton_tables(flg_ctx* ctx, HaltonTables* out);
// The context's tables if they exist (built by flg_ctx_create), without
// building or allocating anything; false otherwise.
bool halton_tables_ready(flg_ctx* ctx, HaltonTables* out);
void release_halton_tables(flg_ctx* ctx);
HaltonParams make_halton_params(int resx, int resy);
constexpr uint64_t kMaxHaltonIndex = 1ull << 32;

// The digit loop on the double unit, two digits per trip.  r = 1/b (1 + d)
// comes from the float reciprocal r0 = (1 + e)/b, |e| <= 2^-24, by two Newton
// steps: 1 - b r0 = -e exactly (b r0 has at most 42 significant bits),
// r1 = r0 (1 - e) rounded has |d1| <= 2^-48 + 2^-53, and the second step
// leaves |d| <= 2^-53 plus terms below 2^-95.  For a < 2^32 and a prime
// b < 2^18:
//  * one digit: q = floor(a r + 2^-19) = floor(a / b).  a/b < 2^31 moves by
//    at most 2^31 |d| <= 2^-22 and the sum's rounding by 2^-22, together
//    below the bias, so an exact multiple stays in (q, q + 1); otherwise
//    a/b - q lies in [1/b, 1 - 1/b] with 1/b > 2^-18, which the bias and the
//    errors cannot cross;
//  * two digits: with B = b^2 (exact) and rB = r r rounded = (1 + D)/B,
//    |D| <= 3 2^-53, q = floor(a rB + 2^-17 rB) = floor(a / B): a/B < 2^32/B
//    moves by at most 2^-19.4/B and the rounding by 2^-21/B, together below
//    the bias 2^-17 rB > 2^-17.01/B, and bias plus errors stay below the
//    1/B that separates a non-multiple from the next integer.
// The remainders, the two digits of a mod B (one-digit step), the reversed
// pair d0 b + d1 and the reversed digits (below b a < 2^50) are exact in
// double fma.  The result is the reference's: (float) of the same integer,
// times the same float product chain of 1 / base.  Two digits per trip halve
// the trips of a wave whose lanes draw dimensions of different bases.
FLG_HD float radical_inverse(int32_t base, float inv_base, uint32_t a) {
  const double bd = (double)base;
  const double r0 = (double)inv_base;
  const double r1 = __builtin_fma(r0, __builtin_fma(-bd, r0, 1.0), r0);
  const double r = __builtin_fma(r1, __builtin_fma(-bd, r1, 1.0), r1);
  const double B = bd * bd;
  const double rB = r * r;
  const double beta = 0x1p-17 * rB;
  double ad = (double)a;
  double reversed = 0.0;
  float inv_base_n = 1.0f;
  while (ad >= bd) {  // at least two digits: the lowest two
    const double q = floor(__builtin_fma(ad, rB, beta));
    const double lo = __builtin_fma(-q, B, ad);
    const double d1 = floor(__builtin_fma(lo, r, 0x1p-19));
    const double d0 = __builtin_fma(-d1, bd, lo);
    reversed = __builtin_fma(reversed, B, __builtin_fma(d0, bd, d1));
    inv_base_n *= inv_base;
    inv_base_n *= inv_base;
    ad = q;
  }
  if (ad != 0.0) {  // the last digit
    reversed = __builtin_fma(reversed, bd, ad);
    inv_base_n *= inv_base;
  }
  return (float)reversed * inv_base_n;
}

// HaltonSampler::get1D (sampler.h:171-174) for global index `index` and
// dimension `dim`.
FLG_D float halton_get1d(const HaltonTables& t, uint32_t index, uint32_t dim, int32_t bs0,
                         int32_t bs1) {
  if (dim >= (uint32_t)t.nprimes) {
    if (t.overflow) *t.overflow = 1;
    return 0.5f;
  }
  uint32_t a = dim == 0 ? index / (uint32_t)bs0 : (dim == 1 ? index / (uint32_t)bs1 : index);
  int32_t p;
  float ib;
#if defined(__HIP_DEVICE_COMPILE__)
  // On the device no vector load from global memory feeds the entry: the
  // staged entries are read as LDS, and a dimension beyond them (or every
  // dimension when nothing is staged) comes from scalar loads, one distinct
  // dimension of the wave at a time.  A vector global load here would share
  // its registers with the LDS read at the join, and the compiler then waits
  // for all of the lane's outstanding global loads (the density lookups in
  // flight) before every draw.
  if (dim < (uint32_t)t.nlocal) {
    const int2 e = ((const __attribute__((address_space(3))) int2*)t.local)[dim];
    p = e.x;
    ib = u2f((uint32_t)e.y);
  } else {
    typedef const __attribute__((address_space(4))) int32_t ConstI32;
    typedef const __attribute__((address_space(4))) float ConstF32;
    p = 0;
    ib = 0.0f;
    bool done = false;
    while (!done) {
      const uint32_t d = __builtin_amdgcn_readfirstlane(dim);
      const int32_t pd = __builtin_amdgcn_readfirstlane(((ConstI32*)t.prime)[d]);
      const int32_t id = __builtin_amdgcn_readfirstlane((int32_t)f2u(((ConstF32*)t.inv)[d]));
      const bool mine = dim == d;
      p = mine ? pd : p;
      ib = mine ? u2f((uint32_t)id) : ib;
      done = done || mine;
    }
  }
#else
  if (dim < (uint32_t)t.nlocal) {
    const int2 e = t.local[dim];
    p = e.x;
    ib = u2f((uint32_t)e.y);
  } else {
    p = t.prime[dim];
    ib = t.inv[dim];
  }
#endif
  float v = radical_inverse(p, ib, a);
  return smin(v, ONE_MINUS_EPSILON);
}

// Copies the first kLocalPrimes table entries into the block's LDS and returns
// tables that use them.  Every thread of the block must call it (it
// synchronises the block).
FLG_D HaltonTables stage_halton(const HaltonTables& t, int2* lds) {
  const int n = t.nprimes < kLocalPrimes ? t.nprimes : kLocalPrimes;
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = make_int2(t.prime[i], (int)f2u(t.inv[i]));
  __syncthreads();
  HaltonTables r = t;
  r.local = lds;
  r.nlocal = n;
  return r;
}

// InverseRadicalInverse (sampler.h:111-120)
FLG_HD uint64_t inverse_radical_inverse(uint32_t base, uint64_t inverse, int n_digits) {
  uint64_t index = 0;
  for (int i = 0; i < n_digits; ++i) {
    uint64_t digit = inverse % base;
    inverse /= base;
    index = index * base + digit;
  }
  return index;
}

// getIndexForSample's pixel offset (sampler.h:206-227), Mod(p, 128).
FLG_HD uint32_t halton_pixel_offset(const HaltonParams& h, int px, int py) {
  if (h.stride <= 1) return 0;
  int r0 = px - (px / 128) * 128, r1 = py - (py / 128) * 128;
  uint64_t pm0 = (uint64_t)(r0 < 0 ? r0 + 128 : r0), pm1 = (uint64_t)(r1 < 0 ? r1 + 128 : r1);
  uint64_t off = inverse_radical_inverse(2, pm0, h.exp0) * (uint64_t)(h.stride / (uint32_t)h.bs0) *
                 (uint64_t)h.mulinv0;
  off += inverse_radical_inverse(3, pm1, h.exp1) * (uint64_t)(h.stride / (uint32_t)h.bs1) *
         (uint64_t)h.mulinv1;
  return (uint32_t)(off % h.stride);
}

// A path's sampler state: global index plus the next dimension.
struct PathSampler {
  uint32_t index;
  uint32_t dim;
  FLG_D float next(const HaltonTables& t, const HaltonParams& h) {
    return halton_get1d(t, index, dim++, h.bs0, h.bs1);
  }
};

}  // namespace flg
