// Device-side primitives of the walker path: Philox4x32-10 draws, the
// u64 -> f64 maps of rand 0.5, exp() and the 3-vector norm.
//
// Every function here is compiled with -ffp-contract=off; fused
// multiply-adds appear only where fma() is written.  exp_m and norm3 spell
// the same operations, in the same order, as the CPU restatement used to
// check them, so both round identically.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mole {

// Philox4x32 multipliers and Weyl key increments (Salmon et al., SC'11).
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
  }
  return c;
}

// A walker's u64 draw stream.  Draw i lives in Philox block i >> 1 with
// counter (lo32(i>>1), hi32(i>>1), walker, epoch) and key
// (lo32(seed), hi32(seed) ^ system); its value is word[2j] | word[2j+1]<<32
// with j = i & 1.  All walkers of a launch consume draws in lockstep, so
// the "need a new block" test is wave-uniform.
struct Stream {
  uint32_t k0, k1, walker, epoch;
  uint64_t blk;  // block held in `words`, ~0 if none
  u32x4 words;

  __device__ __forceinline__ void init(uint64_t seed, uint32_t sys,
                                       uint32_t w, uint32_t ep) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32) ^ sys;
    walker = w;
    epoch = ep;
    blk = ~(uint64_t)0;
  }
  __device__ __forceinline__ uint64_t draw(uint64_t i) {
    const uint64_t b = i >> 1;
    if (b != blk) {
      blk = b;
      words = philox4x32_10({(uint32_t)b, (uint32_t)(b >> 32), walker, epoch},
                            k0, k1);
    }
    return (i & 1) ? ((uint64_t)words.z | ((uint64_t)words.w << 32))
                   : ((uint64_t)words.x | ((uint64_t)words.y << 32));
  }
};

// rand 0.5 Standard f64: 53 high bits times 2^-53.
__device__ __forceinline__ double u64_standard(uint64_t x) {
  return (1.0 / 9007199254740992.0) * (double)(x >> 11);
}

// rand 0.5 Range(lo, hi) f64: v in [1,2) from 52 high bits, v*scale+offset.
__device__ __forceinline__ double u64_range(uint64_t x, double scale,
                                            double offset) {
  const double v = __longlong_as_double(
      (long long)(0x3FF0000000000000ull | (x >> 12)));
  const double t = v * scale;
  return t + offset;
}

// exp(x): Cody-Waite reduction with the fdlibm split of ln2, degree-13
// Taylor polynomial by Horner with fma, two exact power-of-two scalings.
__device__ __forceinline__ double exp_m(double x) {
  const double kLn2Hi = 6.93147180369123816490e-01;
  const double kLn2Lo = 1.90821492927058770002e-10;
  const double kInvLn2 = 1.44269504088896338700e+00;
  const double kOvf = 709.782712893383973096;
  const double kUnf = -745.133219101941108420;
  const double xc = fmin(fmax(x, kUnf), kOvf);
  const double kd = rint(xc * kInvLn2);
  double r = fma(-kd, kLn2Hi, xc);
  r = fma(-kd, kLn2Lo, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const int k = (int)kd;
  const int k1 = k >> 1;
  const int k2 = k - k1;
  const double s1 = __longlong_as_double((long long)(k1 + 1023) << 52);
  const double s2 = __longlong_as_double((long long)(k2 + 1023) << 52);
  double res = (p * s1) * s2;
  // out-of-range and NaN inputs: +inf above, 0 below, NaN through
  res = (x > kOvf) ? __longlong_as_double(0x7FF0000000000000ll) : res;
  res = (x < kUnf) ? 0.0 : res;
  res = (x != x) ? x : res;
  return res;
}

// |v| as sqrt((x*x + y*y) + z*z), two roundings per term.
__device__ __forceinline__ double norm3(double x, double y, double z) {
  double s = x * x;
  s = s + y * y;
  s = s + z * z;
  return sqrt(s);
}

// Rust f64::min(a, 1.0): NaN yields the other operand.
__device__ __forceinline__ double min_one(double a) {
  return (a < 1.0) ? a : 1.0;
}

// c ? a : b on values.  `v[i] = c ? w[i] : v[i]` with array elements is a
// conditional on lvalues: it is compiled as a branch over the two element
// addresses and one load, which can leave the arrays in scratch memory (a
// memory round trip and a full wait per select); the arguments here are
// loaded first, so the arrays stay in registers and the select is a
// v_cndmask.
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }

// ln x for finite x > 0 with a normal exponent: x = 2^k m, m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) by its series to s^20
// (Horner with fma), ln x = k ln2_hi + (ln m + k ln2_lo), fdlibm split.
__device__ __forceinline__ double log_m(double x) {
  const double kLn2Hi = 6.93147180369123816490e-01;
  const double kLn2Lo = 1.90821492927058770002e-10;
  const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  int k = (int)(bits >> 52) - 1023;
  double m = __longlong_as_double((long long)((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
  if (m > 1.41421356237309504880) {
    m = m * 0.5;
    k = k + 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  p = fma(p, s2, 1.0);
  const double lnm = (2.0 * s) * p;
  const double kd = (double)k;
  return fma(kd, kLn2Hi, fma(kd, kLn2Lo, lnm));
}

// sin(pi x), cos(pi x) for 0 <= x < 2: n = rint(2x), y = x - n/2 (exact),
// Taylor polynomials of sin (to t^17) and cos (to t^18) at t = pi y, then
// the quadrant n.
__device__ __forceinline__ void sincospi_m(double x, double *sn, double *cs) {
  const double nd = rint(2.0 * x);
  const double y = x - 0.5 * nd;
  const double t = 3.14159265358979323846 * y;
  const double t2 = t * t;
  double ps = -1.0 / 355687428096000.0;
  ps = fma(ps, t2, 1.0 / 1307674368000.0);
  ps = fma(ps, t2, -1.0 / 6227020800.0);
  ps = fma(ps, t2, 1.0 / 39916800.0);
  ps = fma(ps, t2, -1.0 / 362880.0);
  ps = fma(ps, t2, 1.0 / 5040.0);
  ps = fma(ps, t2, -1.0 / 120.0);
  ps = fma(ps, t2, 1.0 / 6.0);
  const double s = fma(-(t * t2), ps, t);
  double pc = 1.0 / 6402373705728000.0;
  pc = fma(pc, t2, -1.0 / 20922789888000.0);
  pc = fma(pc, t2, 1.0 / 87178291200.0);
  pc = fma(pc, t2, -1.0 / 479001600.0);
  pc = fma(pc, t2, 1.0 / 3628800.0);
  pc = fma(pc, t2, -1.0 / 40320.0);
  pc = fma(pc, t2, 1.0 / 720.0);
  pc = fma(pc, t2, -1.0 / 24.0);
  pc = fma(pc, t2, 0.5);
  const double c = fma(-t2, pc, 1.0);
  const int q = (int)nd & 3;
  *sn = q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
  *cs = q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}

// Two N(0,1) values from two u64 draws: u1 = 1 - Standard(a) in (0, 1],
// u2 = Standard(b), r = sqrt(-2 ln u1), (r cos 2 pi u2, r sin 2 pi u2).
__device__ __forceinline__ void box_muller(uint64_t a, uint64_t b, double *z0, double *z1) {
  const double u1 = 1.0 - u64_standard(a);
  const double u2 = u64_standard(b);
  const double r = sqrt(-2.0 * log_m(u1));
  double sn, cs;
  sincospi_m(2.0 * u2, &sn, &cs);
  *z0 = r * cs;
  *z1 = r * sn;
}

// float -> bfloat16 bits, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

}  // namespace mole
