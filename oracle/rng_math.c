/*
 * CPU oracle (test infrastructure only): Philox4x32-10, the u64 -> f64
 * conversions of rand 0.5, and the exp/norm used bit-for-bit by the device.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* Philox4x32 round constants: multipliers and Weyl key increments from
 * Salmon et al. (SC'11), Table 2 / Random123 philox.h. */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

uint64_t orc_stream_u64(uint64_t seed, uint32_t sys, uint32_t walker,
                        uint32_t epoch, uint64_t i) {
  uint64_t blk = i >> 1;
  uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), walker, epoch};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ sys};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  unsigned j = (unsigned)(i & 1u);
  return (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
}

double orc_u64_standard(uint64_t x) {
  /* 53 most significant bits times 2^-53 */
  const double scale = 1.0 / 9007199254740992.0;
  return scale * (double)(x >> 11);
}

double orc_u64_range(uint64_t x, double lo, double hi) {
  double scale = hi - lo;
  double offset = lo - scale;
  uint64_t bits = UINT64_C(0x3FF0000000000000) | (x >> 12);
  double v;
  memcpy(&v, &bits, sizeof v);
  double t = v * scale;
  return t + offset;
}

/* ln2 split so that k*LN2_HI is exact for |k| < 2^11 (fdlibm e_exp.c). */
static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;
static const double INV_LN2 = 1.44269504088896338700e+00;
static const double EXP_OVF = 709.782712893383973096;  /* ln(DBL_MAX) */
static const double EXP_UNF = -745.133219101941108420; /* ln(2^-1074) */

static double pow2i(int64_t k) {
  uint64_t bits = (uint64_t)(k + 1023) << 52;
  double v;
  memcpy(&v, &bits, sizeof v);
  return v;
}

double orc_exp(double x) {
  if (x != x) return x;
  if (x > EXP_OVF) return INFINITY;
  if (x < EXP_UNF) return 0.0;
  double kd = rint(x * INV_LN2);
  double r = fma(-kd, LN2_HI, x);
  r = fma(-kd, LN2_LO, r);
  /* Taylor coefficients 1/n!, n = 13 .. 0 */
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
  int64_t k = (int64_t)kd;
  int64_t k1 = k >> 1;
  int64_t k2 = k - k1;
  return (p * pow2i(k1)) * pow2i(k2);
}

double orc_norm3(double x, double y, double z) {
  double s = x * x;
  s = s + y * y;
  s = s + z * z;
  return sqrt(s);
}
