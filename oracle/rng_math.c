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

/* ln x for finite x > 0 with a normal exponent: x = 2^k m with m in
 * [sqrt(1/2), sqrt(2)), s = (m - 1)/(m + 1), ln m = 2 atanh(s) =
 * 2 s sum_{j=0..10} s^(2j)/(2j+1) (|s| <= 0.1716, so the series is exhausted
 * at s^20), ln x = k ln2_hi + (ln m + k ln2_lo) with the fdlibm split of ln2.
 * Every step is an IEEE operation, so the device gets the same bits. */
double orc_log(double x) {
  uint64_t bits;
  memcpy(&bits, &x, sizeof bits);
  int64_t k = (int64_t)(bits >> 52) - 1023;
  uint64_t mb = (bits & UINT64_C(0x000FFFFFFFFFFFFF)) | UINT64_C(0x3FF0000000000000);
  double m;
  memcpy(&m, &mb, sizeof m);
  if (m > 1.41421356237309504880) {
    m = m * 0.5;
    k = k + 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double s2 = s * s;
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
  double lnm = (2.0 * s) * p;
  double kd = (double)k;
  return fma(kd, LN2_HI, fma(kd, LN2_LO, lnm));
}

/* sin(pi x), cos(pi x) for 0 <= x < 2: n = rint(2x), y = x - n/2 in
 * [-1/4, 1/4] (exact), t = pi y, Taylor polynomials of sin t (to t^17) and
 * cos t (to t^18) by Horner with fma, then the quadrant n. */
void orc_sincospi(double x, double *sn, double *cs) {
  double nd = rint(2.0 * x);
  double y = x - 0.5 * nd;
  double t = 3.14159265358979323846 * y;
  double t2 = t * t;
  double ps = -1.0 / 355687428096000.0;      /* -1/17! */
  ps = fma(ps, t2, 1.0 / 1307674368000.0);  /* 1/15! */
  ps = fma(ps, t2, -1.0 / 6227020800.0);    /* -1/13! */
  ps = fma(ps, t2, 1.0 / 39916800.0);       /* 1/11! */
  ps = fma(ps, t2, -1.0 / 362880.0);
  ps = fma(ps, t2, 1.0 / 5040.0);
  ps = fma(ps, t2, -1.0 / 120.0);
  ps = fma(ps, t2, 1.0 / 6.0);
  double s = fma(-(t * t2), ps, t); /* t - t^3 (1/6 - t^2/120 + ...) */
  double pc = 1.0 / 6402373705728000.0;     /* 1/18! */
  pc = fma(pc, t2, -1.0 / 20922789888000.0); /* -1/16! */
  pc = fma(pc, t2, 1.0 / 87178291200.0);     /* 1/14! */
  pc = fma(pc, t2, -1.0 / 479001600.0);      /* -1/12! */
  pc = fma(pc, t2, 1.0 / 3628800.0);
  pc = fma(pc, t2, -1.0 / 40320.0);
  pc = fma(pc, t2, 1.0 / 720.0);
  pc = fma(pc, t2, -1.0 / 24.0);
  pc = fma(pc, t2, 0.5);
  double c = fma(-t2, pc, 1.0);
  switch ((int)nd & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

/* Box-Muller pair from two u64 draws: u1 = 1 - Standard(a) in (0, 1],
 * u2 = Standard(b); r = sqrt(-2 ln u1); (r cos 2 pi u2, r sin 2 pi u2). */
void orc_box_muller(uint64_t a, uint64_t b, double *z0, double *z1) {
  double u1 = 1.0 - orc_u64_standard(a);
  double u2 = orc_u64_standard(b);
  double r = sqrt(-2.0 * orc_log(u1));
  double sn, cs;
  orc_sincospi(2.0 * u2, &sn, &cs);
  *z0 = r * cs;
  *z1 = r * sn;
}
