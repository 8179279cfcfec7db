/* cl_math.h - FP64 elementary functions and error-free transforms shared by the gfx950 kernels
 * and by host code.
 *
 * Everything here is written with IEEE-754 binary64 operations that are correctly rounded on
 * both x86-64 and CDNA4 (+, -, *, /, fma, sqrt, floor, rint, comparisons) plus exact integer
 * arithmetic. Compiled with -ffp-contract=off on both sides, a call returns the same bits on
 * the CPU and on the GPU, which is what lets a chaotic FP64 orbit computed by a kernel be
 * compared bit for bit with the same orbit computed on the host.
 *
 * Methods (published, not taken from any vendor math library):
 *  - sin/cos: argument reduction x = q*pi/2 + r with r carried as a double-double.
 *    |x| < 2^50: Cody and Waite (1980) with pi/2 split over four doubles and the leading
 *    product made exact with fma (the reduced value is exact by Sterbenz' lemma).
 *    |x| >= 2^50: Payne and Hanek (1983): a 192-bit window of the bits of 2/pi selected by the
 *    exponent of x, an exact 53x192-bit integer product, the quadrant from its integer bits.
 *    Then sin(r) = r + r^3 S(r^2) and cos(r) = 1 - r^2/2 + r^4 C(r^2) on |r| <= 1 (near 2^50 the
 *    quadrant from the rounded x * 2/pi can be off by one, so |r| reaches about 0.9), with the
 *    first-order correction for the low part of r.
 *  - log: x = 2^k m, m in [sqrt(2)/2, sqrt(2)), f = m - 1, s = f/(2+f), and
 *    log(1+f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R(s^2)), R(z) = z P(z) ~ sum_i 2 z^i/(2i+1),
 *    with ln 2 split so that k*LN2_HI is exact.
 *  S, C and P (degree 6) are minimax polynomials for relative error on the reduced intervals, fitted
 *  by the Remez exchange algorithm in 200-bit arithmetic (chaos_amd/csrc/minimax.py; relative errors
 *  after rounding to double 2^-53.8, 2^-54.0, 2^-50.9) and evaluated by Horner's rule with separate
 *  multiplies and adds (a constant addend then stays a scalar operand instead of taking a vector register
 *  across the kernels' step loops, as the tied accumulator of a fused multiply-add would).
 *  - exp: Cody and Waite reduction by ln 2 and the degree-13 Taylor polynomial of exp(r) - 1 - r (see cl_exp).
 *  - TwoSum (Knuth), TwoProd with fma (Dekker's product made exact by the fused multiply-add).
 * The constants come from chaos_amd/csrc/gen_math_constants.py.
 */
#ifndef CL_MATH_H
#define CL_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CL_HD __host__ __device__ inline
#else
#include <math.h>
#define CL_HD static inline
#endif

/* pi/2 = CL_PIO2_1 + CL_PIO2_2 + CL_PIO2_3 + CL_PIO2_4 (to about 2^-216) */
#define CL_PIO2_1 0x1.921fb54442d18p+0
#define CL_PIO2_2 0x1.1a62633145c07p-54
#define CL_PIO2_3 -0x1.f1976b7ed8fbcp-110
#define CL_PIO2_4 0x1.4cf98e804177dp-164
#define CL_INV_PIO2 0x1.45f306dc9c883p-1
/* ln 2 = CL_LN2_HI (42 significant bits) + CL_LN2_LO */
#define CL_LN2_HI 0x1.62e42fefa3800p-1
#define CL_LN2_LO 0x1.ef35793c76730p-45

/* ---------------------------------------------------------------- bit helpers */
CL_HD uint64_t cl_bits(double x) {
  union { double d; uint64_t u; } v;
  v.d = x;
  return v.u;
}
CL_HD double cl_from_bits(uint64_t u) {
  union { double d; uint64_t u; } v;
  v.u = u;
  return v.d;
}
CL_HD int cl_isfinite(double x) { return (cl_bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }
CL_HD int cl_isnan(double x) { return (cl_bits(x) & 0x7fffffffffffffffull) > 0x7ff0000000000000ull; }
CL_HD int cl_signbit(double x) { return (int)(cl_bits(x) >> 63); }
CL_HD double cl_fabs(double x) { return cl_from_bits(cl_bits(x) & 0x7fffffffffffffffull); }
CL_HD double cl_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
CL_HD double cl_sqrt(double x) { return __builtin_sqrt(x); }
/* sqrt(x) for x in [1, 2] (the 2-entry norms' sqrt(1 + b^2), |b| <= 1), correctly rounded like cl_sqrt. On the
 * device, Goldschmidt's iteration from the hardware reciprocal square root y: g = x y, h = y / 2, r = 1/2 - h g,
 * g += g r, h += h r, then twice g += (x - g^2) h with a fused residual. This is the general square root's arithmetic
 * without its range scaling and special-value selects, which x in [1, 2] never needs; the device parity test compares
 * it bit for bit with the host's IEEE square root. */
CL_HD double cl_sqrt12(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
#else
  return __builtin_sqrt(x);
#endif
}
CL_HD double cl_floor(double x) { return __builtin_floor(x); }
CL_HD double cl_rint(double x) { return __builtin_rint(x); }
/* 2^e for e in [-1022, 1023] */
CL_HD double cl_pow2i(int e) { return cl_from_bits((uint64_t)(e + 1023) << 52); }

CL_HD uint64_t cl_mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

/* Julia's min/max on floats: NaN wins, -0.0 < +0.0. */
/* (the conditions are combined with | rather than ||, so that no branch splits a reducer's straight-line code;
 * x != x is x's NaN test) */
CL_HD double cl_jmin(double x, double y) {
  const int takey = (y < x) | (cl_signbit(y) > cl_signbit(x));
  return takey ? (x != x ? x : y) : (y != y ? y : x);
}
CL_HD double cl_jmax(double x, double y) {
  const int takey = (y > x) | (cl_signbit(y) < cl_signbit(x));
  return takey ? (x != x ? x : y) : (y != y ? y : x);
}
/* Julia's min and max where neither operand is NaN and the pair is never -0.0 against +0.0: there they equal IEEE
 * minNum/maxNum, one hardware instruction. */
CL_HD double cl_min_num(double x, double y) { return __builtin_fmin(x, y); }
CL_HD double cl_max_num(double x, double y) { return __builtin_fmax(x, y); }
/* Julia's max of two values that are each >= +0 or NaN (absolute values and their running maxima), NaN included:
 * such doubles are ordered as the unsigned integers of their bit patterns, with every NaN pattern above +inf's, so
 * one 64-bit compare and a select replace cl_jmax's tests. (Two NaNs of different bit patterns could give the other
 * NaN than cl_jmax; a device kernel's NaNs all carry the one default pattern.) */
CL_HD double cl_max_nonneg(double x, double y) { return cl_bits(y) > cl_bits(x) ? y : x; }

/* x / y for y in [2^-400, 2^400] and x with |x / y| in [2^-400, 2^400] or x = +0: the device division's arithmetic
 * (the hardware reciprocal, two Newton steps, the product and one fused residual correction) without the operand
 * range scaling and special-value fix-up steps, which are identities on these operands, so the quotient is the
 * correctly rounded one, bit for bit. Outside that range (a smaller nonzero quotient, a -0 numerator) the result may
 * differ from x / y; callers use it only where such a quotient cannot change what they compute. */
CL_HD double cl_div_plain(double x, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(y);
  double e = __builtin_fma(-y, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-y, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = x * r;
  const double d = __builtin_fma(-y, q, x);
  return __builtin_fma(d, r, q);
#else
  return x / y;
#endif
}

/* ---------------------------------------------------------------- error-free transforms */
/* s + e = a + b exactly (Knuth's TwoSum, no ordering assumption). */
CL_HD void cl_two_sum(double a, double b, double *s, double *e) {
  double t = a + b;
  double bb = t - a;
  *s = t;
  *e = (a - (t - bb)) + (b - bb);
}
/* s + e = a + b exactly when |a| >= |b| or a == 0 (Dekker's FastTwoSum). */
CL_HD void cl_fast_two_sum(double a, double b, double *s, double *e) {
  double t = a + b;
  *s = t;
  *e = b - (t - a);
}
/* p + e = a * b exactly (no overflow/underflow). */
CL_HD void cl_two_prod(double a, double b, double *p, double *e) {
  double t = a * b;
  *p = t;
  *e = cl_fma(a, b, -t);
}

/* ---------------------------------------------------------------- argument reduction */
#define CL_TWO_OVER_PI_WORDS \
  0xa2f9836e4e441529ull, 0xfc2757d1f534ddc0ull, 0xdb6295993c439041ull, \
  0xfe5163abdebbc561ull, 0xb7246e3a424dd2e0ull, 0x06492eea09d1921cull, \
  0xfe1deb1cb129a73eull, 0xe88235f52ebb4484ull, 0xe99c7026b45f7e41ull, \
  0x3991d639835339f4ull, 0x9c845f8bbdf9283bull, 0x1ff897ffde05980full, \
  0xef2f118b5a0a6d1full, 0x6d367ecf27cb09b7ull, 0x4f463f669e5fea2dull, \
  0x7527bac7ebe5f17bull, 0x3d0739f78a5292eaull, 0x6bfb5fb11f8d5d08ull, \
  0x56033046fc7b6babull, 0xf0cfbc209af4361dull, 0xa9e391615ee61b08ull, \
  0x6599855f14a06840ull, 0x8dffd8804d732731ull, 0x06061556ca73a8c9ull, \

/* The table lives in constant memory for the device pass and in ordinary read-only memory for
 * the host pass (a __constant__ variable read from host code would be its uninitialised shadow). */
#if defined(__HIPCC__)
__device__ __constant__ uint64_t cl_two_over_pi_dev[24] = {CL_TWO_OVER_PI_WORDS};
static const uint64_t cl_two_over_pi_host[24] = {CL_TWO_OVER_PI_WORDS};
CL_HD uint64_t cl_two_over_pi(int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return cl_two_over_pi_dev[i];
#else
  return cl_two_over_pi_host[i];
#endif
}
#else
static const uint64_t cl_two_over_pi_host[24] = {CL_TWO_OVER_PI_WORDS};
CL_HD uint64_t cl_two_over_pi(int i) { return cl_two_over_pi_host[i]; }
#endif

/* 64 bits of the 256-bit number p[0..3] (p[0] least significant) starting at bit `pos`. */
CL_HD uint64_t cl_get64(const uint64_t *p, int pos) {
  int w = pos >> 6, sh = pos & 63;
  uint64_t lo = p[w];
  uint64_t hi = (w + 1 < 4) ? p[w + 1] : 0;
  return sh == 0 ? lo : ((lo >> sh) | (hi << (64 - sh)));
}

/* Payne-Hanek reduction of a finite |x| >= 2^50: returns q mod 4 and r = x - q pi/2 as rh + rl,
 * |r| <= pi/4. */
CL_HD int cl_rem_pio2_large(double x, double *rh, double *rl) {
  uint64_t b = cl_bits(x);
  int neg = (int)(b >> 63);
  int e = (int)((b >> 52) & 0x7ff) - 1075;             /* |x| = M 2^e, e >= -2 */
  uint64_t M = (b & 0x000fffffffffffffull) | 0x0010000000000000ull;
  int j0 = e - 1 > 1 ? e - 1 : 1;                       /* first bit of 2/pi that matters mod 4 */
  int sft = j0 + 191 - e;                              /* value = (M * window) 2^-sft */
  int bit = j0 - 1, w = bit >> 6, sh = bit & 63;
  uint64_t win[3];                                      /* win[2] most significant */
  for (int i = 0; i < 3; ++i) {
    uint64_t a = cl_two_over_pi(w + i), c = cl_two_over_pi(w + i + 1);
    win[2 - i] = sh == 0 ? a : ((a << sh) | (c >> (64 - sh)));
  }
  uint64_t l0 = M * win[0], h0 = cl_mulhi64(M, win[0]);
  uint64_t l1 = M * win[1], h1 = cl_mulhi64(M, win[1]);
  uint64_t l2 = M * win[2], h2 = cl_mulhi64(M, win[2]);
  uint64_t p[4];
  p[0] = l0;
  p[1] = h0 + l1;
  uint64_t c1 = p[1] < h0;
  uint64_t t2 = h1 + l2;
  uint64_t c2 = t2 < h1;
  p[2] = t2 + c1;
  c2 += p[2] < t2;
  p[3] = h2 + c2;
  int q = (int)(cl_get64(p, sft) & 3);
  uint64_t f1 = cl_get64(p, sft - 64), f0 = cl_get64(p, sft - 128);
  int negf = 0;
  if (f1 >> 63) { /* fraction >= 1/2: take the next quadrant, fraction - 1 */
    q = (q + 1) & 3;
    negf = 1;
    f0 = ~f0 + 1;
    f1 = ~f1 + (f0 == 0);
  }
  int scale = -64;
  if (f1 == 0) {
    f1 = f0;
    f0 = 0;
    scale -= 64;
  }
  double hi = 0.0, lo = 0.0;
  if (f1 != 0) {
    int lz = __builtin_clzll(f1);
    if (lz) {
      f1 = (f1 << lz) | (f0 >> (64 - lz));
      f0 <<= lz;
    }
    scale -= lz; /* fraction = (f1 . f0) 2^(scale) with the top bit of f1 set */
    hi = (double)(f1 >> 11) * cl_pow2i(scale + 11);
    lo = (double)(((f1 & 0x7ffull) << 42) | (f0 >> 22)) * cl_pow2i(scale - 42);
  }
  /* r = fraction * pi/2 as a double-double */
  double ph, pe;
  cl_two_prod(hi, CL_PIO2_1, &ph, &pe);
  pe = pe + (hi * CL_PIO2_2 + lo * CL_PIO2_1);
  double s, t;
  cl_fast_two_sum(ph, pe, &s, &t);
  if (negf) { s = -s; t = -t; }
  if (neg) { s = -s; t = -t; q = (4 - q) & 3; }
  *rh = s;
  *rl = t;
  return q;
}

/* x = q pi/2 + (rh + rl) for finite x; returns q mod 4. */
/* Cody and Waite's reduction for pi/4 < |x| < 2^50 */
CL_HD int cl_rem_pio2_mid(double x, double *rh, double *rl) {
  double qd = cl_rint(x * CL_INV_PIO2);
  double p1, e1;
  cl_two_prod(qd, CL_PIO2_1, &p1, &e1);
  double t = x - p1; /* exact (Sterbenz) */
  double s, se;
  cl_two_sum(t, -e1, &s, &se);
  double p2, e2;
  cl_two_prod(qd, CL_PIO2_2, &p2, &e2);
  double s2, se2;
  cl_two_sum(s, -p2, &s2, &se2);
  double lo = ((se + se2) - e2) - qd * CL_PIO2_3;
  lo = lo - qd * CL_PIO2_4;
  cl_two_sum(s2, lo, rh, rl);
  int64_t qi = (int64_t)qd;
  return (int)(qi & 3);
}
CL_HD int cl_rem_pio2(double x, double *rh, double *rl) {
  double ax = cl_fabs(x);
  if (ax <= 0x1.921fb54442d18p-1) { /* pi/4 */
    *rh = x;
    *rl = 0.0;
    return 0;
  }
  if (ax >= 0x1p50) return cl_rem_pio2_large(x, rh, rl);
  return cl_rem_pio2_mid(x, rh, rl);
}

/* ---------------------------------------------------------------- sin / cos kernels, |r| <= 1 */
CL_HD double cl_sin_kernel(double rh, double rl) {
  double z = rh * rh;
  double p = -0x1.5555555555555p-3 + z * (0x1.11111111110fep-7 + z * (-0x1.a01a01a015250p-13 +
             z * (0x1.71de3a4e27e9ep-19 + z * (-0x1.ae644c23fee72p-26 + z * (0x1.611ccbda97a69p-33 +
             z * -0x1.a8fb6995f6ab8p-41)))));
  double v = z * rh;
  return rh + (rl * (1.0 - 0.5 * z) + v * p);
}

CL_HD double cl_cos_kernel(double rh, double rl) {
  double z = rh * rh;
  double p = 0x1.5555555555555p-5 + z * (-0x1.6c16c16c16c0ep-10 + z * (0x1.a01a01a017da5p-16 +
             z * (-0x1.27e4fb7450d4fp-22 + z * (0x1.1eed8a6694be0p-29 + z * (-0x1.93908da96df55p-37 +
             z * 0x1.a9983c88bc5b3p-45)))));
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (z * z * p - rh * rl));
}

/* sin and cos of one argument (one reduction). Non-finite x gives x - x (NaN). */
CL_HD void cl_sincos(double x, double *sn, double *cs) {
  /* one test for the usual case pi/4 < |x| < 2^50; then non-finite, small and large arguments */
  double rh, rl;
  int q;
  const double ax = cl_fabs(x);
  if (ax > 0x1.921fb54442d18p-1 && ax < 0x1p50) {
    q = cl_rem_pio2_mid(x, &rh, &rl);
  } else {
    if (!cl_isfinite(x)) {
      *sn = x - x;
      *cs = x - x;
      return;
    }
    q = cl_rem_pio2(x, &rh, &rl);
  }
  double s = cl_sin_kernel(rh, rl), c = cl_cos_kernel(rh, rl);
  switch (q) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}
CL_HD double cl_sin(double x) {
  double s, c;
  cl_sincos(x, &s, &c);
  return s;
}
CL_HD double cl_cos(double x) {
  double s, c;
  cl_sincos(x, &s, &c);
  return c;
}

/* ---------------------------------------------------------------- log */
/* log(x 2^k0) for a positive normal finite x: the arithmetic of cl_log after its special cases, with no branch, so
 * that independent logs interleave. */
CL_HD double cl_log_normal(double x, int k0) {
  int k = k0;
  uint64_t b = cl_bits(x);
  k += (int)(b >> 52) - 1023;
  uint64_t mant = b & 0x000fffffffffffffull;
  double m = cl_from_bits(mant | 0x3ff0000000000000ull); /* [1, 2) */
  if (m > 0x1.6a09e667f3bcdp+0) {                         /* sqrt(2) */
    m = cl_from_bits(mant | 0x3fe0000000000000ull);       /* [0.7, 1) */
    k += 1;
  }
  double f = m - 1.0;
  double s = cl_div_plain(f, 2.0 + f); /* f is +0 or |f| >= 2^-53, 2 + f in [1.7, 2.5] */
  double z = s * s;
  double R = z * (0x1.5555555555558p-1 + z * (0x1.999999999529cp-2 + z * (0x1.2492492df8420p-2 +
             z * (0x1.c71c62dd6bbfcp-3 + z * (0x1.7462b6c858cc6p-3 + z * (0x1.39fe1f2d19f7ap-3 +
             z * 0x1.2b5b3cd07f765p-3))))));
  double hfsq = 0.5 * f * f;
  double dk = (double)k;
  double lo = dk * CL_LN2_LO + s * (hfsq + R);
  return dk * CL_LN2_HI + (f - (hfsq - lo));
}

/* x is positive, normal and finite (cl_log_normal's domain); on the device one class test */
CL_HD int cl_is_pos_normal(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_class(x, 0x100);
#else
  return x >= 0x1p-1022 && x <= 0x1.fffffffffffffp+1023;
#endif
}

CL_HD double cl_log(double x) {
  /* one test for the usual case, a positive normal finite x; then a positive subnormal x, scaled by 2^54 first, and
   * the special values: NaN gives x, x < 0 NaN, +-0 -inf, +inf +inf */
  if (cl_is_pos_normal(x)) return cl_log_normal(x, 0);
  if (x > 0.0 && x < 0x1p-1022) return cl_log_normal(x * 0x1p54, -54);
  if (cl_isnan(x)) return x;
  if (x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  return x;
}


/* ---------------------------------------------------------------- atan2 */
/* atan2(y, x) in (-pi, pi]: u = min(|x|, |y|) / max(|x|, |y|) in [0, 1]; with c = k/8 (k = rint(8u)),
 * atan(u) = atan(c) + atan(t), t = (u - c) / (1 + u c), |t| <= 1/16, atan(c) a double-double constant and atan(t) its
 * Taylor series t - t^3/3 + ... - t^15/15 (the first omitted term is below 2^-64 relative); then pi/2 - a when |y| > |x|,
 * pi - a when x < 0 (double-double pi), and the sign of y. Zero arguments: atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi;
 * a NaN argument gives NaN. Within 2 units in the last place of the correctly rounded value (tested against libm). */
CL_HD double cl_atan2(double y, double x) {
  if (cl_isnan(x) || cl_isnan(y)) return x + y;
  const double ax = cl_fabs(x), ay = cl_fabs(y);
  const double PI_HI = 0x1.921fb54442d18p+1, PI_LO = 0x1.1a62633145c07p-53;
  const double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
  double a;
  if (ay == 0.0) {
    a = cl_signbit(x) ? PI_HI : 0.0;
  } else if (!cl_isfinite(ax) || !cl_isfinite(ay)) {
    /* C99's values at infinity */
    if (cl_isfinite(ax)) a = PIO2_HI;
    else if (cl_isfinite(ay)) a = cl_signbit(x) ? PI_HI : 0.0;
    else a = cl_signbit(x) ? 0x1.2d97c7f3321d2p+1 : 0x1.921fb54442d18p-1; /* 3 pi / 4, pi / 4 */
  } else {
    const int swap = ay > ax;
    const double u = swap ? ax / ay : ay / ax;
    const double AT_HI[9] = {0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2,
                             0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};
    const double AT_LO[9] = {0.0, -0x1.cd37686760c17p-59, 0x1.8ab6e3cf7afbdp-57, -0x1.c63aae6f6e918p-56,
                             0x1.a2b7f222f65e2p-56, -0x1.928df287a668fp-58, 0x1.2419a87f2a458p-56, -0x1.8c34d25aadef6p-56,
                             0x1.1a62633145c07p-55};
    const int k = (int)cl_rint(u * 8.0);
    double ahi = 0.0, alo = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) { /* selects: no indexed register array */
      ahi = j == k ? AT_HI[j] : ahi;
      alo = j == k ? AT_LO[j] : alo;
    }
    const double c = (double)k * 0.125;
    const double t = (u - c) / (1.0 + u * c);
    const double z = t * t;
    double p = -0x1.1111111111111p-4;                  /* -1/15 */
    p = p * z + 0x1.3b13b13b13b14p-4;                  /* 1/13 */
    p = p * z - 0x1.745d1745d1746p-4;                  /* 1/11 */
    p = p * z + 0x1.c71c71c71c71cp-4;                  /* 1/9 */
    p = p * z - 0x1.2492492492492p-3;                  /* 1/7 */
    p = p * z + 0x1.999999999999ap-3;                  /* 1/5 */
    p = p * z - 0x1.5555555555555p-2;                  /* 1/3 */
    a = ahi + (t + (t * (z * p) + alo));
    if (swap) a = PIO2_HI - (a - PIO2_LO);
    if (cl_signbit(x)) a = PI_HI - (a - PI_LO);
  }
  return cl_signbit(y) ? -a : a;
}

/* exp(x): k = rint(x / ln 2), r = (x - k LN2_HI) - k LN2_LO (k LN2_HI is exact: |k| < 2^11 and LN2_HI has 42
 * significant bits), |r| <= ln 2 / 2 plus rounding; exp(r) = 1 + (r + r^2 E(r)) with E the Taylor polynomial
 * sum_{n=2..13} r^(n-2) / n! (the truncation r^12 / 14! is below 2^-60 relative on the interval), the coefficients
 * 1/n! rounded to double. The power 2^k is applied by one multiply; below 2^-1022 by two (an exact one into the normal
 * range, then one rounding into the subnormals), above 2^1020 by two so the exponent field never overflows. */
CL_HD double cl_exp(double x) {
  if (cl_isnan(x)) return x;
  if (x > 710.0) return __builtin_inf();
  if (x < -746.0) return 0.0;
  double k = cl_rint(x * 0x1.71547652b82fep+0);
  double r = (x - k * CL_LN2_HI) - k * CL_LN2_LO;
  double e = 0x1.6124613a86d09p-33;
  e = e * r + 0x1.1eed8eff8d898p-29;
  e = e * r + 0x1.ae64567f544e4p-26;
  e = e * r + 0x1.27e4fb7789f5cp-22;
  e = e * r + 0x1.71de3a556c734p-19;
  e = e * r + 0x1.a01a01a01a01ap-16;
  e = e * r + 0x1.a01a01a01a01ap-13;
  e = e * r + 0x1.6c16c16c16c17p-10;
  e = e * r + 0x1.1111111111111p-7;
  e = e * r + 0x1.5555555555555p-5;
  e = e * r + 0x1.5555555555555p-3;
  e = e * r + 0.5;
  double p = 1.0 + (r + (r * r) * e);
  int ki = (int)k;
  if (ki > 1020) return (p * cl_pow2i(ki - 4)) * 16.0;
  if (ki < -1022) return (p * cl_pow2i(ki + 54)) * 0x1p-54;
  return p * cl_pow2i(ki);
}

#endif /* CL_MATH_H */
