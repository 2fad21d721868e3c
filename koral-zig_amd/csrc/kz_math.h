/* kz_math.h — portable FP64 transcendentals for the koral step, compiled
 * identically by gcc (CPU restatement, CPU baseline) and by hipcc (gfx950
 * kernels).
 *
 * Why this exists: the step's Newton solvers take branch decisions (floor
 * hits, convergence exits, damping retries) on values that pass through
 * log/exp/pow/cbrt/log1p.  glibc, ROCm's ocml and Zig's std.math all round
 * those functions differently in the last place, which would make the GPU's
 * iteration counts and fixup flags drift from the CPU path's.  Every routine
 * here is built only from IEEE-754 +, -, *, /, sqrt and fused multiply-add
 * (all correctly rounded on both targets) and from exact bit manipulation,
 * so with -ffp-contract=off the CPU and the GPU produce bit-identical
 * results.
 *
 * Methods (all textbook; no vendor library is reproduced):
 *   exp   — Cody & Waite (1980) additive argument reduction x = k ln2 + r with
 *           a split ln2 whose high part has 32 significant bits (k ln2_hi is
 *           exact for |k| < 2^21), the reduction error carried as a correction
 *           term, Taylor series of degree 13 for exp(r) - 1 - r (Horner with
 *           fused multiply-adds), scaling by an exponent-field construction
 *           of 2^k.
 *   log   — reduction to m in [sqrt(1/2), sqrt(2)), then
 *           log(m) = 2 atanh(s), s = (m-1)/(m+1), with s carried as a
 *           double-double (one division, the remainder by fused
 *           multiply-add), odd series in s up to s^23.  Result available as a double-double (used by
 *           pow).
 *   pow   — exp(y log x) evaluated in double-double: the product y*log(x) is
 *           formed exactly (TwoProduct by fused multiply-add) and the low part enters
 *           exp as a first-order correction, so the error does not grow with
 *           |y log x|.
 *   cbrt  — division-free: Newton iterations for m^(-1/3) from an
 *           exponent-field seed, t = m y^2, and a final Newton correction
 *           with the residual m - t^3 formed in double-double.
 *   log1p — Goldberg (1991): log1p(x) = log(u) * x / (u - 1), u = 1 + x.
 *
 * The includer must define KZM_QUAL (e.g. `static inline` for C, or
 * `__host__ __device__ static inline` for HIP) before including this file.
 * Accuracy against glibc is tested in tests/test_kz_math.py (<= 2 ulp).
 */
#ifndef KZ_MATH_H
#define KZ_MATH_H

#include <stdint.h>

#ifndef KZM_QUAL
#error "define KZM_QUAL before including kz_math.h"
#endif

KZM_QUAL uint64_t kzm_bits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
KZM_QUAL double kzm_from_bits(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
KZM_QUAL double kzm_fabs(double x) { return __builtin_fabs(x); }
KZM_QUAL double kzm_sqrt(double x) { return __builtin_sqrt(x); }
KZM_QUAL int kzm_isnan(double x) { return x != x; }
KZM_QUAL int kzm_isfinite(double x) { return kzm_fabs(x) <= 0x1.fffffffffffffp+1023; }
/* @min/@max with a NaN operand yield the other operand (koral/math/simd.zig
 * contract).  Written as explicit comparisons so that ties (+0/-0) resolve
 * the same way on both compilers: an equal pair returns b. */
KZM_QUAL double kzm_fmin(double a, double b) { return (a < b || b != b) ? a : b; }
KZM_QUAL double kzm_fmax(double a, double b) { return (a > b || b != b) ? a : b; }

#define KZM_INF (__builtin_inf())
#define KZM_NAN (__builtin_nan(""))

/* Fused multiply-add a*b + c with one rounding: correctly rounded by IEEE
 * 754 on both targets (v_fma_f64 on gfx950; the FMA3 instruction or libm's
 * fma on the host), so it keeps the two builds bit-identical.  A build may
 * define KZM_FMA itself (the flop-counting build of the oracle does). */
#ifndef KZM_FMA
#define KZM_FMA(a, b, c) __builtin_fma((a), (b), (c))
#endif

/* ln 2 = LN2_HI + LN2_LO; LN2_HI has its low 21 mantissa bits clear. */
#define KZM_LN2_HI 0x1.62e42fee00000p-1
#define KZM_LN2_LO 0x1.a39ef35793c76p-33
#define KZM_INV_LN2 0x1.71547652b82fep+0

/* ---- double-double helpers (Knuth TwoSum; TwoProduct by FMA) ------------- */
KZM_QUAL void kzm_two_sum(double a, double b, double *s, double *e) {
  double ss = a + b;
  double bb = ss - a;
  *e = (a - (ss - bb)) + (b - bb);
  *s = ss;
}
KZM_QUAL void kzm_fast_two_sum(double a, double b, double *s, double *e) {
  /* requires |a| >= |b| (or a == 0) */
  double ss = a + b;
  *e = b - (ss - a);
  *s = ss;
}
/* p + e == a*b exactly (the rounding error of a product is representable
 * unless it underflows): e = fma(a, b, -p). */
KZM_QUAL void kzm_two_prod(double a, double b, double *p, double *e) {
  double pp = a * b;
  *e = KZM_FMA(a, b, -pp);
  *p = pp;
}

/* ---- exp ------------------------------------------------------------------ */
/* 2^k as a double for -1022 <= k <= 1023. */
KZM_QUAL double kzm_pow2i(int k) { return kzm_from_bits((uint64_t)(int64_t)(k + 1023) << 52); }

/* exp(x + xl) for a small correction |xl| << ulp(x) scale. */
KZM_QUAL double kzm_exp_dd(double x, double xl) {
  if (x != x) return x;
  if (x > 709.782712893384) return KZM_INF;
  if (x < -745.1332191019412) return 0.0;
  double t = x * KZM_INV_LN2;
  int k = (int)(t >= 0.0 ? t + 0.5 : t - 0.5);
  double kd = (double)k;
  double hi = x - kd * KZM_LN2_HI; /* exact: kd*LN2_HI exact, Sterbenz */
  double lo = KZM_FMA(kd, KZM_LN2_LO, -xl);
  double r = hi - lo;
  double c = (hi - r) - lo; /* r + c == hi - lo to double-double accuracy */
  /* q = exp(r) - 1 - r = r^2 (1/2! + r/3! + ... + r^11/13!), Horner steps
   * fused */
  double q = 0x1.6124613a86d09p-33;              /* 1/13! */
  q = KZM_FMA(r, q, 0x1.1eed8eff8d898p-29);             /* 1/12! */
  q = KZM_FMA(r, q, 0x1.ae64567f544e4p-26);             /* 1/11! */
  q = KZM_FMA(r, q, 0x1.27e4fb7789f5cp-22);             /* 1/10! */
  q = KZM_FMA(r, q, 0x1.71de3a556c734p-19);             /* 1/9!  */
  q = KZM_FMA(r, q, 0x1.a01a01a01a01ap-16);             /* 1/8!  */
  q = KZM_FMA(r, q, 0x1.a01a01a01a01ap-13);             /* 1/7!  */
  q = KZM_FMA(r, q, 0x1.6c16c16c16c17p-10);             /* 1/6!  */
  q = KZM_FMA(r, q, 0x1.1111111111111p-7);              /* 1/5!  */
  q = KZM_FMA(r, q, 0x1.5555555555555p-5);              /* 1/4!  */
  q = KZM_FMA(r, q, 0x1.5555555555555p-3);              /* 1/3!  */
  q = KZM_FMA(r, q, 0.5);                               /* 1/2!  */
  q = (r * r) * q;
  double y = 1.0 + (r + (q + c * (1.0 + r)));
  if (k > 1023) return (y * 2.0) * kzm_pow2i(k - 1);
  if (k >= -1021) return y * kzm_pow2i(k);
  /* subnormal result: scale in two steps */
  return (y * kzm_pow2i(k + 1000)) * kzm_pow2i(-1000);
}

KZM_QUAL double kzm_exp(double x) { return kzm_exp_dd(x, 0.0); }

/* ---- log ------------------------------------------------------------------ */
/* log(x) as a double-double (hi + lo) for finite x > 0. */
KZM_QUAL void kzm_log_dd(double x, double *hi_out, double *lo_out) {
  uint64_t b = kzm_bits(x);
  int eadj = 0;
  if (x < 0x1p-1022) { /* subnormal */
    x *= 0x1p54;
    b = kzm_bits(x);
    eadj = -54;
  }
  int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
  double m = kzm_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
  if (m > 0x1.6a09e667f3bcdp+0) { /* > sqrt(2) */
    m *= 0.5;
    e += 1;
  }
  double f = m - 1.0; /* exact for m in [1/2, 2] */
  double uh, ul;
  kzm_fast_two_sum(2.0, f, &uh, &ul);
  /* s = f / u as sh + sl with one division: sh = f * (1/uh) is within an
   * ulp or two of the quotient, and the remainder f - sh*uh is exact by
   * fma, so sl carries what sh misses */
  double inv = 1.0 / uh;
  double sh = f * inv;
  double sl = (KZM_FMA(-sh, uh, f) - sh * ul) * inv;
  double s2 = sh * sh;
  /* tail = 2 s^3 (1/3 + s^2/5 + ... + s^20/23) */
  double p = 0x1.642c8590b2164p-4;                /* 2/23 */
  p = KZM_FMA(s2, p, 0x1.8618618618618p-4);              /* 2/21 */
  p = KZM_FMA(s2, p, 0x1.af286bca1af28p-4);              /* 2/19 */
  p = KZM_FMA(s2, p, 0x1.e1e1e1e1e1e1ep-4);              /* 2/17 */
  p = KZM_FMA(s2, p, 0x1.1111111111111p-3);              /* 2/15 */
  p = KZM_FMA(s2, p, 0x1.3b13b13b13b14p-3);              /* 2/13 */
  p = KZM_FMA(s2, p, 0x1.745d1745d1746p-3);              /* 2/11 */
  p = KZM_FMA(s2, p, 0x1.c71c71c71c71cp-3);              /* 2/9  */
  p = KZM_FMA(s2, p, 0x1.2492492492492p-2);              /* 2/7  */
  p = KZM_FMA(s2, p, 0x1.999999999999ap-2);              /* 2/5  */
  p = KZM_FMA(s2, p, 0x1.5555555555555p-1);              /* 2/3  */
  double tail = (s2 * sh) * p;
  double ed = (double)e;
  double a = ed * KZM_LN2_HI; /* exact */
  double hi, err;
  kzm_two_sum(a, 2.0 * sh, &hi, &err);
  double lo = err + (ed * KZM_LN2_LO + (2.0 * sl + tail));
  kzm_fast_two_sum(hi, lo, hi_out, lo_out);
}

KZM_QUAL double kzm_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return KZM_NAN;
  if (x == 0.0) return -KZM_INF;
  if (x == KZM_INF) return KZM_INF;
  double h, l;
  kzm_log_dd(x, &h, &l);
  return h;
}

/* ---- pow ------------------------------------------------------------------ */
KZM_QUAL int kzm_is_int(double y) {
  /* y finite: integral iff y == trunc(y); |y| >= 2^52 is always integral */
  double a = kzm_fabs(y);
  if (a >= 0x1p52) return 1;
  double t = (double)(int64_t)a;
  return t == a;
}
KZM_QUAL int kzm_is_odd_int(double y) {
  double a = kzm_fabs(y);
  if (a >= 0x1p53) return 0;
  int64_t t = (int64_t)a;
  return ((double)t == a) && (t & 1);
}

KZM_QUAL double kzm_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (x != x || y != y) return KZM_NAN;
  double sign = 1.0;
  if (x < 0.0) {
    if (!kzm_isfinite(y) || !kzm_is_int(y)) {
      if (x == -KZM_INF && !kzm_isfinite(y)) return y > 0.0 ? KZM_INF : 0.0;
      if (!kzm_isfinite(y)) { /* |x| vs 1 decides */
        double ax = -x;
        if (ax == 1.0) return 1.0;
        return ((ax > 1.0) == (y > 0.0)) ? KZM_INF : 0.0;
      }
      return KZM_NAN;
    }
    if (kzm_is_odd_int(y)) sign = -1.0;
    x = -x;
  }
  if (x == 0.0) {
    if (y > 0.0) return sign * 0.0;
    return sign * KZM_INF;
  }
  if (x == KZM_INF) return y > 0.0 ? sign * KZM_INF : sign * 0.0;
  if (!kzm_isfinite(y)) {
    if (x == 1.0) return 1.0;
    return ((x > 1.0) == (y > 0.0)) ? KZM_INF : 0.0;
  }
  double lh, ll;
  kzm_log_dd(x, &lh, &ll);
  double zh, zl;
  kzm_two_prod(y, lh, &zh, &zl);
  zl += y * ll;
  double th, tl;
  kzm_fast_two_sum(zh, zl, &th, &tl);
  return sign * kzm_exp_dd(th, tl);
}

/* kzm_pow(x, y) for finite x > 0 and finite y, from log(x) as kzm_log_dd
 * returns it: the operations kzm_pow performs after its special cases (and
 * 1.0 for y == 0 and for x == 1, where log_dd gives exactly zero), so a
 * caller taking several powers of one base can form the logarithm once and
 * get kzm_pow's bits for each. */
KZM_QUAL double kzm_pow_from_log(double lh, double ll, double y) {
  if (y == 0.0) return 1.0;
  double zh, zl;
  kzm_two_prod(y, lh, &zh, &zl);
  zl += y * ll;
  double th, tl;
  kzm_fast_two_sum(zh, zl, &th, &tl);
  return kzm_exp_dd(th, tl);
}

/* ---- cbrt ----------------------------------------------------------------- */
KZM_QUAL double kzm_cbrt(double x) {
  if (x == 0.0 || x != x || !kzm_isfinite(x)) return x;
  double sign = x < 0.0 ? -1.0 : 1.0;
  double ax = kzm_fabs(x);
  int eadj = 0;
  if (ax < 0x1p-1022) {
    ax *= 0x1p54;
    eadj = -54;
  }
  /* ax = m * 2^(3j) with m in [1, 8); cbrt(ax) = cbrt(m) * 2^j */
  uint64_t b = kzm_bits(ax);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  int et = e + eadj;
  int j = (et >= 0) ? et / 3 : -((-et + 2) / 3);
  int r3 = et - 3 * j; /* 0, 1, 2 */
  double m = kzm_from_bits((b & 0x000fffffffffffffULL) | ((uint64_t)(1023 + r3) << 52));
  /* y ~ m^(-1/3) without division: seed from the exponent field,
   * bits(y) ~ (4/3)*1023*2^52 - bits(m)/3 (under 10% off), then five Newton
   * steps y <- y + y (1 - m y^3) / 3 (the error e goes to about 2 e^2) */
  double y = kzm_from_bits(0x5540000000000000ULL - kzm_bits(m) / 3);
  for (int i = 0; i < 5; ++i) {
    double y3 = (y * y) * y;
    y = KZM_FMA(y * KZM_FMA(-m, y3, 1.0), 0x1.5555555555555p-2, y);
  }
  double y2 = y * y; /* ~ 1 / cbrt(m)^2 */
  double t = m * y2;
  /* final Newton correction t + (m - t^3) / (3 t^2), the residual formed in
   * double-double and 1 / t^2 taken as y^2 */
  double p2h, p2l, p3h, p3l;
  kzm_two_prod(t, t, &p2h, &p2l);
  kzm_two_prod(p2h, t, &p3h, &p3l);
  p3l += p2l * t;
  double resid = (m - p3h) - p3l;
  t = KZM_FMA(resid, y2 * 0x1.5555555555555p-2, t);
  return sign * (t * kzm_pow2i(j));
}

/* ---- log1p ---------------------------------------------------------------- */
KZM_QUAL double kzm_log1p(double x) {
  if (x != x) return x;
  if (x == -1.0) return -KZM_INF;
  if (x < -1.0) return KZM_NAN;
  if (x == KZM_INF) return KZM_INF;
  double u = 1.0 + x;
  if (u == 1.0) return x;
  return kzm_log(u) * (x / (u - 1.0));
}

#endif /* KZ_MATH_H */
