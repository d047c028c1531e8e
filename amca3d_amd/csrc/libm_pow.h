// pow(x, y) for x > 0, as glibc 2.35 computes it on x86-64, written for host and
// device from one definition.
//
// The remelting growth velocity (ProblemPhysics.cpp:142-158) sums
// a_n * pow(dT, b_n).  glibc's pow is not correctly rounded (its bound is 0.52
// ulp): pow(x, 2) is the neighbour of x*x for about 0.08 % of arguments, and
// pow(x, 3) differs from (x*x)*x for about a quarter of them.  The grain IDs
// depend on the half-diagonals, so the device evaluates glibc's own algorithm.
//
// Algorithm (glibc 2.35 sysdeps/ieee754/dbl-64/e_pow.c, from Arm's
// optimized-routines, 2018): pow = exp(y * log(x)) with
//   log(x) = k ln2 + log(c) + log1p(z/c - 1),  x = 2^k z,  1/c from a 128-entry
//            table, a degree-8 polynomial for log1p, result as hi + lo;
//   exp(u) = 2^(j/128) * exp(r),  u = j ln2/128 + r,  a degree-5 polynomial,
//            2^(j/128) from a 128-entry table as scale * (1 + tail).
// The tables are generated from their published construction by
// gen_libm_pow_tables.py.  The polynomial coefficients are those of the same
// release.  On x86-64 glibc runs its FMA build (ifunc `__pow_fma`, chosen on
// every CPU with FMA and AVX2), and GCC contracts each single-use product into
// the sum it feeds; every such contraction is written here as an explicit fma,
// and this file is compiled with contraction off.
//
// Domain: finite x > 0, subnormals included (the callers pass a positive
// undercooling), and y == 0 or 2^-65 <= |y| < 2^63 (callers check y once with
// pow_exponent_supported).  Results then equal glibc's pow bit for bit, the
// overflow and underflow branches included.
#pragma once

#include <stdint.h>

#include "libm_pow_tables.h"

#ifdef __HIPCC__
#define AMCA_LIBM_HD __host__ __device__
#else
#define AMCA_LIBM_HD
#endif

namespace amca_libm {

AMCA_LIBM_HD inline uint64_t as_u64(double d) { return __builtin_bit_cast(uint64_t, d); }
AMCA_LIBM_HD inline double as_f64(uint64_t u) { return __builtin_bit_cast(double, u); }
AMCA_LIBM_HD inline uint32_t top12(double d) { return (uint32_t)(as_u64(d) >> 52); }

// log1p polynomial, scaled to the evaluation below (A[0] = -1/2, then the
// minimax coefficients of r^3..r^8 times -2, -2, 4, 4, -8, -8).
constexpr double kLogA[7] = {-0x1p-1,
                             0x1.555555555556p-2 * -2,
                             -0x1.0000000000006p-2 * -2,
                             0x1.999999959554ep-3 * 4,
                             -0x1.555555529a47ap-3 * 4,
                             0x1.2495b9b4845e9p-3 * -8,
                             -0x1.0002b8b263fc3p-3 * -8};
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;  // low 11 bits clear: k*Ln2Hi exact
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128;
constexpr double kNegLn2HiN = -0x1.62e42fefa0000p-8;
constexpr double kNegLn2LoN = -0x1.cf79abc9e3b3ap-47;
constexpr double kShift = 0x1.8p52;
// exp(r) - 1 - r polynomial, r^2..r^5
constexpr double kExpC2 = 0x1.ffffffffffdbdp-2;
constexpr double kExpC3 = 0x1.555555555543cp-3;
constexpr double kExpC4 = 0x1.55555cf172b91p-5;
constexpr double kExpC5 = 0x1.1111167a4d017p-7;
constexpr uint64_t kLogOff = 0x3fe6955500000000ull;

// log(x) as hi + *tail for the bit pattern of a positive normal x.
AMCA_LIBM_HD inline double pow_log(uint64_t ix, double* tail) {
  const uint64_t tmp = ix - kLogOff;
  const int i = (int)((tmp >> 45) % 128);
  const int k = (int)((int64_t)tmp >> 52);
  const uint64_t iz = ix - (tmp & (0xfffull << 52));
  const double z = as_f64(iz), kd = (double)k;
  const double invc = kPowLogTab[i][0], logc = kPowLogTab[i][1], logctail = kPowLogTab[i][2];
  // 1/c has few bits and |z/c - 1| < 1/128, so r is exact
  const double r = __builtin_fma(z, invc, -1.0);
  const double t1 = __builtin_fma(kd, kLn2Hi, logc);
  const double t2 = t1 + r;
  const double lo1 = __builtin_fma(kd, kLn2Lo, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = kLogA[0] * r;
  const double ar2 = r * ar;
  const double ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = __builtin_fma(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  const double q3 = __builtin_fma(r, kLogA[6], kLogA[5]);
  const double q2 = __builtin_fma(ar2, q3, __builtin_fma(r, kLogA[4], kLogA[3]));
  const double q1 = __builtin_fma(ar2, q2, __builtin_fma(r, kLogA[2], kLogA[1]));
  const double lo = __builtin_fma(ar3, q1, lo1 + lo2 + lo3 + lo4);
  const double y = hi + lo;
  *tail = hi - y + lo;
  return y;
}

// Results of exp that leave the normal range: 512 <= |u| < 1024.
AMCA_LIBM_HD inline double pow_exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  if ((ki & 0x80000000ull) == 0) {
    // k > 0: the exponent of the scale may have overflowed by up to 460
    sbits -= 1009ull << 52;
    const double scale = as_f64(sbits);
    return 0x1p1009 * __builtin_fma(scale, tmp, scale);
  }
  // k < 0: round to the final precision before scaling into the subnormal range
  sbits += 1022ull << 52;
  const double scale = as_f64(sbits);
  // scale * tmp feeds two sums here, so GCC keeps the product unfused
  const double st = scale * tmp;
  double y = scale + st;
  if (__builtin_fabs(y) < 1.0) {
    const double one = y < 0.0 ? -1.0 : 1.0;
    double lo = scale - y + st;
    const double hi = one + y;
    lo = one - hi + y + lo;
    y = (hi + lo) - one;
    if (y == 0) y = as_f64(sbits & 0x8000000000000000ull);
  }
  return 0x1p-1022 * y;
}

// exp(x + xtail)
AMCA_LIBM_HD inline double pow_exp(double x, double xtail) {
  uint32_t abstop = top12(x) & 0x7ff;
  if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
    if (abstop - top12(0x1p-54) >= 0x80000000u) return 1.0 + x;  // |x| < 2^-54
    if (abstop >= top12(1024.0)) return (as_u64(x) >> 63) ? 0.0 : __builtin_inf();
    abstop = 0;  // 512 <= |x| < 1024
  }
  // x = j ln2/N + r, r in [-ln2/2N, ln2/2N]
  double kd = __builtin_fma(kInvLn2N, x, kShift);
  const uint64_t ki = as_u64(kd);
  kd -= kShift;
  double r = __builtin_fma(kd, kNegLn2LoN, __builtin_fma(kd, kNegLn2HiN, x));
  r += xtail;
  const uint64_t idx = 2 * (ki % 128);
  const uint64_t top = ki << 45;
  const double tail = as_f64(kExpTab[idx]);
  const uint64_t sbits = kExpTab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, kExpC5, kExpC4),
                                   __builtin_fma(r2, __builtin_fma(r, kExpC3, kExpC2), tail + r));
  if (abstop == 0) return pow_exp_special(tmp, sbits, ki);
  const double scale = as_f64(sbits);
  return __builtin_fma(scale, tmp, scale);
}

// Exponents the restatement covers (everything but |y| < 2^-65 other than 0,
// |y| >= 2^63, inf and nan).
AMCA_LIBM_HD inline bool pow_exponent_supported(double y) {
  if (y == 0.0) return true;
  const uint32_t t = top12(y) & 0x7ff;
  return t - 0x3beu < 0x43eu - 0x3beu;
}

// glibc's pow(x, y) for finite x > 0 and a supported y.
AMCA_LIBM_HD inline double glibc_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  uint64_t ix = as_u64(x);
  if (top12(x) == 0) {
    // subnormal x: normalise so that the exponent becomes negative
    ix = as_u64(x * 0x1p52);
    ix &= 0x7fffffffffffffffull;
    ix -= 52ull << 52;
  }
  double lo;
  const double hi = pow_log(ix, &lo);
  const double ehi = y * hi;
  const double elo = __builtin_fma(y, lo, __builtin_fma(y, hi, -ehi));
  return pow_exp(ehi, elo);
}

// Whether r is the correctly rounded value of an exact r + res (|res| <= 1/2
// ulp) that lies more than 1/16 ulp from a rounding midpoint: r normal, not a
// power of two (the spacing below it halves) and not in the top binade.
AMCA_LIBM_HD inline bool clear_of_midpoint(double r, double res) {
  const uint32_t e = top12(r);
  if (e < 53 || e >= 0x7fe) return false;
  if ((as_u64(r) & 0xfffffffffffffull) == 0) return false;
  const double ulp = as_f64((uint64_t)(e - 52) << 52);
  return __builtin_fabs(res) < 0.4375 * ulp;
}

// glibc's pow(x, y) for finite x > 0, evaluated from exact products for the
// exponents 1, 2 and 3 where that is glibc's result (pow_shortcut says whether
// it is).  glibc's pow is within
// 0.54 ulp of the exact x^y (its source bounds the error of exp by 0.509 ulp
// and the log's contribution by 1024 ln2 1.3 2^-68 2^53 < 0.03 ulp), so when
// the exact x^y lies more than 0.04 ulp from a rounding midpoint glibc returns
// the correctly rounded x^y.  x^1 is x.  x^2 = hi + lo exactly with
// lo = fma(x, x, -hi).  x^3 = x hi + x lo, where x hi = p + pe exactly and the
// rounded remainder t = pe + x lo is within 2^-100 of x^3 of the exact one, so
// r = p + t rounds like x^3 and res = (p - r) + t is its residual.  Results
// within 1/16 ulp of a midpoint, and everything else, take glibc_pow.
AMCA_LIBM_HD inline bool pow_shortcut(double x, double y, double* out) {
  if (top12(x) - 1u >= 0x7feu) return false;  // x not normal
  if (y == 1.0) {
    *out = x;
    return true;
  }
  if (y == 2.0) {
    const double hi = x * x;
    const double lo = __builtin_fma(x, x, -hi);
    *out = hi;
    return clear_of_midpoint(hi, lo);
  }
  if (y == 3.0) {
    const double h2 = x * x;
    const double l2 = __builtin_fma(x, x, -h2);
    const double p = x * h2;
    const double pe = __builtin_fma(x, h2, -p);
    const double t = pe + x * l2;
    const double r = p + t;
    const double res = (p - r) + t;
    *out = r;
    return clear_of_midpoint(r, res);
  }
  return false;
}

AMCA_LIBM_HD inline double glibc_pow_fast(double x, double y) {
  double r;
  if (pow_shortcut(x, y, &r)) return r;
  return glibc_pow(x, y);
}

}  // namespace amca_libm
