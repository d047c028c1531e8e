// Test driver (built by tests/test_libm_pow.py): the device velocity's
// restatement of glibc's pow (amca3d_amd/csrc/libm_pow.h), compiled for the
// host, counted against the host's own glibc pow on random arguments.
#include <math.h>
#include <stdint.h>

#include "../amca3d_amd/csrc/libm_pow.h"

using amca_libm::as_f64;
using amca_libm::as_u64;

namespace {

uint64_t g_state;
uint64_t next() {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return g_state;
}

// a random positive double with binary exponent in [lo_e, hi_e] (below -1022:
// subnormal)
double draw(int lo_e, int hi_e) {
  const int64_t e = lo_e + (int64_t)(next() % (uint64_t)(hi_e - lo_e + 1));
  const uint64_t m = next() & ((1ull << 52) - 1);
  if (e < -1022) return as_f64(m >> (-1022 - e));
  return as_f64(((uint64_t)(e + 1023) << 52) | m);
}

// distance, in ulps of the rounded result, of the exact x^y (y = 2 or 3) from
// the nearest rounding midpoint; x^2 = hi + lo exactly, x^3 = x*hi + x*lo with
// x*hi = p + e exactly and x*lo rounded (its error is ~2^-52 of lo)
double midpoint_distance(double x, int y) {
  double hi = x * x, lo = fma(x, x, -hi);
  if (y == 3) {
    const double p = x * hi, e = fma(x, hi, -p);
    hi = p;
    lo = e + x * lo;
  }
  const double ulp = nextafter(hi, INFINITY) - hi;
  return fabs(fabs(lo) / ulp - 0.5);
}

}  // namespace

extern "C" {

// Mismatches between glibc's pow(x, y) and the restatement over n arguments
// with exponents in [lo_e, hi_e].  hard > 0 keeps only arguments whose exact
// x^y (y 2 or 3) lies within `band` ulp of a rounding midpoint; naive counts
// how often the product form (x*x or x*x*x) misses glibc on the same arguments.
long pow_mismatches(long n, double y, int lo_e, int hi_e, uint64_t seed, int hard, double band,
                    long* naive) {
  g_state = 0x9e3779b97f4a7c15ull ^ (seed * 0xbf58476d1ce4e5b9ull) ^ 1;
  long bad = 0, kept = 0, nv = 0;
  while (kept < n) {
    const double x = draw(lo_e, hi_e);
    if (!(x > 0)) continue;
    if (hard && midpoint_distance(x, (int)y) > band) continue;
    kept++;
    const double ref = pow(x, y);
    if (as_u64(ref) != as_u64(amca_libm::glibc_pow(x, y))) bad++;
    const double prod = y == 2.0 ? x * x : y == 3.0 ? x * x * x : ref;
    if (as_u64(prod) != as_u64(ref)) nv++;
  }
  if (naive) *naive = nv;
  return bad;
}

// The same draws against glibc_pow_fast (the form the device velocity calls);
// *shortcut counts the arguments it answered from exact products.
long pow_fast_mismatches(long n, double y, int lo_e, int hi_e, uint64_t seed, int hard,
                         double band, long* shortcut) {
  g_state = 0x9e3779b97f4a7c15ull ^ (seed * 0xbf58476d1ce4e5b9ull) ^ 1;
  long bad = 0, kept = 0, sc = 0;
  while (kept < n) {
    const double x = draw(lo_e, hi_e);
    if (!(x > 0)) continue;
    if (hard && midpoint_distance(x, (int)y) > band) continue;
    kept++;
    double r;
    if (amca_libm::pow_shortcut(x, y, &r)) sc++;
    if (as_u64(pow(x, y)) != as_u64(amca_libm::glibc_pow_fast(x, y))) bad++;
  }
  if (shortcut) *shortcut = sc;
  return bad;
}

// The arguments pow_mismatches(n, ...) draws, written to out[n].
void pow_args(long n, double y, int lo_e, int hi_e, uint64_t seed, int hard, double band,
              double* out) {
  g_state = 0x9e3779b97f4a7c15ull ^ (seed * 0xbf58476d1ce4e5b9ull) ^ 1;
  long kept = 0;
  while (kept < n) {
    const double x = draw(lo_e, hi_e);
    if (!(x > 0)) continue;
    if (hard && midpoint_distance(x, (int)y) > band) continue;
    out[kept++] = x;
  }
}

// glibc's own pow(x[i], y) (overflow gives inf, as the C function does).
void glibc_pow_array(long n, const double* x, double y, double* out) {
  for (long i = 0; i < n; i++) out[i] = pow(x[i], y);
}

}  // extern "C"
