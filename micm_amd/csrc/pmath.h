// pmath.h -- portable double-precision exp, log, log10 and pow.
//
// Built only from IEEE-754 +, -, *, / and exact bit manipulation, in a fixed
// operation order, so the same source compiled for the host (g++) and for gfx950
// (hipcc), both without FMA contraction, returns bit-identical results.  The rate
// constants and the step-size controller of the device path use these functions;
// the CPU oracle can be built with them as well, which makes whole solves
// comparable bit for bit.
//
// Each function is accurate to about 0.51 ulp (tests/test_pmath.py measures it
// against 60-digit values), so its result is the correctly rounded one -- and the
// same as a good C library's -- except in rare near-halfway cases.
//
// Methods (published):
//   exp: table-driven reduction after P. T. P. Tang, "Table-driven implementation of
//        the exponential function in IEEE floating-point arithmetic", ACM TOMS 15
//        (1989): x = (128 k + j) ln2/128 + r, |r| <= ln2/256, exp(x) = 2^k 2^(j/128)
//        (1 + p(r)) with 2^(j/128) held as a sum of two doubles and p the degree-5
//        Taylor polynomial of exp(r) - 1.
//   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), s = (m - 1)/(m + 1) carried to
//        double-double precision with Dekker's exact product (T. J. Dekker, "A
//        floating-point technique for extending the available precision", Numer.
//        Math. 18 (1971)), log m = 2 atanh(s) = 2s + 2s^3/3 + ... ; the result is
//        a double-double e ln2 + log m.
//   log10: fdlibm's decomposition (see log10 below) over this log.
//   pow: exp(y log|x|) with y log|x| formed in double-double and its low part
//        folded into exp's reduced argument; IEEE special cases first.
#pragma once

#include <cstdint>

#include "pmath_table.h"

#if defined(__HIPCC__)
#define PM_FN __host__ __device__ inline
#else
#define PM_FN inline
#endif

namespace pmath
{
#if defined(__HIP_DEVICE_COMPILE__)
  __constant__ static const double kExp2Hi[128] = { PMATH_EXP2_TABLE_HI };
  __constant__ static const double kExp2Lo[128] = { PMATH_EXP2_TABLE_LO };
#else
  static const double kExp2Hi[128] = { PMATH_EXP2_TABLE_HI };
  static const double kExp2Lo[128] = { PMATH_EXP2_TABLE_LO };
#endif

  PM_FN double from_bits(uint64_t b)
  {
    return __builtin_bit_cast(double, b);
  }

  PM_FN uint64_t to_bits(double d)
  {
    return __builtin_bit_cast(uint64_t, d);
  }

  PM_FN bool is_nan(double x)
  {
    return x != x;
  }

  constexpr double kInf = __builtin_huge_val();
  constexpr double kLn2Hi = 6.93147180369123816490e-01;  // ln 2 to 31 significant bits: e * kLn2Hi is exact
  constexpr double kLn2Lo = 1.9082149292705877e-10;      // ln 2 - kLn2Hi
  constexpr double k128InvLn2 = 184.6649652337873;       // 128 / ln 2
  constexpr double kLn2Over128Hi = 0.00541521234663378;  // kLn2Hi / 128 (exact)
  constexpr double kLn2Over128Lo = 1.4907929134926466e-12;
  constexpr double kInvLn10Hi = 0.4342944819032518;  // 1 / ln 10
  constexpr double kLog10Of2Hi = 3.01029995663611771306e-01;  // log10(2) to 33 significant bits
  constexpr double kLog10Of2Lo = 3.69423907715893078616e-13;  // log10(2) - kLog10Of2Hi
  constexpr double kExpMax = 709.78271289338397;   // largest x with exp(x) finite
  constexpr double kExpMin = -745.13321910194122;  // below: exp(x) rounds to zero

  // 2^k for -1022 <= k <= 1023
  PM_FN double pow2i(int k)
  {
    return from_bits(static_cast<uint64_t>(k + 1023) << 52);
  }

  // a + b = s + err exactly (Knuth)
  PM_FN double two_sum(double a, double b, double& err)
  {
    const double s = a + b;
    const double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
  }

  // a * b = p + err exactly (Dekker, with Veltkamp's split; |a|, |b| < 2^996)
  PM_FN double two_prod(double a, double b, double& err)
  {
    const double p = a * b;
    const double ca = 134217729.0 * a;  // 2^27 + 1
    const double ah = ca - (ca - a);
    const double al = a - ah;
    const double cb = 134217729.0 * b;
    const double bh = cb - (cb - b);
    const double bl = b - bh;
    err = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
    return p;
  }

  // exp(x + xl) for kExpMin <= x <= kExpMax and |xl| <= 1e-12 |x|
  PM_FN double exp_core(double x, double xl)
  {
    const double t = x * k128InvLn2;
    const int n = static_cast<int>(t < 0.0 ? t - 0.5 : t + 0.5);
    const double nd = static_cast<double>(n);
    const double r = ((x - nd * kLn2Over128Hi) - nd * kLn2Over128Lo) + xl;
    const int j = n & 127;
    const int k = (n - j) / 128;
    // exp(r) - 1, |r| <= ln2/256: the degree-6 term is below 1e-18 relative
    double q = 1.0 / 120.0;
    q = 1.0 / 24.0 + r * q;
    q = 1.0 / 6.0 + r * q;
    q = 0.5 + r * q;
    const double p = r + (r * r) * q;
    const double th = kExp2Hi[j];
    const double m = th + (kExp2Lo[j] + th * p);
    if (k < -1000)
      return (m * pow2i(k + 1000)) * pow2i(-1000);
    if (k > 1000)
      return (m * pow2i(k - 1000)) * pow2i(1000);
    return m * pow2i(k);
  }

  PM_FN double exp(double x)
  {
    if (is_nan(x))
      return x;
    if (x > kExpMax)
      return kInf;
    if (x < kExpMin)
      return 0.0;
    return exp_core(x, 0.0);
  }

  // log(x) = hi + lo for finite x > 0, |lo| <= ulp(hi)
  PM_FN double log_dd(double x, double& lo)
  {
    int e = 0;
    if (x < 2.2250738585072014e-308)  // subnormal: scale into the normal range
    {
      x *= 18014398509481984.0;  // 2^54
      e = -54;
    }
    const uint64_t b = to_bits(x);
    e += static_cast<int>((b >> 52) & 0x7ff) - 1023;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m > 1.41421356237309504880)
    {
      m = m * 0.5;
      e += 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    double u_lo;
    const double u = two_sum(m, 1.0, u_lo);
    // s = f / (u + u_lo) as s + s_lo
    const double s = f / u;
    double pe;
    const double p = two_prod(s, u, pe);
    const double s_lo = (((f - p) - pe) - s * u_lo) / u;
    // 2 atanh(s) = 2s + 2s (s^2/3 + s^4/5 + ... + s^22/23); |s| <= 0.1716
    const double s2 = s * s;
    double q = 1.0 / 23.0;
    q = 1.0 / 21.0 + s2 * q;
    q = 1.0 / 19.0 + s2 * q;
    q = 1.0 / 17.0 + s2 * q;
    q = 1.0 / 15.0 + s2 * q;
    q = 1.0 / 13.0 + s2 * q;
    q = 1.0 / 11.0 + s2 * q;
    q = 1.0 / 9.0 + s2 * q;
    q = 1.0 / 7.0 + s2 * q;
    q = 1.0 / 5.0 + s2 * q;
    q = 1.0 / 3.0 + s2 * q;
    const double tail = (2.0 * s) * (s2 * q);
    const double ed = static_cast<double>(e);
    double l0;
    const double h = two_sum(ed * kLn2Hi, 2.0 * s, l0);
    const double l = l0 + ((ed * kLn2Lo + 2.0 * s_lo) + tail);
    const double hi = h + l;
    lo = l - (hi - h);
    return hi;
  }

  PM_FN double log(double x)
  {
    if (is_nan(x))
      return x;
    if (x < 0.0)
      return __builtin_nan("");
    if (x == 0.0)
      return -kInf;
    if (x == kInf)
      return kInf;
    double lo;
    return log_dd(x, lo);
  }

  // log10 as in fdlibm's e_log10.c (Sun Microsystems, 1993): x = 2^n x' with
  // x' in [1, 2) for n >= 0 and in [1/2, 1) for n < 0, then
  // log10(x) = (n log10(2)_lo + log(x') / ln 10) + n log10(2)_hi.  With a log that
  // is correctly rounded in all but rare cases this reproduces C libraries built
  // on that method bit for bit, and log10(10^k) == k for k = 0 .. 22.
  PM_FN double log10(double x)
  {
    if (is_nan(x) || x <= 0.0 || x == kInf)
      return log(x);
    int n = 0;
    if (x < 2.2250738585072014e-308)
    {
      x *= 18014398509481984.0;  // 2^54
      n = -54;
    }
    const uint64_t b = to_bits(x);
    n += static_cast<int>((b >> 52) & 0x7ff) - 1023;
    const int neg = n < 0 ? 1 : 0;
    const double xr = from_bits((b & 0x000fffffffffffffull) | (static_cast<uint64_t>(1023 - neg) << 52));
    const double nd = static_cast<double>(n + neg);
    double lo;
    const double z = nd * kLog10Of2Lo + kInvLn10Hi * log_dd(xr, lo);
    return z + nd * kLog10Of2Hi;
  }

  PM_FN bool is_integer(double y)
  {
    if (!(y > -4503599627370496.0 && y < 4503599627370496.0))  // |y| >= 2^52: an integer (or inf)
      return y == y && y != kInf && y != -kInf;
    return static_cast<double>(static_cast<int64_t>(y)) == y;
  }

  PM_FN bool is_odd_integer(double y)
  {
    if (!(y > -9007199254740992.0 && y < 9007199254740992.0))
      return false;
    const int64_t i = static_cast<int64_t>(y);
    return static_cast<double>(i) == y && (i & 1) != 0;
  }

  PM_FN double pow(double x, double y)
  {
    if (y == 0.0 || x == 1.0)
      return 1.0;
    if (is_nan(x) || is_nan(y))
      return x + y;
    if (y == 1.0)
      return x;
    if (y == 2.0)
      return x * x;
    if (x == 0.0)
    {
      const double s0 = ((to_bits(x) >> 63) != 0 && is_odd_integer(y)) ? -1.0 : 1.0;
      return y > 0.0 ? s0 * 0.0 : s0 * kInf;
    }
    double sign = 1.0;
    if (x < 0.0)
    {
      if (y == kInf || y == -kInf)
        return (x == -1.0) ? 1.0 : ((x > -1.0) == (y > 0.0) ? 0.0 : kInf);
      if (!is_integer(y))
        return __builtin_nan("");
      if (is_odd_integer(y))
        sign = -1.0;
      x = -x;
    }
    if (x == kInf)
      return y > 0.0 ? sign * kInf : sign * 0.0;
    if (y == kInf)
      return x > 1.0 ? kInf : 0.0;
    if (y == -kInf)
      return x > 1.0 ? 0.0 : kInf;
    double l;
    const double h = log_dd(x, l);
    const double t = y * h;
    if (!(t < 1000.0))
      return sign * kInf;
    if (!(t > -1000.0))
      return sign * 0.0;
    double pe;
    const double p = two_prod(y, h, pe);
    double wl;
    const double w = two_sum(p, pe + y * l, wl);
    if (w > kExpMax)
      return sign * kInf;
    if (w < kExpMin)
      return sign * 0.0;
    return sign * exp_core(w, wl);
  }
}  // namespace pmath
