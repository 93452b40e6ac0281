// pmath.h -- double-precision exp, log, log10 and pow that return, bit for bit,
// what the reference's CPU solver gets from its platform C library.
//
// The reference calls std::exp, std::pow and std::log10 (rate constants,
// include/micm/process/rate_constant/rate_constant_functions.hpp:51-126, and the
// step-size controller, include/micm/solver/rosenbrock.inl:212-216).  On x86-64
// Linux those are the GNU C library's (2.28 and later), whose exp, log and pow are
// ARM Optimized Routines' algorithms (math/exp.c, math/log.c, math/pow.c; MIT OR
// Apache-2.0 WITH LLVM-exception), selected in the variant built with fused
// multiply-add on CPUs that have it; log10 is fdlibm's decomposition (e_log10.c,
// Sun Microsystems 1993) over that log.  This header restates those algorithms --
// same reductions, same tables (pmath_table.h, generated from their defining
// formulas by gen_pmath_table.py), same polynomial coefficients and the same
// rounding of every operation, with each multiply-add that the C library's build
// fuses written as an explicit fma -- so the host build (g++) and the gfx950 build
// (hipcc) both return exactly the library's results; everything else is compiled
// without FMA contraction.  tests/test_pmath.py checks each function bit for bit
// against the C library on millions of arguments and against 60-digit values.
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
#define PM_TABLE __constant__ static const
#else
#define PM_TABLE static const
#endif
  PM_TABLE uint64_t kExpTable[256] = { PMATH_EXP_TABLE };
  PM_TABLE double kPowLogTable[128][3] = { PMATH_POW_LOG_TABLE };
  PM_TABLE double kLogTable[128][2] = { PMATH_LOG_TABLE };
#undef PM_TABLE

  PM_FN double from_bits(uint64_t b)
  {
    return __builtin_bit_cast(double, b);
  }

  PM_FN uint64_t to_bits(double d)
  {
    return __builtin_bit_cast(uint64_t, d);
  }

  PM_FN uint32_t top12(double x)
  {
    return static_cast<uint32_t>(to_bits(x) >> 52);
  }

  PM_FN double fma(double a, double b, double c)
  {
    return __builtin_fma(a, b, c);
  }

  constexpr double kInf = __builtin_huge_val();

  // exp: x = k ln2/128 + r, exp(x) = 2^(k/128) exp(r), 2^(k/128) = scale (1 + tail)
  constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128;
  constexpr double kNegLn2HiN = -0x1.62e42fefa0000p-8;
  constexpr double kNegLn2LoN = -0x1.cf79abc9e3b3ap-47;
  constexpr double kShift = 0x1.8p52;
  constexpr double kExpC2 = 0x1.ffffffffffdbdp-2;
  constexpr double kExpC3 = 0x1.555555555543cp-3;
  constexpr double kExpC4 = 0x1.55555cf172b91p-5;
  constexpr double kExpC5 = 0x1.1111167a4d017p-7;
  constexpr uint64_t kSignBias = 0x800ull << 7;

  // |x| >= 512: scale's exponent may leave the normal range, so form the result
  // with a shifted scale (and round once into the subnormal range).
  PM_FN double exp_special(double tmp, uint64_t sbits, uint64_t ki)
  {
    if ((ki & 0x80000000u) == 0)
    {
      const double scale = from_bits(sbits - (1009ull << 52));
      return 0x1p1009 * fma(scale, tmp, scale);
    }
    const double scale = from_bits(sbits + (1022ull << 52));
    const double st = scale * tmp;  // not fused: it has a second use below
    double y = scale + st;
    if ((y < 0.0 ? -y : y) < 1.0)
    {
      const double one = y < 0.0 ? -1.0 : 1.0;
      double lo = (scale - y) + st;
      const double hi = one + y;
      lo = ((one - hi) + y) + lo;
      y = (hi + lo) - one;
      if (y == 0.0)
        y = from_bits((sbits + (1022ull << 52)) & 0x8000000000000000ull);
    }
    return 0x1p-1022 * y;
  }

  // exp(x + xtail) with the sign of sign_bias (pow's odd powers of negative x)
  PM_FN double exp_core(double x, double xtail, uint64_t sign_bias)
  {
    uint32_t abstop = top12(x) & 0x7ff;
    if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54))
    {
      if (abstop - top12(0x1p-54) >= 0x80000000u)  // |x| < 2^-54
      {
        const double one = 1.0 + x;
        return sign_bias ? -one : one;
      }
      if (abstop >= top12(1024.0))
      {
        const double s = sign_bias ? -1.0 : 1.0;
        return (to_bits(x) >> 63) ? s * 0.0 : s * kInf;
      }
      abstop = 0;  // 512 <= |x| < 1024
    }
    double kd = fma(kInvLn2N, x, kShift);
    const uint64_t ki = to_bits(kd);
    kd -= kShift;
    double r = fma(kd, kNegLn2LoN, fma(kd, kNegLn2HiN, x));
    r += xtail;
    const uint64_t idx = 2 * (ki % 128);
    const uint64_t top = (ki + sign_bias) << 45;
    const double tail = from_bits(kExpTable[idx]);
    const uint64_t sbits = kExpTable[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, kExpC5, kExpC4), fma(r2, fma(r, kExpC3, kExpC2), tail + r));
    if (abstop == 0)
      return exp_special(tmp, sbits, ki);
    const double scale = from_bits(sbits);
    return fma(scale, tmp, scale);
  }

  PM_FN double exp(double x)
  {
    if ((top12(x) & 0x7ff) >= top12(1024.0))
    {
      if (to_bits(x) == to_bits(-kInf))
        return 0.0;
      if ((top12(x) & 0x7ff) == 0x7ff)  // +inf or nan
        return 1.0 + x;
    }
    return exp_core(x, 0.0, 0);
  }

  // log: x = 2^k z, z in [0x1.6p-1, 0x1.6p0), log x = k ln2 + log c + log1p(z/c - 1)
  constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
  constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
  constexpr double kLogA[5] = { -0x1.0000000000001p-1, 0x1.555555551305bp-2, -0x1.fffffffeb459p-3,
                                0x1.999b324f10111p-3, -0x1.55575e506c89fp-3 };
  // log1p(r) near x = 1, |x - 1| < 0x1p-4
  constexpr double kLogB[11] = { -0x1p-1,
                                 0x1.5555555555577p-2,
                                 -0x1.ffffffffffdcbp-3,
                                 0x1.999999995dd0cp-3,
                                 -0x1.55555556745a7p-3,
                                 0x1.24924a344de3p-3,
                                 -0x1.fffffa4423d65p-4,
                                 0x1.c7184282ad6cap-4,
                                 -0x1.999eb43b068ffp-4,
                                 0x1.78182f7afd085p-4,
                                 -0x1.5521375d145cdp-4 };

  PM_FN double log(double x)
  {
    uint64_t ix = to_bits(x);
    const uint32_t top = static_cast<uint32_t>(ix >> 48);
    constexpr uint64_t kLo = 0x3fee000000000000ull;  // 1 - 2^-4
    constexpr uint64_t kHi = 0x3ff1090000000000ull;  // 1 + 0x1.09p-4
    if (ix - kLo < kHi - kLo)
    {
      if (ix == 0x3ff0000000000000ull)
        return 0.0;
      const double r = x - 1.0;
      const double r2 = r * r;
      const double r3 = r * r2;
      const double q3 = fma(r3, kLogB[10], fma(r2, kLogB[9], fma(r, kLogB[8], kLogB[7])));
      const double q2 = fma(r3, q3, fma(r2, kLogB[6], fma(r, kLogB[5], kLogB[4])));
      const double q1 = fma(r3, q2, fma(r2, kLogB[3], fma(r, kLogB[2], kLogB[1])));
      // r r / 2 in double-double: r = rhi + rlo with rhi of 26 bits
      const double w0 = r * 0x1p27;
      const double rhi = (r + w0) - w0;
      const double rlo = r - rhi;
      const double w = (rhi * rhi) * kLogB[0];
      const double hi = r + w;
      double lo = (r - hi) + w;
      lo = fma(kLogB[0] * rlo, rhi + r, lo);
      return fma(r3, q1, lo) + hi;
    }
    if (top - 0x0010 >= 0x7ff0 - 0x0010)
    {
      if (ix * 2 == 0)
        return -kInf;
      if (ix == to_bits(kInf))
        return x;
      if ((top & 0x8000) || (top & 0x7ff0) == 0x7ff0)
        return (x - x) / (x - x);
      ix = to_bits(x * 0x1p52) - (52ull << 52);  // subnormal
    }
    const uint64_t tmp = ix - 0x3fe6000000000000ull;
    const int i = static_cast<int>((tmp >> 45) % 128);
    const int k = static_cast<int>(static_cast<int64_t>(tmp) >> 52);
    const double z = from_bits(ix - (tmp & (0xfffull << 52)));
    const double invc = kLogTable[i][0];
    const double logc = kLogTable[i][1];
    const double r = fma(z, invc, -1.0);
    const double kd = static_cast<double>(k);
    const double w = fma(kd, kLn2Hi, logc);
    const double hi = w + r;
    const double lo = fma(kd, kLn2Lo, (w - hi) + r);
    const double r2 = r * r;
    const double p = fma(r2, fma(r, kLogA[4], kLogA[3]), fma(r, kLogA[2], kLogA[1]));
    return fma(r * r2, p, fma(r2, kLogA[0], lo)) + hi;
  }

  // log10 as fdlibm's e_log10.c: x = 2^n x' with x' in [1, 2) for n >= 0 and in
  // [1/2, 1) for n < 0, log10(x) = (n log10(2)_lo + log(x') / ln 10) + n log10(2)_hi.
  constexpr double kInvLn10 = 4.34294481903251816668e-01;
  constexpr double kLog10Of2Hi = 3.01029995663611771306e-01;
  constexpr double kLog10Of2Lo = 3.69423907715893078616e-13;

  PM_FN double log10(double x)
  {
    uint64_t b = to_bits(x);
    int32_t hx = static_cast<int32_t>(b >> 32);
    int n = 0;
    if (hx < 0x00100000)
    {
      if (((hx & 0x7fffffff) | static_cast<uint32_t>(b)) == 0)
        return -kInf;
      if (hx < 0)
        return (x - x) / (x - x);
      n -= 54;
      x *= 0x1p54;  // subnormal
      b = to_bits(x);
      hx = static_cast<int32_t>(b >> 32);
    }
    if (hx >= 0x7ff00000)
      return x + x;
    n += (hx >> 20) - 1023;
    const int neg = n < 0 ? 1 : 0;
    hx = (hx & 0x000fffff) | ((0x3ff - neg) << 20);
    const double y = static_cast<double>(n + neg);
    const double xr = from_bits((static_cast<uint64_t>(static_cast<uint32_t>(hx)) << 32) | (b & 0xffffffffull));
    const double z = y * kLog10Of2Lo + kInvLn10 * log(xr);
    return z + y * kLog10Of2Hi;
  }

  // pow's logarithm: log|x| = hi + tail to about 2^-68 relative
  constexpr double kPowA[7] = { -0x1p-1,
                                0x1.555555555556p-2 * -2,
                                -0x1.0000000000006p-2 * -2,
                                0x1.999999959554ep-3 * 4,
                                -0x1.555555529a47ap-3 * 4,
                                0x1.2495b9b4845e9p-3 * -8,
                                -0x1.0002b8b263fc3p-3 * -8 };

  PM_FN double pow_log(uint64_t ix, double& tail)
  {
    const uint64_t tmp = ix - 0x3fe6955500000000ull;
    const int i = static_cast<int>((tmp >> 45) % 128);
    const int k = static_cast<int>(static_cast<int64_t>(tmp) >> 52);
    const double z = from_bits(ix - (tmp & (0xfffull << 52)));
    const double kd = static_cast<double>(k);
    const double invc = kPowLogTable[i][0];
    const double logc = kPowLogTable[i][1];
    const double logctail = kPowLogTable[i][2];
    const double r = fma(z, invc, -1.0);  // exact: 1/c has few bits
    const double t1 = fma(kd, kLn2Hi, logc);
    const double t2 = t1 + r;
    const double lo1 = fma(kd, kLn2Lo, logctail);
    const double lo2 = (t1 - t2) + r;
    const double ar = kPowA[0] * r;
    const double ar2 = r * ar;
    const double ar3 = r * ar2;
    const double hi = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = (t2 - hi) + ar2;
    const double q = fma(ar2, fma(ar2, fma(r, kPowA[6], kPowA[5]), fma(r, kPowA[4], kPowA[3])),
                         fma(r, kPowA[2], kPowA[1]));
    const double lo = fma(ar3, q, ((lo1 + lo2) + lo3) + lo4);
    const double y = hi + lo;
    tail = (hi - y) + lo;
    return y;
  }

  // 0: not an integer, 1: odd integer, 2: even integer
  PM_FN int check_int(uint64_t iy)
  {
    const int e = static_cast<int>((iy >> 52) & 0x7ff);
    if (e < 0x3ff)
      return 0;
    if (e > 0x3ff + 52)
      return 2;
    if (iy & ((1ull << (0x3ff + 52 - e)) - 1))
      return 0;
    if (iy & (1ull << (0x3ff + 52 - e)))
      return 1;
    return 2;
  }

  PM_FN bool zero_inf_nan(uint64_t i)
  {
    return 2 * i - 1 >= 2 * to_bits(kInf) - 1;
  }

  PM_FN double pow(double x, double y)
  {
    uint64_t sign_bias = 0;
    uint64_t ix = to_bits(x);
    const uint64_t iy = to_bits(y);
    uint32_t topx = top12(x);
    const uint32_t topy = top12(y);
    if (topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be)
    {
      // x subnormal, zero, negative, inf or nan; or |y| < 2^-65, |y| >= 2^63 or nan
      if (zero_inf_nan(iy))
      {
        if (2 * iy == 0)
          return 1.0;
        if (ix == 0x3ff0000000000000ull)
          return 1.0;
        if (2 * ix > 2 * to_bits(kInf) || 2 * iy > 2 * to_bits(kInf))
          return x + y;
        if (2 * ix == 2 * 0x3ff0000000000000ull)
          return 1.0;
        if ((2 * ix < 2 * 0x3ff0000000000000ull) == !(iy >> 63))
          return 0.0;
        return y * y;
      }
      if (zero_inf_nan(ix))
      {
        double x2 = x * x;
        if ((ix >> 63) && check_int(iy) == 1)
          x2 = -x2;
        return (iy >> 63) ? 1.0 / x2 : x2;
      }
      if (ix >> 63)
      {
        const int yint = check_int(iy);
        if (yint == 0)
          return (x - x) / (x - x);
        if (yint == 1)
          sign_bias = kSignBias;
        ix &= 0x7fffffffffffffffull;
        topx &= 0x7ff;
      }
      if ((topy & 0x7ff) - 0x3be >= 0x43e - 0x3be)
      {
        if (ix == 0x3ff0000000000000ull)
          return 1.0;
        if ((topy & 0x7ff) < 0x3be)
          return ix > 0x3ff0000000000000ull ? 1.0 + y : 1.0 - y;
        return ((ix > 0x3ff0000000000000ull) == (topy < 0x800)) ? kInf : 0.0;
      }
      if (topx == 0)
        ix = (to_bits(x * 0x1p52) & 0x7fffffffffffffffull) - (52ull << 52);  // subnormal
    }
    double lo;
    const double hi = pow_log(ix, lo);
    const double ehi = y * hi;
    const double elo = fma(y, lo, fma(y, hi, -ehi));
    return exp_core(ehi, elo, sign_bias);
  }
}  // namespace pmath
