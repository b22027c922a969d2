// expf as the reference's CPU build evaluates it: every std::exp(float) of the
// reference (the LogLoss sigmoid, loss.hpp:99-101 and loss.cpp:157-162; the
// softmax, loss.cpp:273-285; the probability transforms, loss.cpp:213-215)
// resolves to glibc's expf (glibc >= 2.28, sysdeps/ieee754/flt-32/e_expf.c,
// the algorithm published in ARM's optimized-routines).  That algorithm:
//   k = round-to-nearest(x * 32/ln2) by adding and subtracting 1.5 * 2^52,
//   r = x * 32/ln2 - k in [-1/2, 1/2];
//   exp(x) = 2^(k/32) * 2^(r/32), 2^(r/32) ~= 1 + C2 r + C1 r^2 + C0 r^3,
//   2^(k/32) = 2^(k mod 32 / 32) from a 32-entry table with the exponent
//   field raised by k >> 5; everything in double, rounded once to float.
// The table entries are the correctly rounded doubles 2^(i/32) with i << 47
// subtracted from their bits (so that adding k << 47 restores i's share of
// the exponent and adds k >> 5); the coefficients are the published minimax
// coefficients of 2^r on [-1/64, 1/64], scaled by 32^-3, 32^-2, 32^-1.
// Special cases as glibc: |x| >= 88 first checks -inf (0), inf / NaN
// (x + x), overflow above ln(2^128) (inf) and underflow below ln(2^-150) (0).
// Pinned against the host's glibc expf on every float (tests/test_expf.py).
// Host and device: the same source compiles under g++ for that check.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define TC_EXPF_HD __host__ __device__
#else
#define TC_EXPF_HD
#endif

namespace tc {

TC_EXPF_HD inline float ref_expf(float x) {
  constexpr uint64_t kTab[32] = {
      0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
      0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
      0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
      0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
      0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
      0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
      0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
      0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
  constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32;
  constexpr double kShift = 0x1.8p+52;
  constexpr double kC0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
  constexpr double kC1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
  constexpr double kC2 = 0x1.62e42ff0c52d6p-1 / 32;
  const uint32_t ux = __builtin_bit_cast(uint32_t, x);
  const uint32_t abstop = (ux >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) {  // |x| >= 88, inf or NaN
    if (ux == 0xff800000u) return 0.0f;
    if (abstop >= 0x7f8u) return x + x;
    if (x > 0x1.62e42ep6f) return __builtin_inff();
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  // the fused multiply-adds of glibc's FMA build (the one a CPU with FMA runs:
  // the reference's AVX2 build requires one), the range reduction included
  const double xd = static_cast<double>(x);
  double kd = __builtin_fma(kInvLn2N, xd, kShift);
  const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
  kd -= kShift;
  const double r = __builtin_fma(kInvLn2N, xd, -kd);
  const double s = __builtin_bit_cast(double, kTab[ki % 32] + (ki << 47));
  const double p = __builtin_fma(kC0, r, kC1);
  const double r2 = r * r;
  double y = __builtin_fma(kC2, r, 1.0);
  y = __builtin_fma(p, r2, y);
  return static_cast<float>(y * s);
}

}  // namespace tc
