// Device helpers shared by the HIP translation units.
#pragma once
#include "tc_common.h"
#include "tc_expf.h"

namespace tc {

// Order-preserving map of a float to uint32 (larger float -> larger key).
__device__ inline uint32_t float_order(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float float_unorder(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __uint_as_float(u);
}

// Split key: max over keys = max gain, ties to the lowest (feature, bin).
__device__ inline unsigned long long split_key(float gain, uint32_t feature, uint32_t bin) {
  const uint32_t lo = 0xFFFFFFFFu - (feature * 256u + bin);
  return (static_cast<unsigned long long>(float_order(gain)) << 32) | lo;
}

// Exp as the reference's AVX2 build evaluates it inside the sigmoid:
// round-to-nearest range reduction t = x*log2(e), degree-5 Taylor
// polynomial of 2^f in Horner form with fused multiply-adds, scaled by
// 2^round(t) through the exponent bits (src/boosting/loss.cpp:34-69).  The
// reference compiler fuses the final scale into the sigmoid's 1 + e, so this
// returns 1 + exp(x) with that single rounding.  t - round(t) is not fused
// (t has a second use), which needs the build's -ffp-contract=off.
__device__ inline float one_plus_exp_ref(float x) {
  x = fmaxf(x, -88.0f);
  x = fminf(x, 88.0f);
  const float t = __fmul_rn(x, 1.4426950408889634f);
  const float ti = rintf(t);
  const float tf = __fsub_rn(t, ti);
  float p = __fmaf_rn(0.0013333558146428443f, tf, 0.009618129107628477f);
  p = __fmaf_rn(p, tf, 0.05550410866482046f);
  p = __fmaf_rn(p, tf, 0.2402265069591007f);
  p = __fmaf_rn(p, tf, 0.6931471805599453f);
  p = __fmaf_rn(p, tf, 1.0f);
  const float pw = __uint_as_float(static_cast<uint32_t>(static_cast<int32_t>(ti) + 127) << 23);
  return __fmaf_rn(p, pw, 1.0f);
}

// A6/A7 gradient of one row (src/boosting/loss.cpp:123-163, :348-362).
// `vec_row` selects the reference's AVX2 formula (rows 0 .. 8*floor(N/8)-1);
// the scalar tail uses the library exp the reference calls (glibc's expf,
// ref_expf).
__device__ inline float2 row_gradient(int loss, float pred, float label, bool vec_row) {
  if (loss == kMSE) {
    return make_float2(__fmul_rn(2.0f, __fsub_rn(pred, label)), 2.0f);
  }
  float x = fmaxf(pred, -15.0f);
  x = fminf(x, 15.0f);
  float s;
  if (vec_row) {
    s = __fdiv_rn(1.0f, one_plus_exp_ref(__fsub_rn(0.0f, x)));
  } else {
    s = __fdiv_rn(1.0f, __fadd_rn(1.0f, ref_expf(-x)));
  }
  const float g = __fsub_rn(s, label);
  const float hh = __fmul_rn(s, __fsub_rn(1.0f, s));
  return make_float2(g, hh > 1e-6f ? hh : 1e-6f);
}

__device__ inline double warp_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
// One DPP move of a wave64 scan step: the value of the lane CTRL selects
// (0 where the control leaves a lane without a source, or outside ROW_MASK).
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ inline uint32_t dpp_mov(uint32_t v) {
  return __builtin_amdgcn_update_dpp(0u, v, CTRL, ROW_MASK, 0xf, BOUND);
}

// Wave64 inclusive scan on the DPP crossbar (no LDS traffic; every lane of
// the wave active): row_shr 1, 2, 4, 8 inside each 16-lane row, then the
// row_bcast:15 and row_bcast:31 broadcasts of the GFX9 family carry the row
// results on.  OP must be order-free with identity 0: an integer sum, or the
// max of unsigned or non-negative values.
template <typename Op>
__device__ inline uint32_t wave_scan_u32(uint32_t v, Op op) {
  v = op(v, dpp_mov<0x111, 0xf, true>(v));
  v = op(v, dpp_mov<0x112, 0xf, true>(v));
  v = op(v, dpp_mov<0x114, 0xf, true>(v));
  v = op(v, dpp_mov<0x118, 0xf, true>(v));
  v = op(v, dpp_mov<0x142, 0xa, false>(v));
  v = op(v, dpp_mov<0x143, 0xc, false>(v));
  return v;
}
__device__ inline uint32_t lane63(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}
__device__ inline uint32_t wave_incl_scan_u32(uint32_t v) {
  return wave_scan_u32(v, [](uint32_t a, uint32_t b) { return a + b; });
}
// Sum over the wave, in every lane.
__device__ inline uint32_t wave_sum_u32(uint32_t v) { return lane63(wave_incl_scan_u32(v)); }
// Max over the wave, in every lane.
__device__ inline uint32_t wave_max_u32(uint32_t v) {
  return lane63(wave_scan_u32(v, [](uint32_t a, uint32_t b) { return a > b ? a : b; }));
}
// Max of non-negative floats over the wave (fmaxf), in every lane.
__device__ inline float wave_max_f32(float v) {
  return __uint_as_float(lane63(wave_scan_u32(__float_as_uint(v), [](uint32_t a, uint32_t b) {
    return __float_as_uint(fmaxf(__uint_as_float(a), __uint_as_float(b)));
  })));
}

// 64-bit steps: both halves moved by the same control.
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ inline unsigned long long dpp_mov64(unsigned long long u) {
  const uint32_t lo = dpp_mov<CTRL, ROW_MASK, BOUND>(static_cast<uint32_t>(u));
  const uint32_t hi = dpp_mov<CTRL, ROW_MASK, BOUND>(static_cast<uint32_t>(u >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}
template <typename Op>
__device__ inline unsigned long long wave_scan_u64(unsigned long long v, Op op) {
  v = op(v, dpp_mov64<0x111, 0xf, true>(v));
  v = op(v, dpp_mov64<0x112, 0xf, true>(v));
  v = op(v, dpp_mov64<0x114, 0xf, true>(v));
  v = op(v, dpp_mov64<0x118, 0xf, true>(v));
  v = op(v, dpp_mov64<0x142, 0xa, false>(v));
  v = op(v, dpp_mov64<0x143, 0xc, false>(v));
  return v;
}
__device__ inline unsigned long long lane63_u64(unsigned long long v) {
  return (static_cast<unsigned long long>(lane63(static_cast<uint32_t>(v >> 32))) << 32) |
         lane63(static_cast<uint32_t>(v));
}
// Inclusive prefix sum of an int64 over the wave: exact in any order.
__device__ inline long long wave_incl_scan_i64(long long v) {
  return static_cast<long long>(wave_scan_u64(
      static_cast<unsigned long long>(v),
      [](unsigned long long a, unsigned long long b) { return a + b; }));
}
// Sum of an int64 over the wave, in every lane.
__device__ inline long long wave_sum_i64(long long v) {
  return static_cast<long long>(lane63_u64(static_cast<unsigned long long>(wave_incl_scan_i64(v))));
}
// Max of a u64 over the wave, in every lane.
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
  return lane63_u64(wave_scan_u64(v, [](unsigned long long a, unsigned long long b) {
    return a > b ? a : b;
  }));
}

// Rank a logical work index so that consecutive indices (the feature tiles
// of one row chunk) run on the same XCD back to back: workgroups are dealt
// round-robin to the 8 XCDs, so XCD x gets blockIdx x, x+8, ...  Returns
// false when this workgroup has no work.
__device__ inline bool xcd_rank(uint32_t block, uint32_t count, uint32_t* logical) {
  const uint32_t per = (count + 7u) >> 3;
  const uint32_t k = block >> 3, x = block & 7u;
  if (k >= per) return false;
  const uint32_t l = x * per + k;
  if (l >= count) return false;
  *logical = l;
  return true;
}

}  // namespace tc
