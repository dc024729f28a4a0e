// Philox4x32-10 counter-based RNG and libm-free log/exp, usable on host and device.
//
// Replaces the clock-seeded streams of the reference (curand_rand_mat.cu:42 time(NULL),
// stpm.c:1123 srand(time(NULL)), CPU Spectral.cpp:120 srand(clock())). A draw is a pure function
// of (seed, stream, counter words), so results are independent of thread and block counts.
// Philox4x32-10 follows Salmon et al., SC'11 (constants and round structure from the paper).
// det_log/det_exp evaluate fixed series with IEEE +,-,*,/ and exact exponent manipulation only,
// so translation units compiled with -ffp-contract=off reproduce the oracle's bits.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define MMSB_HD __host__ __device__ inline
#else
#define MMSB_HD inline
#endif

namespace mmsb {

enum : uint32_t {
  kStreamTheta = 1,
  kStreamPow = 2,
  kStreamGamma = 3,
  kStreamSplit = 4,
  kStreamEdge = 5,
  kStreamBall = 6,
  kStreamOmega = 16,  // + index of the Gaussian test matrix
};

struct u32x4 {
  uint32_t x, y, z, w;
};

MMSB_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// Four words for counter (i0, i1, stream, sub) under key = seed.
MMSB_HD u32x4 draw(uint64_t seed, uint32_t stream, uint32_t i0, uint32_t i1, uint32_t sub) {
  return philox4x32_10(u32x4{i0, i1, stream, sub}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Uniform in [0, 1) with 53 significant bits.
MMSB_HD double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
// Uniform in (0, 1].
MMSB_HD double u53_open0(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 1.0) * (1.0 / 9007199254740992.0);
}

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kSqrtHalf = 0.70710678118654752440;

// log(x), x > 0 finite: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1).
MMSB_HD double det_log(double x) {
  int e;
  double m = std::frexp(x, &e);
  if (m < kSqrtHalf) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;  // 12 terms: 1/(2n+1), n = 11 .. 0
  for (int n = 10; n >= 0; --n) p = p * s2 + 1.0 / (double)(2 * n + 1);
  const double logm = (2.0 * s) * p;
  const double ed = (double)e;
  return ed * kLn2Hi + (ed * kLn2Lo + logm);
}

MMSB_HD double det_fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= (double)i;
  return f;
}

// exp(x): x = k ln2 + r, |r| <= ln2/2, 14-term Taylor series, exact ldexp.
MMSB_HD double det_exp(double x) {
  const double k = std::floor(x * kInvLn2 + 0.5);
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.0 / det_fact(13);
  for (int n = 12; n >= 0; --n) p = p * r + 1.0 / det_fact(n);
  double kc = k < -1100.0 ? -1100.0 : (k > 1100.0 ? 1100.0 : k);
  return std::ldexp(p, (int)kc);
}

}  // namespace mmsb
