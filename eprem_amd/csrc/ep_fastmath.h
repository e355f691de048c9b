// ep_fastmath.h -- table-driven FP64 exp and log for the adiabatic-change
// kernel's fused precision modes (the reference-order mode keeps the
// device libm).  Both follow P. T. P. Tang's table-driven methods (ACM
// TOMS 15(2) 1989 for exp, 16(4) 1990 for log) with tables held in LDS:
//
//   exp(x) = 2^m * 2^(j/64) * exp(r),  x = (64m + j) ln2/64 + r,
//            |r| <= ln2/128, exp(r) - 1 by a degree-6 polynomial
//            (truncation < 4e-18);
//   log(x) = m ln2 + ln F + ln(1 + u),  x = 2^m Y, Y in [1, 2),
//            F = 1 + j/128 nearest to Y, u = (Y - F)/F, |u| <= 1/256,
//            ln(1 + u) by a degree-6 polynomial (truncation < 3e-18),
//            ln F and ln 2 each split into a high part on a 2^-42 grid
//            and the rest.
//
// The functions are __host__ __device__ and use only correctly rounded
// operations (fma, rint, frexp, ldexp, integer ops), so the host build
// gives the same bits as the device one; the library exports the host
// form for the error tests.  Inputs: exp for any x (NaN in, NaN out;
// overflow to inf, underflow through ldexp), log for x > 0 (inf -> inf,
// NaN -> NaN); the change operator only takes logs of floored f.
#pragma once

#include <stdint.h>

#include <cmath>

namespace ep {

constexpr int kExpN = 64;    // 2^(j/64), j = 0..63
constexpr int kLogN = 128;   // F = 1 + j/128, j = 0..128
// table layout (doubles): exp hi[64], exp lo[64], log 1/F[129],
// ln F hi[129], ln F lo[129]
constexpr int kFmExpHi = 0, kFmExpLo = kExpN, kFmLogInv = 2 * kExpN, kFmLogHi = kFmLogInv + kLogN + 1,
              kFmLogLo = kFmLogHi + kLogN + 1, kFmTable = kFmLogLo + kLogN + 1;

__host__ __device__ inline uint32_t fm_hi_word(double x) {
  return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32);
}

// exp(x); t = the table.  k = round(64 x / ln 2) is clamped before the
// integer conversion (defined for every x, NaN included); with the
// clamped k, x = +inf and every finite x above 709.79 reach infinity
// through the polynomial or ldexp, x below -745.2 (where the clamped
// reduction can overflow the polynomial) and -inf give 0, and NaN stays NaN.
__host__ __device__ inline double fm_exp(double x, const double *t) {
  const double inv = 92.332482616893656768;          // 64 / ln 2
  const double l1 = 0.010830424696223417;            // ln2/64 to 36 bits: kd * l1 exact
  const double l2 = 2.572804622327669e-14;           // ln2/64 - l1
  const double kd = fmin(fmax(rint(x * inv), -80000.0), 80000.0);
  const int k = (int)kd;
  const double r = fma(-kd, l2, fma(-kd, l1, x));
  const int j = k & (kExpN - 1);
  const int m = k >> 6;  // floor(k / 64): k = 64 m + j
  double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  q = fma(r, q, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  const double p = fma(r * r, q, r);  // exp(r) - 1
  const double th = t[kFmExpHi + j], tl = t[kFmExpLo + j];
  const double s = th + fma(th, p, tl);
  double y = ldexp(s, m);  // exact scaling, or correctly rounded into the subnormals / inf
  if (x < -745.2) y = 0.0;
  return y;
}

// log(x) for x > 0 (normal or not).  ln F's high part is a multiple of
// 2^-42 and so is ln2_hi, so m ln2_hi + ln F_hi is exact for every
// exponent m (Tang's split) and needs no error-free sum.
__host__ __device__ inline double fm_log(double x, const double *t) {
  const double ln2_hi = 0.6931471805598903;           // ln 2 to a multiple of 2^-42
  const double ln2_lo = 5.497923018708371e-14;        // ln 2 - ln2_hi
  int e;
  const double y = frexp(x, &e);  // x = y 2^e, y in [0.5, 1)
  const double Y = y + y;         // [1, 2)
  const int m = e - 1;
  const uint32_t hw = fm_hi_word(Y);
  const int j = (int)(((hw & 0xFFFFFu) + (1u << 12)) >> 13);  // nearest j/128
  const double F = fma((double)j, 1.0 / kLogN, 1.0);
  const double f = Y - F;  // exact
  const double u = f * t[kFmLogInv + j];
  double q = fma(u, -1.0 / 6.0, 1.0 / 5.0);
  q = fma(u, q, -0.25);
  q = fma(u, q, 1.0 / 3.0);
  q = fma(u, q, -0.5);
  const double p = fma(u * u, q, u);  // ln(1 + u)
  const double md = (double)m;
  const double hi = fma(md, ln2_hi, t[kFmLogHi + j]);  // exact
  const double lo = fma(md, ln2_lo, t[kFmLogLo + j]) + p;
  double r = hi + lo;
  if (x == INFINITY) r = INFINITY;
  if (x != x) r = x;
  return r;
}

}  // namespace ep
