// pow(x, y) of the C library, restated so the device reproduces it bitwise.
//
// The reference computes the new step size with one libm call,
// h * fmin(2, fmax(0.5, 0.9 * pow(tol / err, 1 / (ord + 1))))
// (include/tools.h:221-233).  An ulp of difference in that pow changes h,
// which the error estimate then amplifies until accepted/rejected sequences
// part over thousands of steps, so the step control needs the C library's
// exact result, not just an accurate one.
//
// The C library here is glibc 2.35, whose pow is the algorithm of ARM's
// optimized-routines math/pow.c (Szabolcs Nagy, 2018; glibc
// sysdeps/ieee754/dbl-64/e_pow.c): x = 2^k z, log(x) = k ln2 + log(c) +
// log1p(z/c - 1) with a 128-entry table of c near z and a degree-8
// polynomial, carried as hi + lo; then exp(y log x) = 2^(k/128) exp(r) with a
// 128-entry table of 2^(k/128) and a degree-5 polynomial.  On x86-64 CPUs
// with FMA and AVX2 glibc selects the build of that source compiled with
// -mfma, in which the compiler fuses every a*b + c whose product feeds only
// additions; the fma() calls below sit exactly there (and at the source's own
// __builtin_fma), and every other operation rounds as written (this file is
// compiled with -ffp-contract=off).  The tables are generated from their
// published definitions by gen_pow_tables.py.
//
// Domain: x >= 0 (subnormal, +0 and +inf included), 0 < y <= 1/2 -- the
// step control's 1/(ord+1) for ord >= 1.  Then |y log x| < 512, so exp's
// over/underflow branch of the original is never taken and is not restated.
// Checked bitwise against the C library: tests/test_pow.py (host evaluation
// of this same code) and tests/test_gpu_pow.py (device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rkx {

struct pow_tables {
  double logt[128][3];            // invc, logc, logctail
  unsigned long long expt[256];   // bits of tail T_k, bits of H_k - (k << 45)
};

__host__ __device__ inline double pow_asd(uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ inline uint64_t pow_asu(double x) { return __builtin_bit_cast(uint64_t, x); }

// log(x) as hi + *tail for the bits ix of a normal positive x (e_pow.c log_inline)
__host__ __device__ inline double pow_log_inline(uint64_t ix, double *tail, const pow_tables &T)
{
  constexpr double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
  // polynomial coefficients, scaled as the evaluation below uses them
  constexpr double A0 = -0x1p-1, A1 = 0x1.555555555556p-2 * -2,
                   A2 = -0x1.0000000000006p-2 * -2, A3 = 0x1.999999959554ep-3 * 4,
                   A4 = -0x1.555555529a47ap-3 * 4, A5 = 0x1.2495b9b4845e9p-3 * -8,
                   A6 = -0x1.0002b8b263fc3p-3 * -8;
  const uint64_t tmp = ix - 0x3fe6955500000000ULL;
  const int i = static_cast<int>((tmp >> 45) & 127u);
  const int k = static_cast<int>(static_cast<int64_t>(tmp) >> 52);
  const uint64_t iz = ix - (tmp & (0xfffULL << 52));
  const double z = pow_asd(iz);
  const double kd = static_cast<double>(k);
  const double invc = T.logt[i][0], logc = T.logt[i][1], logctail = T.logt[i][2];
  const double r = __builtin_fma(z, invc, -1.0);        // exact: 1/c has few bits
  const double t1 = __builtin_fma(kd, Ln2hi, logc);
  const double t2 = t1 + r;
  const double lo1 = __builtin_fma(kd, Ln2lo, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = A0 * r;
  const double ar2 = r * ar;
  const double ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = __builtin_fma(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  double q = __builtin_fma(ar2, __builtin_fma(r, A6, A5), __builtin_fma(r, A4, A3));
  q = __builtin_fma(ar2, q, __builtin_fma(r, A2, A1));
  const double lo = __builtin_fma(ar3, q, lo1 + lo2 + lo3 + lo4);
  const double y = hi + lo;
  *tail = hi - y + lo;
  return y;
}

// exp(x + xtail) for |x| < 512 (e_pow.c exp_inline, sign bias 0)
__host__ __device__ inline double pow_exp_inline(double x, double xtail, const pow_tables &T)
{
  constexpr double InvLn2N = 0x1.71547652b82fep0 * 128, NegLn2hiN = -0x1.62e42fefa0000p-8,
                   NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  constexpr double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3,
                   C4 = 0x1.55555cf172b91p-5, C5 = 0x1.1111167a4d017p-7;
  const uint32_t abstop = static_cast<uint32_t>(pow_asu(x) >> 52) & 0x7ffu;
  if (abstop < 0x3c9u)                 // |x| < 2^-54: exp(x) rounds to 1 + x
    return 1.0 + x;
  double kd = __builtin_fma(InvLn2N, x, Shift);   // x*N/ln2 rounded to an integer
  const uint64_t ki = pow_asu(kd);
  kd -= Shift;
  double r = __builtin_fma(kd, NegLn2loN, __builtin_fma(kd, NegLn2hiN, x));
  r += xtail;
  const uint64_t idx = 2u * (ki % 128u);
  const uint64_t top = ki << 45;
  const double tail = pow_asd(T.expt[idx]);
  const uint64_t sbits = T.expt[idx + 1] + top;
  const double r2 = r * r;
  const double p = __builtin_fma(r2 * r2, __builtin_fma(r, C5, C4),
                                 __builtin_fma(r2, __builtin_fma(r, C3, C2), tail + r));
  const double scale = pow_asd(sbits);
  return __builtin_fma(scale, p, scale);
}

// pow(x, y) for x >= 0 and 0 < y <= 1/2 (e_pow.c __pow on that domain)
__host__ __device__ inline double libm_pow(double x, double y, const pow_tables &T)
{
  uint64_t ix = pow_asu(x);
  if ((ix >> 52) - 1u >= 0x7feu) {     // zero, subnormal, inf or nan
    if (x == 0.0)
      return 0.0;
    if (ix == 0x7ff0000000000000ULL)
      return x;
    if (x != x)
      return x;
    ix = pow_asu(x * 0x1p52);          // subnormal: normalise, exponent in the bits
    ix &= 0x7fffffffffffffffULL;
    ix -= 52ULL << 52;
  }
  double lo;
  const double hi = pow_log_inline(ix, &lo, T);
  const double ehi = y * hi;
  const double elo = __builtin_fma(y, lo, __builtin_fma(y, hi, -ehi));
  return pow_exp_inline(ehi, elo, T);
}

}  // namespace rkx
