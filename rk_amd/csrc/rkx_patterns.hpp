// Compile-time zero patterns of the method tables the fused kernels are
// instantiated for, and small compile-time helpers shared by those kernels.
#pragma once

#include "rkx_device.hpp"

#include <type_traits>
#include <utility>

namespace rkx {

// Zero patterns of the compiled methods (bit i of row l: A[l][i] != 0), the
// occupancy (waves per SIMD) the fused kernels are register-allocated for,
// and whether the fused kernel runs whole trips of its unrolled row ring
// (full_trips: no exit inside a trip, band heights rounded so that a band's
// row iterations fill whole trips).  Exits inside a trip keep more values
// live: VERNER65 needs 209 VGPRs without them and more than 256 with them,
// i.e. two waves per SIMD instead of one; DOPRI54 fits two waves either way
// and keeps the exits, which spare it the rounding of its band heights.
struct pat_dopri54 {
  static constexpr bool full_trips = false;
  static constexpr int occ = 2;
  static constexpr int s = 7;
  static constexpr unsigned rowA[7] = {0x0, 0x1, 0x3, 0x7, 0xF, 0x1F, 0x3D};
  static constexpr unsigned b = 0x3D, e = 0x7D;
  static constexpr bool fsal = true;      // A[6][j] == b[j]
};
struct pat_verner65 {
  static constexpr bool full_trips = true;
  static constexpr int occ = 2;
  static constexpr int s = 8;
  static constexpr unsigned rowA[8] = {0x0, 0x1, 0x3, 0x7, 0xF, 0x1F, 0x1F, 0x5F};
  static constexpr unsigned b = 0x3D, e = 0xFD;
  static constexpr bool fsal = false;
};
struct pat_rkf23 {
  static constexpr bool full_trips = false;
  static constexpr int occ = 2;
  static constexpr int s = 3;
  static constexpr unsigned rowA[3] = {0x0, 0x1, 0x3};
  static constexpr unsigned b = 0x3, e = 0x7;
  static constexpr bool fsal = false;
};
struct pat_heun2 {
  static constexpr bool full_trips = false;
  static constexpr int occ = 2;
  static constexpr int s = 2;
  static constexpr unsigned rowA[2] = {0x0, 0x1};
  static constexpr unsigned b = 0x3, e = 0x3;
  static constexpr bool fsal = false;
};
struct pat_rk4 {
  static constexpr bool full_trips = false;
  static constexpr int occ = 2;
  static constexpr int s = 4;
  static constexpr unsigned rowA[4] = {0x0, 0x1, 0x2, 0x4};
  static constexpr unsigned b = 0xF, e = 0xF;
  static constexpr bool fsal = false;
};
// DOPRI87 (13 stages): its register ring needs 256 VGPRs and about 130 AGPRs,
// one wave per SIMD, so strips of at most 2 waves per component; whole trips
// (the exit test inside a trip would cost more registers still).
struct pat_dopri87 {
  static constexpr bool full_trips = true;
  static constexpr int occ = 1;
  static constexpr int s = 13;
  static constexpr unsigned rowA[13] = {0x0,  0x1,  0x3,   0x5,   0xD,   0x19,  0x39,
                                        0x79, 0xF9, 0x1F9, 0x3F9, 0x7F9, 0x7F9};
  static constexpr unsigned b = 0x1FE1, e = 0x1FE1;
  static constexpr bool fsal = false;
};
struct pat_ssprk3 {
  static constexpr bool full_trips = false;
  static constexpr int occ = 2;
  static constexpr int s = 3;
  static constexpr unsigned rowA[3] = {0x0, 0x1, 0x3};
  static constexpr unsigned b = 0x7, e = 0x7;
  static constexpr bool fsal = false;
};

// y rows the fused kernel loads ahead of its first stage (bounded by the ring)
constexpr int prefetch_rows(int s) { return (s - 2 < 2) ? s - 2 : 2; }

// Row iterations at the start of a band that the fused kernel runs outside
// its loop (the first 2(s-1), in which the stages whose rows no stored row
// needs are left out); none for the 13-stage tables, whose peeled code would
// be too long.  Tables whose kernel runs whole trips of the ring then fill
// whole trips with the rest of the band.
constexpr int peeled_iters(bool full_trips, int s) { return (full_trips && s > 9) ? 0 : 2 * (s - 1); }

// compile-time loop: f(integral_constant<int, I>) for I = 0 .. N-1, so ring
// slots and stage indices are constants and the register ring stays in VGPRs
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>)
{
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <class P>
__device__ __forceinline__ constexpr bool nzA(int l, int i)
{
  return (P::rowA[l] >> i) & 1u;
}


// Which compiled pattern a method table has (0: none -- the fused kernels do
// not cover it and the stage kernels run instead).  Zero coefficients are
// skipped by the fused kernels (as by the reference's tiled stage kernels,
// block-inline.h:324-428), which leaves every sum bitwise unchanged.
inline int pattern_of(const rkx_method &m)
{
  // 1 DOPRI54, 2 VERNER65, 3 RKF23, 4 HEUN2, 5 RK4, 6 SSPRK3, 7 DOPRI87; 0 none
  auto rowmask = [&](int l) {
    unsigned r = 0;
    for (int i = 0; i < l; ++i)
      if (m.A[l][i] != 0.0)
        r |= 1u << i;
    return r;
  };
  auto vmask = [&](const double *v, const double *w) {
    unsigned r = 0;
    for (int i = 0; i < m.s; ++i)
      if ((w ? v[i] - w[i] : v[i]) != 0.0)
        r |= 1u << i;
    return r;
  };
  auto match = [&](int s, const unsigned *rows, unsigned b, unsigned e) {
    if (m.s != s)
      return false;
    for (int l = 0; l < s; ++l)
      if (rowmask(l) != rows[l])
        return false;
    return vmask(m.b, nullptr) == b && vmask(m.b, m.b_hat) == e;
  };
  bool fsal = true;
  for (int j = 0; j < m.s; ++j)
    fsal = fsal && (m.A[m.s - 1][j] == m.b[j] || j == m.s - 1);
  if (match(7, pat_dopri54::rowA, pat_dopri54::b, pat_dopri54::e) && fsal) return 1;
  if (match(8, pat_verner65::rowA, pat_verner65::b, pat_verner65::e)) return 2;
  if (match(3, pat_rkf23::rowA, pat_rkf23::b, pat_rkf23::e)) return 3;
  if (match(2, pat_heun2::rowA, pat_heun2::b, pat_heun2::e)) return 4;
  if (match(4, pat_rk4::rowA, pat_rk4::b, pat_rk4::e)) return 5;
  if (match(3, pat_ssprk3::rowA, pat_ssprk3::b, pat_ssprk3::e)) return 6;
  if (match(13, pat_dopri87::rowA, pat_dopri87::b, pat_dopri87::e)) return 7;
  return 0;
}


// Calls f(pat_X{}) for the compiled pattern id 1..6 (the tables the LV chain
// kernel is built for; see chain_supported).
template <class F>
inline hipError_t dispatch_pattern(int pattern, F &&f)
{
  switch (pattern) {
  case 1: return f(pat_dopri54{});
  case 2: return f(pat_verner65{});
  case 3: return f(pat_rkf23{});
  case 4: return f(pat_heun2{});
  case 5: return f(pat_rk4{});
  case 6: return f(pat_ssprk3{});
  default: return hipErrorInvalidValue;
  }
}

}  // namespace rkx
