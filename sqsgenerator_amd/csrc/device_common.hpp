// Device helpers shared by the scoring kernels (wave64, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sqsg {
namespace {

constexpr uint64_t kSecret0 = 0x2d358dccaa6c78a5ull;  // rapidhash.h:129-131
constexpr uint64_t kSecret1 = 0x8bb84b93962eacc9ull;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rapidrand (shuffle.h:23-26): state += s0; mix(state, state ^ s1) = lo ^ hi of the
// 128-bit product.  random_bounded (shuffle.h:28-32) only uses the low 32 bits.
__device__ __forceinline__ uint32_t draw32(uint64_t &st) {
  st += kSecret0;
  const uint64_t b = st ^ kSecret1;
  return static_cast<uint32_t>(st * b) ^ static_cast<uint32_t>(__umul64hi(st, b));
}

__device__ __forceinline__ double absolute(double a) { return a > 0 ? a : -a; }  // numeric.h:15

// |v - tgt| of an objective term as an accumulated objective needs it.  absolute() returns
// -0.0 for +0.0 where fabs returns +0.0, and nothing else differs for finite values.  A
// term's zero sign never changes the sum it joins: the sum starts at +0.0, so it is never
// -0.0 (an exact zero sum rounds to +0.0), and x + 0 = x for either zero.  fabs folds into
// the following multiply as a source modifier instead of a compare and select.
__device__ __forceinline__ double term_abs(double a) { return __builtin_fabs(a); }

// One objective term (optimization.h:85-96): v = 1 - count * pref, pw * |v - tgt|, in double,
// or in float for a configuration<float> (`single`; the parameters then hold float values).
// Returns the term and sets v (the SRO entry), both as double.
__device__ __forceinline__ double term_value(bool single, uint32_t count, double pref, double pw,
                                             double tgt, double &v) {
  if (single) {
    const float vf = 1.0f - static_cast<float>(count) * static_cast<float>(pref);
    const float d = vf - static_cast<float>(tgt);
    v = static_cast<double>(vf);
    return static_cast<double>(static_cast<float>(pw) * (d > 0 ? d : -d));
  }
  v = 1.0 - static_cast<double>(count) * pref;
  return pw * absolute(v - tgt);
}

// objective += term in the configuration's type
__device__ __forceinline__ double add_term(bool single, double obj, double t) {
  return single ? static_cast<double>(static_cast<float>(obj) + static_cast<float>(t)) : obj + t;
}

__device__ __forceinline__ unsigned long long okey(double d) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(d));
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// a * m / d for a result below 2^64 (then the high word of a * m is below d)
__device__ __forceinline__ uint64_t muldiv(uint64_t a, uint32_t m, uint32_t d) {
  const uint64_t lo = a * m, hi = __umul64hi(a, static_cast<uint64_t>(m));
  const uint64_t x1 = (hi << 32) | (lo >> 32);
  const uint64_t q1 = x1 / d, r1 = x1 % d;
  const uint64_t x0 = (r1 << 32) | (lo & 0xffffffffull);
  return (q1 << 32) | (x0 / d);
}

}  // namespace
}  // namespace sqsg
