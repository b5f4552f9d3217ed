// CPU oracle: scalar float math of the reference render core, restated.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker.
//
// Every operator below keeps the reference's evaluation order so that float
// results are reproducible bit for bit (built with -ffp-contract=off, so no
// fused multiply-add is formed):
//   vector ops, dot, cross, bclerp, reflect, refract: src/util/math.h:91-246
//   normalize uses rsqrtf on the GPU (src/util/math.h:210-211); rsqrtf's
//   approximation is CUDA-specific, so the restatement is 1.0f/sqrtf(x),
//   which is what the reference's host build uses (src/util/math.h:45-46).
//   float min/max on the device are fminf/fmaxf (NaN-ignoring), as here.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

// src/util/math.h:17-22
constexpr float kInfinity = 1000000.0f;
constexpr float kEpsilon = 0.0000001f;
constexpr float kPi = 3.1415926536f;
constexpr float kTwoPi = 6.2831853072f;
constexpr float kInvPi = 0.3183098862f;

struct f3 {
  float x, y, z;
};

inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
inline f3 mk3(float v) { return f3{v, v, v}; }
inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator*(f3 a, float b) { return f3{a.x * b, a.y * b, a.z * b}; }
inline f3 operator*(float a, f3 b) { return f3{a * b.x, a * b.y, a * b.z}; }
inline f3 operator/(f3 a, float b) { return f3{a.x / b, a.y / b, a.z / b}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(f3 v) { return sqrtf(dot(v, v)); }
inline f3 normalize(f3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
inline f3 bclerp(f3 a, f3 b, f3 c, float u, float v) {
  return (1.0f - u - v) * a + u * b + v * c;
}
inline f3 reflect(f3 i, f3 n) { return i - 2.0f * n * dot(n, i); }
// src/util/math.h:235-246 (kept literally, including its unusual form)
inline bool refract(f3& r, f3 i, f3 n, float eta1, float eta2) {
  const float eta = eta1 / eta2;
  const float cosi = dot(n, i);
  const float sint = eta * eta * (1.0f - cosi * cosi);
  if (sint > 1.0f) return false;
  r = eta * i - (eta + sqrtf(1.0f - sint)) * n;
  return true;
}
// src/util/math.h:86-87
inline bool is_zero(f3 v) { return v.x < kEpsilon && v.y < kEpsilon && v.z < kEpsilon; }

// CUDA's device fminf/fmaxf: a NaN operand is ignored.
inline float fmin_nn(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}
inline float fmax_nn(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}
inline float comp(f3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// src/util/math.h:253-272
inline uint32_t log2i(uint32_t x) {
  uint32_t ret = ~0u;
  while (x) {
    x >>= 1;
    ret++;
  }
  return ret;
}
inline uint32_t exp2i(uint32_t x) {
  if (x == 0) return 1;
  uint32_t ret = 2;
  while (--x) ret *= 2;
  return ret;
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  It replaces the
// reference's per-pixel cuRAND state (src/kernels/utility.cu:182-197).
struct u4 {
  uint32_t x, y, z, w;
};
inline u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}
// 24-bit uniform in [0, 1): exact in float.
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

}  // namespace orc
