// Shared device/host helpers for the gfx950 render core.
//
// Float semantics: the library is built with -ffp-contract=off and HIP's
// default correctly rounded f32 division and square root, so every
// expression below rounds exactly like the scalar IEEE restatement of the
// reference.  Vector operators keep the reference's evaluation order
// (src/util/math.h:91-246); normalize is v * (1/sqrtf(v.v)) rather than the
// reference's device rsqrtf (src/util/math.h:210-211), whose approximation
// is not specified.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aur {

// src/util/math.h:17-22
constexpr float kInfinity = 1000000.0f;
constexpr float kEpsilon = 0.0000001f;
constexpr float kPi = 3.1415926536f;
constexpr float kInvPi = 0.3183098862f;

struct V3 {
  float x, y, z;
};

__host__ __device__ inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 mk(float v) { return V3{v, v, v}; }
__host__ __device__ inline V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
__host__ __device__ inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline V3 operator*(V3 a, float b) { return V3{a.x * b, a.y * b, a.z * b}; }
__host__ __device__ inline V3 operator*(float a, V3 b) { return V3{a * b.x, a * b.y, a * b.z}; }
__host__ __device__ inline V3 operator/(V3 a, float b) { return V3{a.x / b, a.y / b, a.z / b}; }
__host__ __device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float length(V3 v) { return sqrtf(dot(v, v)); }
__host__ __device__ inline V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
__host__ __device__ inline V3 bclerp(V3 a, V3 b, V3 c, float u, float v) {
  return (1.0f - u - v) * a + u * b + v * c;
}
__host__ __device__ inline V3 reflect(V3 i, V3 n) { return i - 2.0f * n * dot(n, i); }

// RGBA8 texture: texels at byte offset `offset` of the texel buffer,
// row-major, row y at offset + 4 * width * y.
struct TexDesc {
  unsigned long long offset;
  uint32_t width, height;
};

// Bilinear fetch with wrap addressing on normalised coordinates, texels
// read as value / 255 (the filtering the reference configures for tex2D,
// src/kernels/lib/texunit.cuh:12-19): x = frac(u) * width - 0.5, neighbours
// wrap, weights are float products.  (The texture unit quantises the weights
// to 1/256; that rounding is not reproduced.)
__host__ __device__ inline V3 tex_sample(const TexDesc& t, const uint8_t* texels, float u, float v) {
  float uf = u - floorf(u), vf = v - floorf(v);
  if (!(uf >= 0.0f && uf <= 1.0f)) uf = 0.0f;  // non-finite coordinates
  if (!(vf >= 0.0f && vf <= 1.0f)) vf = 0.0f;
  const float x = uf * (float)t.width - 0.5f, y = vf * (float)t.height - 0.5f;
  const float x0f = floorf(x), y0f = floorf(y);
  const float ax = x - x0f, ay = y - y0f;
  const int W = (int)t.width, H = (int)t.height;
  int x0 = (int)x0f % W, y0 = (int)y0f % H;
  if (x0 < 0) x0 += W;
  if (y0 < 0) y0 += H;
  const int x1 = x0 + 1 == W ? 0 : x0 + 1, y1 = y0 + 1 == H ? 0 : y0 + 1;
  const uint8_t* base = texels + t.offset;
  const uint8_t* p00 = base + 4 * ((size_t)y0 * W + x0);
  const uint8_t* p10 = base + 4 * ((size_t)y0 * W + x1);
  const uint8_t* p01 = base + 4 * ((size_t)y1 * W + x0);
  const uint8_t* p11 = base + 4 * ((size_t)y1 * W + x1);
  const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
  const float k = 1.0f / 255.0f;
  V3 c;
  c.x = (((w00 * (float)p00[0] + w10 * (float)p10[0]) + w01 * (float)p01[0]) + w11 * (float)p11[0]) * k;
  c.y = (((w00 * (float)p00[1] + w10 * (float)p10[1]) + w01 * (float)p01[1]) + w11 * (float)p11[1]) * k;
  c.z = (((w00 * (float)p00[2] + w10 * (float)p10[2]) + w01 * (float)p01[2]) + w11 * (float)p11[2]) * k;
  return c;
}
__host__ __device__ inline bool refract(V3& r, V3 i, V3 n, float eta1, float eta2) {
  const float eta = eta1 / eta2;
  const float cosi = dot(n, i);
  const float sint = eta * eta * (1.0f - cosi * cosi);
  if (sint > 1.0f) return false;
  r = eta * i - (eta + sqrtf(1.0f - sint)) * n;
  return true;
}
__host__ __device__ inline bool is_zero(V3 v) {
  return v.x < kEpsilon && v.y < kEpsilon && v.z < kEpsilon;
}
__host__ __device__ inline float comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Order-preserving map of a non-NaN float (with -0.0 already folded to +0.0)
// onto uint32: a < b as floats iff key(a) < key(b).  Never returns 0.
__host__ __device__ inline uint32_t fkey(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
#endif
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the
// multiply/xor round function with the published constants.
struct U4 {
  uint32_t x, y, z, w;
};
__host__ __device__ inline U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}
__host__ __device__ inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// PBR sample points: Owen-scrambled Sobol points (Burley, "Practical
// Hash-based Owen Scrambling", JCGT 9(4), 2020).  Each (pixel, dimension
// pair) has its own scrambles, and the sample index is shuffled by a nested
// uniform scramble so that dimension pairs are decorrelated; the first 2^m
// points of each pixel and pair still form a (0,m,2)-net.  Every point is
// uniformly distributed, so a render is an unbiased estimate, and renders
// with different seeds are independent randomisations.
//
// lowbias32 integer hash (C. Wellons, hash-prospector).
__host__ __device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__host__ __device__ inline uint32_t rev32(uint32_t x) { return __builtin_bitreverse32(x); }
// Laine-Karras style permutation with the constants of Burley's paper: each
// step flips bits from lower bits only, so rev32 around it scrambles a
// binary fraction from its most significant digit down (Owen scrambling).
__host__ __device__ inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  x = rev32(x);
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return rev32(x);
}
// Sobol dimension 1 (primitive polynomial x + 1, direction numbers m_k =
// 1, 3, 5, 15, 17, ...): the generator matrix is Pascal's triangle mod 2, so
// by Lucas' theorem output digit m (from the most significant) is the XOR of
// index bits n with n a bit superset of m: a superset-sum over GF(2), five
// masked shift-XOR steps, then a bit reversal.
__host__ __device__ inline uint32_t sobol_dim1(uint32_t i) {
  i ^= (i >> 1) & 0x55555555u;
  i ^= (i >> 2) & 0x33333333u;
  i ^= (i >> 4) & 0x0F0F0F0Fu;
  i ^= (i >> 8) & 0x00FF00FFu;
  i ^= (i >> 16) & 0x0000FFFFu;
  return rev32(i);
}
// Point `sample` of dimension pair `dims` in pixel `pixel`: x, y = 2D
// Owen-scrambled Sobol; z = a 1D Owen-scrambled van der Corput point with
// its own shuffle.  Fixed-point in [0, 2^32) (u01 takes the top 24 bits).
__host__ __device__ inline U4 sobol_owen(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dims) {
  const uint32_t key = hash32(seed ^ hash32(pixel ^ hash32(dims + 0x9E3779B9u)));
  const uint32_t i = nested_uniform_scramble(sample, hash32(key ^ 0xA511E9B3u));
  const uint32_t x = nested_uniform_scramble(rev32(i), hash32(key ^ 0x63D83595u));
  const uint32_t y = nested_uniform_scramble(sobol_dim1(i), hash32(key ^ 0x8CB92BA7u));
  const uint32_t j = nested_uniform_scramble(sample, hash32(key ^ 0x2F7A1C3Du));
  const uint32_t z = nested_uniform_scramble(rev32(j), hash32(key ^ 0x5A3E9F17u));
  return U4{x, y, z, 0u};
}

// Compat mode's RNG stream key: it draws sequentially per pixel sample.
constexpr uint32_t kKeyCompat = 0x2545F491u;


// Traversal node record: the two triangles of node k as (v1, e1 = v2-v1,
// e2 = v3-v1) -- the same float values Ray::intersect computes on the fly
// (src/kernels/lib/ray.cuh:25-26) -- followed by the slabs of the node's two
// children on the children's axis (min/max of their six axis coordinates,
// src/kernels/lib/intersect.cuh:106-111), and flags.  96 bytes, 16-byte
// aligned.  One extra record after the last node (index n) holds the root's
// own slab (axis 0) in cslab[0..1].
//
// A tree's records are kNodeBytes apart, or kNodeLineBytes (zero padding
// after each record) when the array is too large to stay in the Infinity
// Cache: unpadded, half the records straddle two 128-byte lines, and once
// node visits go to HBM every extra line is paid for; while the array is
// cache-resident the compact form keeps more records in L2.
constexpr uint32_t kNodeSlabsNormal = 1u;  // every cslab value passes normal_coord

// Coordinates x = 0 or 2^-72 <= |x| <= 2^59.  For a slab value s and a ray
// origin coordinate p that both pass, the numerator n = RN(s - p) is 0 or
// 2^-96 <= |n| <= 2^60: |s - p| <= 2^60, and a nonzero difference of two such
// floats is a multiple of 2^(-72-23).
__host__ __device__ inline bool normal_coord(float x) {
  const float a = fabsf(x);
  return (x == 0.0f) | ((a >= 0x1p-72f) & (a <= 0x1p59f));
}
struct alignas(16) NodeRec {
  float tri[2][9];
  float cslab[4];  // left lo, left hi, right lo, right hi
  uint32_t flags;
  uint32_t spare;
};
static_assert(sizeof(NodeRec) == 96, "node record");
constexpr uint32_t kNodeBytes = 96;       // records back to back
constexpr uint32_t kNodeLineBytes = 128;  // one record per 128-byte line

// Record k of a tree whose records are rec_bytes apart.
__host__ __device__ inline const NodeRec* node_at(const NodeRec* nodes, uint32_t k, uint32_t rec_bytes) {
  return reinterpret_cast<const NodeRec*>(reinterpret_cast<const char*>(nodes) + (size_t)k * rec_bytes);
}
__host__ __device__ inline NodeRec* node_at(NodeRec* nodes, uint32_t k, uint32_t rec_bytes) {
  return reinterpret_cast<NodeRec*>(reinterpret_cast<char*>(nodes) + (size_t)k * rec_bytes);
}

// Subtree arithmetic of a left-complete binary tree of n nodes in heap order.
// Number of nodes in the subtrees of the consecutive same-depth range [a, b).
__host__ __device__ inline uint64_t desc_count(uint64_t a, uint64_t b, uint64_t n) {
  uint64_t total = 0;
  uint64_t lo = a, hi = b;  // range at the current relative depth
  while (lo < n) {
    total += (hi < n ? hi : n) - lo;
    lo = 2 * lo + 1;
    hi = 2 * hi + 1;
  }
  return total;
}

__host__ __device__ inline uint32_t depth_of(uint32_t k) {
  uint32_t d = 0;
  uint32_t x = k + 1;
  while (x > 1) {
    x >>= 1;
    ++d;
  }
  return d;
}

}  // namespace aur
