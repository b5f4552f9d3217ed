// CPU oracle: camera rays, tangent frames, direct lighting and the path loop.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.h).
//
// Reference-compat mode restates, in order of execution:
//   Camera                     src/util/camera.h:18-27 (built by the caller)
//   cudaGenerateRays(+Multisample) src/kernels/utility.cu:18-94
//   cudaRaytraceKernel / raytrace  src/kernels/monteCarloRaytracer.cu:20-113
//   Shader::getBSDF            src/kernels/lib/shader.cuh:8-27
//   getBasisVectors            src/kernels/lib/common.cuh:53-66
//   BSDF::f                    src/kernels/lib/bsdf.cuh:13-41 (Phong evaluates as Lambert)
//   estimate*Radiance          src/kernels/lib/radiance.cuh:8-80
//   Light::sampleL / visiblePoint src/kernels/lib/light.cuh:24-60
//   sampleRectangle            src/kernels/lib/sampling.cuh:148-152
//   cudaDrawPixels             src/kernels/utility.cu:199-229
//   MonteCarloRaytracer::render src/render/monteCarloRaytracer.cpp:81-120
// cuRAND (utility.cu:182-197, light.cuh:47) is replaced by Philox4x32-10
// keyed by (seed, pixel, sample, draw index): reference images are not
// reproducible (tick-count seeding), so compat images match the reference
// only statistically while matching this restatement per pixel.
//
// PBR mode is new semantics that sits beside compat mode: a path loop of
// max_depth bounces with next-event estimation on every light, cosine
// sampling of the Lambert BSDF (concentric map, Shirley & Chiu 1997, as in
// sampling.cuh:51-82), Russian roulette from bounce rr_start, and
// Owen-scrambled Sobol sample points (Burley, JCGT 2020; sobol_owen below)
// instead of independent draws.  Point, directional and ambient lights keep
// compat mode's formulas (direct_delta), an ambient light with both its
// unshadowed and its positional part.  The same float and integer operations
// run on the GPU.
//
// cudaGenerateTB (utility.cu:136-180) is restated as orc_generate_tb.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_math.h"

namespace {
using namespace orc;

// Layouts identical to the C-ABI structs (80 B each, src/util/shader.h:23-37,
// src/util/light.h:24-35).
struct Shader {
  int32_t type;
  float diffuseColor[3], emissionColor[3], specularColor[3], reflectionColor[3];
  float diffuse, emission, exponent, reflectivity, translucence, refractiveIndex;
  uint32_t texture;
};
struct Light {
  int32_t type;  // 0 ambient, 1 point, 2 directional, 3 area
  float intensity, ambient, area;
  int16_t samples, pad;
  float position[3], color[3], direction[3], e1[3], e2[3];
};
struct Camera {
  float position[3], forward[3], right[3], up[3], tanfov[2], aspect;
};
struct Opts {
  uint32_t mode, spp, max_depth, rr_start, seed;
};
static_assert(sizeof(Shader) == 80, "shader layout");
static_assert(sizeof(Light) == 80, "light layout");

inline f3 v3(const float* p) { return mk3(p[0], p[1], p[2]); }

struct TexDesc {  // layout of the C ABI's aur_texture: RGBA8 texels at byte offset, row-major
  uint32_t width, height;
  uint64_t offset;
};
static_assert(sizeof(TexDesc) == 16, "texture descriptor layout");

struct Scene {
  uint32_t count;
  const float *verts, *normals, *tangents, *bitangents;
  const uint16_t* shaderIds;
  uint32_t nShaders;
  const Shader* shaders;
  uint32_t nLights;
  const Light* lights;
  const float* texcoords = nullptr;  // tree order, 6 per triangle
  uint32_t nTextures = 0;
  const TexDesc* textures = nullptr;
  const uint8_t* texels = nullptr;
};

// texfetch (src/kernels/lib/texunit.cuh:54-94) with the filtering the
// reference configures (linear, wrap, normalised coordinates, :12-19) as an
// own bilinear: x = frac(u) * width - 0.5, wrapped neighbours, float weights
// (the texture unit's 1/256 weight quantisation is not reproduced).
f3 tex_sample(const TexDesc& t, const uint8_t* texels, float u, float v) {
  float uf = u - floorf(u), vf = v - floorf(v);
  if (!(uf >= 0.0f && uf <= 1.0f)) uf = 0.0f;
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
  const uint8_t* p00 = base + 4 * (size_t(y0) * W + x0);
  const uint8_t* p10 = base + 4 * (size_t(y0) * W + x1);
  const uint8_t* p01 = base + 4 * (size_t(y1) * W + x0);
  const uint8_t* p11 = base + 4 * (size_t(y1) * W + x1);
  const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
  const float k = 1.0f / 255.0f;
  f3 c;
  c.x = (((w00 * (float)p00[0] + w10 * (float)p10[0]) + w01 * (float)p01[0]) + w11 * (float)p11[0]) * k;
  c.y = (((w00 * (float)p00[1] + w10 * (float)p10[1]) + w01 * (float)p01[1]) + w11 * (float)p11[1]) * k;
  c.z = (((w00 * (float)p00[2] + w10 * (float)p10[2]) + w01 * (float)p01[2]) + w11 * (float)p11[2]) * k;
  return c;
}

// Texture colour of a hit (getTexCoord, src/kernels/lib/common.cuh:45-51,
// bclerp's order); white when the texture index exceeds the uploaded count.
f3 tex_color(const Scene& sc, uint32_t texture, uint32_t tri, float u, float v) {
  const uint32_t id = texture - 1;
  if (id >= sc.nTextures || !sc.texcoords) return mk3(1.0f);
  const float* t = sc.texcoords + size_t(tri) * 6;
  const float w = 1.0f - u - v;
  const float tu = (w * t[0] + u * t[2]) + v * t[4];
  const float tv = (w * t[1] + u * t[3]) + v * t[5];
  return tex_sample(sc.textures[id], sc.texels, tu, tv);
}

// ---- traversal (same restatement as trace.cpp, kept local so that the
// path loop below reads like the reference) ----
struct Ray {
  f3 pos, dir;
  float t, weight, eta;
};

inline float vmin3(const float* v, int a) {
  return fmin_nn(fmin_nn(v[3 * a], v[3 * a + 1]), v[3 * a + 2]);
}
inline float vmax3(const float* v, int a) {
  return fmax_nn(fmax_nn(v[3 * a], v[3 * a + 1]), v[3 * a + 2]);
}
inline bool ray_tri(const Ray& r, const float* v, float& u, float& vv, float& t) {
  const f3 v1 = mk3(v[0], v[3], v[6]), v2 = mk3(v[1], v[4], v[7]), v3_ = mk3(v[2], v[5], v[8]);
  const f3 e1 = v2 - v1, e2 = v3_ - v1;
  const f3 P = cross(r.dir, e2);
  const float det = dot(e1, P);
  if (det < kEpsilon) return false;
  const float invdet = 1.0f / det;
  const f3 T = r.pos - v1;
  u = dot(T, P) * invdet;
  if (u < 0.0f || u > 1.0f) return false;
  const f3 Q = cross(T, e1);
  vv = dot(r.dir, Q) * invdet;
  if (vv < 0.0f || u + vv > 1.0f) return false;
  t = dot(e2, Q) * invdet;
  return true;
}
template <bool kAny>
bool traverse(const Scene& sc, Ray& ray, int32_t& id, float& hu, float& hv) {
  struct S {
    uint32_t index;
    int axis;
    float rx, ry;
  } st[32];
  uint32_t n = 0;
  bool hit = false;
  st[n++] = S{0, 0, kEpsilon, kInfinity};
  while (n > 0) {
    S s = st[--n];
    if (s.index >= sc.count) continue;
    const float* a = sc.verts + size_t(s.index) * 9;
    const float* b = a + 9;
    const float slo = fmin_nn(vmin3(a, s.axis), vmin3(b, s.axis));
    const float shi = fmax_nn(vmax3(a, s.axis), vmax3(b, s.axis));
    const float p = comp(ray.pos, s.axis), d = comp(ray.dir, s.axis);
    const float d0 = (slo - p) / d, d1 = (shi - p) / d;
    s.rx = fmax_nn(fmin_nn(d0, d1), s.rx);
    s.ry = fmin_nn(fmax_nn(d0, d1), s.ry);
    if (!((s.rx <= s.ry) && (s.rx <= ray.t))) continue;
    float u, v, t;
    if (ray_tri(ray, a, u, v, t) && t > kEpsilon && t < ray.t) {
      if (kAny) return true;
      hit = true;
      ray.t = t;
      hu = u;
      hv = v;
      id = (int32_t)s.index;
    }
    if (ray_tri(ray, b, u, v, t) && t > kEpsilon && t < ray.t) {
      if (kAny) return true;
      hit = true;
      ray.t = t;
      hu = u;
      hv = v;
      id = (int32_t)s.index + 1;
    }
    const int ca = (s.axis + 1) % 3;
    uint32_t L = 2 * s.index + 2, R = 2 * s.index + 4;
    if (comp(ray.dir, ca) < 0.0f) {
      const uint32_t tmp = L;
      L = R;
      R = tmp;
    }
    if (n + 2 > 32) return hit;
    st[n++] = S{R, ca, s.rx, s.ry};
    st[n++] = S{L, ca, s.rx, s.ry};
  }
  return hit;
}
inline bool occluded(const Scene& sc, const Ray& r) {
  Ray tmp = r;
  int32_t id = -1;
  float u, v;
  return traverse<true>(sc, tmp, id, u, v);
}

inline f3 interp(const float* base, uint32_t tri, float u, float v) {
  const float* p = base + size_t(tri) * 9;  // "values" layout v1xyz v2xyz v3xyz
  return bclerp(mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5]), mk3(p[6], p[7], p[8]), u, v);
}

inline const Shader& shader_of(const Scene& sc, uint32_t tri) {
  uint32_t sid = sc.shaderIds[tri];
  sid = sid > 0 ? sid - 1 : 0;  // getSafeID, src/util/array.h:18
  if (sid >= sc.nShaders) sid = 0;
  return sc.shaders[sid];
}

// ---- random numbers ----
constexpr uint32_t kKeyCompat = 0x2545F491u;

struct CompatRng {
  uint32_t seed, pixel, sample, k = 0;
  u4 cache{};
  float next() {
    if ((k & 3u) == 0)
      cache = philox4x32_10(u4{pixel, sample, 0x40000000u | (k >> 2), 0u}, seed, kKeyCompat);
    const uint32_t w = (k & 3u) == 0 ? cache.x : (k & 3u) == 1 ? cache.y : (k & 3u) == 2 ? cache.z : cache.w;
    ++k;
    return u01(w);
  }
};

// PBR sample points: Owen-scrambled Sobol points, Burley, "Practical
// Hash-based Owen Scrambling", JCGT 9(4), 2020 (nested uniform scrambling by
// a Laine-Karras permutation between bit reversals, with the paper's
// constants; the sample index shuffled the same way).  Sobol dimension 0 is
// the bit-reversed index; dimension 1 (polynomial x + 1, Joe-Kuo direction
// numbers 1, 3, 5, 15, ...) is written here as the direction-number XOR sum,
// one index bit at a time.  Keys: lowbias32 (C. Wellons, hash-prospector).
inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
inline uint32_t rev32(uint32_t x) {
  uint32_t r = 0;
  for (int k = 0; k < 32; ++k) r |= ((x >> k) & 1u) << (31 - k);
  return r;
}
inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  x = rev32(x);
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return rev32(x);
}
inline uint32_t sobol_dim1(uint32_t i) {
  uint32_t v = 1u << 31, r = 0;  // v_k = m_k / 2^k as a 32-bit fraction
  for (int k = 0; k < 32; ++k, v ^= v >> 1)
    if ((i >> k) & 1u) r ^= v;
  return r;
}
inline u4 sobol_owen(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dims) {
  const uint32_t key = hash32(seed ^ hash32(pixel ^ hash32(dims + 0x9E3779B9u)));
  const uint32_t i = nested_uniform_scramble(sample, hash32(key ^ 0xA511E9B3u));
  const uint32_t x = nested_uniform_scramble(rev32(i), hash32(key ^ 0x63D83595u));
  const uint32_t y = nested_uniform_scramble(sobol_dim1(i), hash32(key ^ 0x8CB92BA7u));
  const uint32_t j = nested_uniform_scramble(sample, hash32(key ^ 0x2F7A1C3Du));
  const uint32_t z = nested_uniform_scramble(rev32(j), hash32(key ^ 0x5A3E9F17u));
  return u4{x, y, z, 0u};
}
// dimension pair (bounce, slot): slot 0 of bounce 0 = pixel position, slot 1
// = BSDF direction (x, y) and Russian roulette (z), slots 16.. = light samples
inline u4 pbr_draw(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t slot) {
  return sobol_owen(seed, pixel, sample, (bounce << 16) | slot);
}

// ---- camera rays ----
f3 camera_dir_compat(const Camera& c, uint32_t w, uint32_t h, uint32_t x, uint32_t y,
                     uint32_t sampleID, uint32_t samples) {
  float px, py;
  if (samples == 1) {
    px = float(x) / float(w) - 0.5f;
    py = float(y) / float(h) - 0.5f;
  } else {
    const short swidth = (short)sqrtf((float)(unsigned short)samples);
    const short sx = (short)(sampleID % swidth), sy = (short)(sampleID / swidth);
    const float dx = 1.0f / float(w), dy = 1.0f / float(h);
    const float sdx = dx / swidth, sdy = dy / swidth;
    px = x * dx - 0.5f + (sx * sdx);
    py = y * dy - 0.5f + (sy * sdy);
  }
  return normalize(px * v3(c.right) * c.tanfov[0] + py * v3(c.up) * c.tanfov[1] + v3(c.forward));
}

// ---- compat shading ----
f3 direct_delta(const Scene& sc, const Light& L, f3 color1, f3 N, f3 pos) {
  Ray wi{pos, mk3(0.0f), kInfinity, 1.0f, 1.0f};
  f3 Li;
  const f3 CI = v3(L.color) * L.intensity;
  if (L.type == 0) {
    wi.dir = v3(L.position) - wi.pos;
    wi.t = length(wi.dir);
    wi.dir = wi.dir / wi.t;
    Li = CI * (1.0f - L.ambient);
  } else if (L.type == 1) {
    wi.dir = v3(L.position) - wi.pos;
    wi.t = length(wi.dir);
    wi.dir = wi.dir / wi.t;
    Li = CI / (wi.t * wi.t);
  } else {
    wi.dir = -v3(L.direction);
    wi.t = kInfinity;
    Li = CI;
  }
  const float pdf = 1.0f;
  if (pdf > 0.0f && !is_zero(Li)) {
    const f3 f = color1 * kInvPi;
    if (!is_zero(f) && !occluded(sc, wi)) return f * Li * fmaxf(0.0f, dot(wi.dir, N));
  }
  return mk3(0.0f);
}

f3 direct_area(const Scene& sc, const Light& L, f3 color1, f3 N, f3 pos, CompatRng& rng) {
  f3 acc = mk3(0.0f);
  const f3 CI = v3(L.color) * L.intensity;
  for (int i = 0; i < L.samples; ++i) {
    f3 Ls = mk3(0.0f);
    Ray wi{pos, mk3(0.0f), kInfinity, 1.0f, 1.0f};
    const float u1 = rng.next();
    const float u2 = rng.next();
    const f3 P = v3(L.position) + ((v3(L.e1) * u1) + (v3(L.e2) * u2));
    wi.dir = P - wi.pos;
    wi.t = length(wi.dir);
    wi.dir = wi.dir / wi.t;
    const float lightPdf = 1.0f / fabsf(dot(v3(L.direction), -wi.dir));
    const f3 Li = CI * fmaxf(0.0f, dot(v3(L.direction), -wi.dir));
    // Ray::offset, src/kernels/lib/ray.cuh:8-15
    wi.pos.x += kEpsilon * wi.dir.x;
    wi.pos.y += kEpsilon * wi.dir.y;
    wi.pos.z += kEpsilon * wi.dir.z;
    wi.t += kEpsilon;
    if (lightPdf > 0.0f && !is_zero(Li)) {
      const f3 f = color1 * kInvPi;
      if (!is_zero(f) && !occluded(sc, wi)) Ls = Ls + f * Li * fmaxf(0.0f, dot(wi.dir, N));
    }
    acc = acc + Ls;
  }
  const float ns = (float)L.samples;
  return mk3(acc.x / ns, acc.y / ns, acc.z / ns);
}

struct Sample {
  f3 color;
  bool covered;
};

Sample shade_compat(const Scene& sc, f3 origin, f3 dir, CompatRng& rng) {
  const int kMaxDepth = 8;
  Ray stack[kMaxDepth + 2];
  int n = 0;
  stack[n++] = Ray{origin, dir, kInfinity, 1.0f, 1.0f};
  int depth = 0;
  f3 color = mk3(0.0f);
  int32_t hpId = -1;
  float hpU = 0.0f, hpV = 0.0f;
  while (depth <= kMaxDepth && n > 0) {
    Ray ray = stack[--n];
    if (!traverse<false>(sc, ray, hpId, hpU, hpV)) continue;
    const f3 position = ray.pos + (ray.dir * ray.t);
    const f3 wo = -ray.dir;
    const Shader& sh = shader_of(sc, (uint32_t)hpId);
    const f3 N = normalize(interp(sc.normals, hpId, hpU, hpV));
    f3 color1 = v3(sh.diffuseColor) * sh.diffuse;
    if (sh.texture > 0) color1 = tex_color(sc, sh.texture, (uint32_t)hpId, hpU, hpV) * sh.diffuse;
    f3 Li = v3(sh.emissionColor);
    for (uint32_t i = 0; i < sc.nLights; ++i) {
      const Light& L = sc.lights[i];
      if (L.type == 0) Li = Li + (color1 * kInvPi) * (v3(L.color) * L.intensity * L.ambient);
      if (L.type != 3)
        Li = Li + direct_delta(sc, L, color1, N, position);
      else
        Li = Li + direct_area(sc, L, color1, N, position, rng);
    }
    if (sh.reflectivity > 0.0f || sh.translucence > 0.0f) depth++;
    if (depth <= kMaxDepth && sh.reflectivity > 0.0f) {
      Ray nr{position, reflect(ray.dir, N), kInfinity, sh.reflectivity, ray.eta};
      ray.weight *= (1.0f - sh.reflectivity);
      stack[n++] = nr;
    }
    if (depth <= kMaxDepth && sh.translucence > 0.0f) {
      f3 rN = N;
      float etai = ray.eta, etat = sh.refractiveIndex;
      if (dot(rN, wo) < 0.0f) {
        rN = -rN;
        etai = sh.refractiveIndex;
        etat = 1.0f;
      }
      f3 tdir;
      if (refract(tdir, ray.dir, rN, etai, etat)) {
        Ray nr{position, tdir, kInfinity, sh.translucence, etat};
        ray.weight *= (1.0f - sh.translucence);
        stack[n++] = nr;
      }
    }
    color = color + (Li * ray.weight);
  }
  return Sample{color, hpId != -1};
}

// ---- PBR path loop ----
// sin/cos of phi = (pi/4) q for |q| <= 1: Taylor polynomials in Horner form
// (|error| < 2e-9 on the range); only + and * so that the CPU and the GPU
// round identically.
inline void sincos_quarter(float q, float& s, float& c) {
  const float phi = 0.785398163f * q;
  const float z = phi * phi;
  s = phi * (1.0f + z * (-0.166666667f + z * (0.00833333333f + z * (-0.000198412698f + z * 2.75573192e-6f))));
  c = 1.0f + z * (-0.5f + z * (0.0416666667f + z * (-0.00138888889f + z * (2.48015873e-5f + z * -2.75573192e-7f))));
}

// Concentric square-to-disk map, branch structure of sampling.cuh:51-73 with
// the angle written as (pi/4)(k + q) so that only the quarter-range
// polynomial is needed; multiples of pi/2 are applied as exact rotations.
inline void concentric_disk(float u1, float u2, float& dx, float& dy) {
  const float x = 2.0f * u1 - 1.0f;
  const float y = 2.0f * u2 - 1.0f;
  float r, s, c;
  if (x >= -y) {
    if (x > y) {
      r = x;
      sincos_quarter(y / x, s, c);
      dx = r * c;
      dy = r * s;
    } else if (y != 0.0f) {
      r = y;
      sincos_quarter(-x / y, s, c);  // angle pi/2 + (pi/4) q
      dx = r * -s;
      dy = r * c;
    } else {
      // x = y = 0: the reference divides 0/0 here and returns NaN
      dx = 0.0f;
      dy = 0.0f;
    }
  } else {
    if (x < y) {
      r = -x;
      sincos_quarter(y / x, s, c);  // angle pi + (pi/4) q
      dx = r * -c;
      dy = r * -s;
    } else {
      r = -y;
      if (y != 0.0f) {
        sincos_quarter(-x / y, s, c);  // angle 3pi/2 + (pi/4) q
        dx = r * s;
        dy = r * -c;
      } else {
        dx = r;
        dy = 0.0f;
      }
    }
  }
}

// Orthonormal basis around n (Duff, Burgess, Christensen, Hery, Kensler,
// Liani, Villemin, "Building an Orthonormal Basis, Revisited", JCGT 2017).
inline void onb(f3 n, f3& t1, f3& t2) {
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  t1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  t2 = mk3(b, sign + n.y * n.y * a, -n.y);
}

// Cosine-weighted direction around N: the concentric disk point lifted to
// the hemisphere (Malley's method) in the onb() frame.
inline f3 sample_cosine(f3 N, float u1, float u2) {
  float dx, dy;
  concentric_disk(u1, u2, dx, dy);
  const float dz = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
  f3 t1, t2;
  onb(N, t1, t2);
  return t1 * dx + t2 * dy + N * dz;
}

// One path vertex as the adjoint needs it: shader index, Russian-roulette
// probability applied after it (1 when none), and per light the unoccluded
// next-event weight W with contribution beta * albedo * radiance * W / pi.
struct Vertex {
  uint32_t shader;
  float q;
  std::vector<float> W;
};

inline uint32_t shader_index(const Scene& sc, uint32_t tri) {
  uint32_t sid = sc.shaderIds[tri];
  sid = sid > 0 ? sid - 1 : 0;
  return sid < sc.nShaders ? sid : 0;
}

Sample shade_pbr(const Scene& sc, const Opts& o, f3 origin, f3 dir, uint32_t pixel, uint32_t sample,
                 std::vector<Vertex>* rec = nullptr) {
  f3 beta = mk3(1.0f), L = mk3(0.0f);
  bool covered = false;
  Ray ray{origin, dir, kInfinity, 1.0f, 1.0f};
  for (uint32_t b = 0; b < o.max_depth; ++b) {
    int32_t id = -1;
    float hu = 0.0f, hv = 0.0f;
    ray.t = kInfinity;
    if (!traverse<false>(sc, ray, id, hu, hv)) break;
    if (b == 0) covered = true;
    const f3 P = ray.pos + (ray.dir * ray.t);
    const f3 wo = -ray.dir;
    const Shader& sh = shader_of(sc, (uint32_t)id);
    f3 N = normalize(interp(sc.normals, id, hu, hv));
    if (dot(N, wo) < 0.0f) N = -N;
    if (rec) rec->push_back(Vertex{shader_index(sc, (uint32_t)id), 1.0f, std::vector<float>(sc.nLights, 0.0f)});
    f3 albedo = v3(sh.diffuseColor) * sh.diffuse;
    if (sh.texture > 0) albedo = tex_color(sc, sh.texture, (uint32_t)id, hu, hv) * sh.diffuse;
    const f3 f = albedo * kInvPi;
    L = L + beta * v3(sh.emissionColor);
    const f3 bf = beta * f;
    uint32_t slot = 16;
    for (uint32_t li = 0; li < sc.nLights; ++li) {
      const Light& Lt = sc.lights[li];
      const f3 CI = v3(Lt.color) * Lt.intensity;
      if (Lt.type == 3) {
        const int ns = Lt.samples > 0 ? Lt.samples : 1;
        const float inv = 1.0f / (float)ns;
        for (int j = 0; j < ns; ++j) {
          const u4 r = pbr_draw(o.seed, pixel, sample, b, slot++);
          const f3 Q = v3(Lt.position) + ((v3(Lt.e1) * u01(r.x)) + (v3(Lt.e2) * u01(r.y)));
          f3 wi = Q - P;
          const float dist = length(wi);
          wi = wi / dist;
          const float cosL = dot(v3(Lt.direction), -wi);
          const float cosS = dot(N, wi);
          if (cosL > 0.0f && cosS > 0.0f) {
            Ray sr{P + (kEpsilon * wi), wi, dist + kEpsilon, 1.0f, 1.0f};
            if (!occluded(sc, sr)) {
              const float w = (cosS * cosL * Lt.area) / (dist * dist) * inv;
              L = L + bf * (CI * w);
              if (rec) rec->back().W[li] += w;
            }
          }
        }
      } else if (Lt.type == 1) {
        f3 wi = v3(Lt.position) - P;
        const float dist = length(wi);
        wi = wi / dist;
        const float cosS = dot(N, wi);
        if (cosS > 0.0f) {
          Ray sr{P, wi, dist, 1.0f, 1.0f};
          if (!occluded(sc, sr)) {
            L = L + bf * ((CI / (dist * dist)) * cosS);
            if (rec) rec->back().W[li] += cosS / (dist * dist);
          }
        }
      } else if (Lt.type == 2) {
        const f3 wi = -v3(Lt.direction);
        const float cosS = dot(N, wi);
        if (cosS > 0.0f) {
          Ray sr{P, wi, kInfinity, 1.0f, 1.0f};
          if (!occluded(sc, sr)) {
            L = L + bf * (CI * cosS);
            if (rec) rec->back().W[li] += cosS;
          }
        }
      } else {
        // ambient part unshadowed, the rest from the position without falloff
        // (direct_delta's ambient case)
        L = L + bf * (CI * Lt.ambient);
        if (rec) rec->back().W[li] += Lt.ambient;
        f3 wi = v3(Lt.position) - P;
        const float dist = length(wi);
        wi = wi / dist;
        const float cosS = dot(N, wi);
        const float rest = 1.0f - Lt.ambient;
        if (cosS > 0.0f && rest != 0.0f) {
          Ray sr{P, wi, dist, 1.0f, 1.0f};
          if (!occluded(sc, sr)) {
            L = L + bf * ((CI * rest) * cosS);
            if (rec) rec->back().W[li] += rest * cosS;
          }
        }
      }
    }
    if (b + 1 >= o.max_depth) break;
    const u4 r = pbr_draw(o.seed, pixel, sample, b, 1);
    if (b >= o.rr_start) {
      const float q = fminf(0.95f, fmaxf(beta.x, fmaxf(beta.y, beta.z)));
      if (u01(r.z) >= q) break;
      beta = beta / q;
      if (rec) rec->back().q = q;
    }
    const f3 wi = sample_cosine(N, u01(r.x), u01(r.y));
    beta = beta * albedo;
    ray = Ray{P, wi, kInfinity, 1.0f, 1.0f};
  }
  return Sample{L, covered};
}

f3 camera_dir_pbr(const Camera& c, uint32_t w, uint32_t h, uint32_t x, uint32_t y, uint32_t seed,
                  uint32_t pixel, uint32_t s) {
  const u4 r = pbr_draw(seed, pixel, s, 0, 0);
  const float fx = u01(r.x), fy = u01(r.y);
  const float px = ((float)x + fx) / (float)w - 0.5f;
  const float py = ((float)y + fy) / (float)h - 0.5f;
  return normalize(px * v3(c.right) * c.tanfov[0] + py * v3(c.up) * c.tanfov[1] + v3(c.forward));
}

}  // namespace

extern "C" {

// cudaGenerateTB restated (utility.cu:136-180): per vertex tangent n x up,
// falling back to n x forward when shorter than Epsilon; bitangent n x T.
void orc_generate_tb(uint32_t count, const float* normals, float* tangents, float* bitangents) {
  const f3 up = mk3(0.0f, 1.0f, 0.0f), fwd = mk3(0.0f, 0.0f, 1.0f);
  for (uint32_t t = 0; t < count; ++t) {
    for (int k = 0; k < 3; ++k) {
      const float* n = normals + size_t(t) * 9 + 3 * k;
      const f3 nv = mk3(n[0], n[1], n[2]);
      f3 T = cross(nv, up);
      if (length(T) < kEpsilon) T = cross(nv, fwd);
      const f3 B = cross(nv, T);
      float* tp = tangents + size_t(t) * 9 + 3 * k;
      float* bp = bitangents + size_t(t) * 9 + 3 * k;
      tp[0] = T.x;
      tp[1] = T.y;
      tp[2] = T.z;
      bp[0] = B.x;
      bp[1] = B.y;
      bp[2] = B.z;
    }
  }
}

// Camera ray direction for one pixel sample (compat: utility.cu:18-77).
void orc_camera_ray(const void* cam, uint32_t w, uint32_t h, uint32_t x, uint32_t y,
                    uint32_t sampleID, uint32_t samples, float* dir_out) {
  const f3 d = camera_dir_compat(*(const Camera*)cam, w, h, x, y, sampleID, samples);
  dir_out[0] = d.x;
  dir_out[1] = d.y;
  dir_out[2] = d.z;
}

// Render a list of pixels (or all when pixels == NULL, npix ignored) and
// write float4 RGBA accumulations per listed pixel into rgba (npix*4).
// Geometry arrays are in tree order.  opts: {mode, spp, max_depth, rr_start,
// seed}; mode 0 = reference-compat, 1 = PBR.
// The same with texture maps: texcoords in tree order (6 per triangle) and
// ntex textures (descriptor {uint32 width, uint32 height, uint64 offset},
// RGBA8 texels).
int orc_render_tex(uint32_t count, const float* verts, const float* normals, const float* tangents,
                   const float* bitangents, const uint16_t* shaderIds, uint32_t nShaders,
                   const void* shaders, uint32_t nLights, const void* lights, const void* camera,
                   uint32_t w, uint32_t h, const uint32_t* opts, const uint32_t* pixels, uint32_t npix,
                   float* rgba, int threads, const float* texcoords, uint32_t ntex, const void* texdesc,
                   const uint8_t* texels);

int orc_render(uint32_t count, const float* verts, const float* normals, const float* tangents,
               const float* bitangents, const uint16_t* shaderIds, uint32_t nShaders,
               const void* shaders, uint32_t nLights, const void* lights, const void* camera,
               uint32_t w, uint32_t h, const uint32_t* opts, const uint32_t* pixels, uint32_t npix,
               float* rgba, int threads) {
  return orc_render_tex(count, verts, normals, tangents, bitangents, shaderIds, nShaders, shaders, nLights,
                        lights, camera, w, h, opts, pixels, npix, rgba, threads, nullptr, 0, nullptr, nullptr);
}

int orc_render_tex(uint32_t count, const float* verts, const float* normals, const float* tangents,
                   const float* bitangents, const uint16_t* shaderIds, uint32_t nShaders,
                   const void* shaders, uint32_t nLights, const void* lights, const void* camera,
                   uint32_t w, uint32_t h, const uint32_t* opts, const uint32_t* pixels, uint32_t npix,
                   float* rgba, int threads, const float* texcoords, uint32_t ntex, const void* texdesc,
                   const uint8_t* texels) {
  if (count < 2 || (count & 1u) || nShaders == 0 || w == 0 || h == 0) return -1;
  Scene sc{count, verts, normals, tangents, bitangents, shaderIds, nShaders,
           (const Shader*)shaders, nLights, (const Light*)lights};
  sc.texcoords = texcoords;
  sc.nTextures = ntex;
  sc.textures = (const TexDesc*)texdesc;
  sc.texels = texels;
  const Opts o{opts[0], opts[1], opts[2], opts[3], opts[4]};
  if (o.spp == 0) return -1;
  const Camera& cam = *(const Camera*)camera;
  const uint32_t total = pixels ? npix : w * h;
  const float weight = 1.0f / (float)o.spp;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
  for (int64_t q = 0; q < (int64_t)total; ++q) {
    const uint32_t pix = pixels ? pixels[q] : (uint32_t)q;
    const uint32_t x = pix % w, y = pix / w;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t s = 0; s < o.spp; ++s) {
      Sample smp;
      if (o.mode == 0) {
        const f3 d = camera_dir_compat(cam, w, h, x, y, s, o.spp);
        CompatRng rng{o.seed, pix, s};
        smp = shade_compat(sc, v3(cam.position), d, rng);
      } else {
        const f3 d = camera_dir_pbr(cam, w, h, x, y, o.seed, pix, s);
        smp = shade_pbr(sc, o, v3(cam.position), d, pix, s);
      }
      if (smp.covered) {
        acc[0] += smp.color.x * weight;
        acc[1] += smp.color.y * weight;
        acc[2] += smp.color.z * weight;
        acc[3] += weight;
      }
    }
    std::memcpy(rgba + size_t(q) * 4, acc, sizeof(acc));
  }
  return 0;
}

// Gradient of sum_p <dimg_p, I_p> (dimg = dLoss/dI, float4 per pixel, RGB
// used) for the PBR image of orc_render, with respect to each shader's
// albedo (diffuseColor * diffuse) and each light's radiance (color *
// intensity), in double.  Product-rule form: every estimator term is a
// product of the albedos of the vertices before it (and of its own vertex
// for a next-event term) with constant factors, so its derivative with
// respect to albedo m is term * (times m occurs) / albedo_m; Russian-roulette
// probabilities are constants.  Requires albedo channels > 0.
int orc_adjoint(uint32_t count, const float* verts, const float* normals, const float* tangents,
                const float* bitangents, const uint16_t* shaderIds, uint32_t nShaders,
                const void* shaders, uint32_t nLights, const void* lights, const void* camera,
                uint32_t w, uint32_t h, const uint32_t* opts, const float* dimg, double* g_albedo,
                double* g_light, int threads) {
  if (count < 2 || (count & 1u) || nShaders == 0 || w == 0 || h == 0) return -1;
  Scene sc{count, verts, normals, tangents, bitangents, shaderIds, nShaders,
           (const Shader*)shaders, nLights, (const Light*)lights};
  const Opts o{opts[0], opts[1], opts[2], opts[3], opts[4]};
  if (o.spp == 0 || o.mode != 1) return -1;
  const Camera& cam = *(const Camera*)camera;
  const int nt = threads > 0 ? threads : 1;
  std::vector<std::vector<double>> ga(nt, std::vector<double>(size_t(nShaders) * 3, 0.0));
  std::vector<std::vector<double>> gl(nt, std::vector<double>(size_t(nLights) * 3, 0.0));
  auto rho_of = [&](uint32_t m, int c) {
    const Shader& sh = sc.shaders[m];
    return (double)sh.diffuseColor[c] * (double)sh.diffuse;
  };
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (int64_t pix = 0; pix < (int64_t)w * h; ++pix) {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    const uint32_t x = (uint32_t)pix % w, y = (uint32_t)pix / w;
    std::vector<Vertex> path;
    for (uint32_t s = 0; s < o.spp; ++s) {
      path.clear();
      const f3 d = camera_dir_pbr(cam, w, h, x, y, o.seed, (uint32_t)pix, s);
      shade_pbr(sc, o, v3(cam.position), d, (uint32_t)pix, s, &path);
      for (int c = 0; c < 3; ++c) {
        const double dl = (double)dimg[pix * 4 + c] / (double)o.spp;
        if (dl == 0.0) continue;
        double beta = 1.0;
        for (size_t b = 0; b < path.size(); ++b) {
          const Vertex& v = path[b];
          const double rho = rho_of(v.shader, c);
          const Shader& sh = sc.shaders[v.shader];
          // emission term: albedos of vertices 0..b-1
          const double te = beta * (double)sh.emissionColor[c];
          // next-event terms: albedos of vertices 0..b
          double tn = 0.0;
          for (uint32_t l = 0; l < nLights; ++l) {
            if (v.W[l] == 0.0f) continue;
            const Light& Lt = sc.lights[l];
            const double ci = (double)Lt.color[c] * (double)Lt.intensity;
            const double k = beta * rho * (double)v.W[l] / M_PI;
            tn += k * ci;
            gl[t][size_t(l) * 3 + c] += dl * k;
          }
          for (size_t i = 0; i <= b; ++i) {
            const uint32_t m = path[i].shader;
            const double r = rho_of(m, c);
            double term = tn;
            if (i < b) term += te;
            ga[t][size_t(m) * 3 + c] += dl * term / r;
          }
          beta = beta * rho / (double)v.q;
        }
      }
    }
  }
  for (size_t i = 0; i < size_t(nShaders) * 3; ++i) {
    double acc = 0.0;
    for (int t = 0; t < nt; ++t) acc += ga[t][i];
    g_albedo[i] = acc;
  }
  for (size_t i = 0; i < size_t(nLights) * 3; ++i) {
    double acc = 0.0;
    for (int t = 0; t < nt; ++t) acc += gl[t][i];
    g_light[i] = acc;
  }
  return 0;
}

// cudaFilterPixelsKernel restated (src/kernels/utility.cu:231-258): 3x3
// tent over the inclusive region, taps clamped to the region, float4 sums in
// the reference's order; the row above is clamped to region.top (the
// reference clamps it to region.bottom, a bug).  in/out: float4 per pixel, w*h; out's
// pixels outside the region are copied from in.
void orc_filter_pixels(uint32_t w, uint32_t h, const uint32_t* region, const float* in, float* out) {
  std::memcpy(out, in, size_t(w) * h * 4 * sizeof(float));
  const int left = (int)region[0], right = (int)region[1], bottom = (int)region[2], top = (int)region[3];
  const float wt[9] = {0.0625f, 0.125f, 0.0625f, 0.125f, 0.25f, 0.125f, 0.0625f, 0.125f, 0.0625f};
  for (int y = bottom; y <= top; ++y) {
    for (int x = left; x <= right; ++x) {
      const int xs[3] = {x - 1 < left ? left : x - 1, x, x + 1 > right ? right : x + 1};
      const int ys[3] = {y - 1 < bottom ? bottom : y - 1, y, y + 1 > top ? top : y + 1};
      float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int i = 0; i < 9; ++i) {
        const float* p = in + (size_t(ys[i / 3]) * w + xs[i % 3]) * 4;
        for (int c = 0; c < 4; ++c) v[c] = i == 0 ? p[c] * wt[0] : v[c] + p[c] * wt[i];
      }
      std::memcpy(out + (size_t(y) * w + x) * 4, v, sizeof(v));
    }
  }
}

// Exposed pieces for closed-form tests.
void orc_tex_sample(const void* desc, const uint8_t* texels, float u, float v, float* out) {
  const f3 c = tex_sample(*(const TexDesc*)desc, texels, u, v);
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
}
void orc_concentric_disk(float u1, float u2, float* out) { concentric_disk(u1, u2, out[0], out[1]); }
// onb and sample_cosine of the PBR path loop on float[3] vectors
void orc_onb(const float* n, float* t1, float* t2) {
  f3 a, b;
  onb(v3(n), a, b);
  t1[0] = a.x, t1[1] = a.y, t1[2] = a.z;
  t2[0] = b.x, t2[1] = b.y, t2[2] = b.z;
}
void orc_sample_cosine(const float* n, float u1, float u2, float* out) {
  const f3 w = sample_cosine(v3(n), u1, u2);
  out[0] = w.x, out[1] = w.y, out[2] = w.z;
}
// reflect (src/util/math.h:231-234) on float[3] vectors
void orc_reflect(const float* i, const float* n, float* out) {
  const f3 r = reflect(v3(i), v3(n));
  out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
// refract (src/util/math.h:235-246) on float[3] vectors; out is written only
// when it returns 1 (no total internal reflection)
int orc_refract(const float* i, const float* n, float eta1, float eta2, float* out) {
  f3 r;
  if (!refract(r, v3(i), v3(n), eta1, eta2)) return 0;
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  return 1;
}
// PBR sample point (x, y, z) of (seed, pixel, sample, dimension pair).
void orc_sobol_owen(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dims, uint32_t* out) {
  const u4 r = sobol_owen(seed, pixel, sample, dims);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
uint32_t orc_sobol_dim1(uint32_t i) { return sobol_dim1(i); }
uint32_t orc_nested_uniform_scramble(uint32_t x, uint32_t seed) { return nested_uniform_scramble(x, seed); }
void orc_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
  const u4 r = philox4x32_10(u4{ctr[0], ctr[1], ctr[2], ctr[3]}, k0, k1);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
}

}  // extern "C"
