// CPU oracle: scene bounds, light-power CDF and emitter table.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.h).
//
// Restates, in order of execution:
//   computeBounds            src/kernels/rebuilder.cu:291-309 (over every
//                            vertex; the reference reads the first two
//                            triangles only)
//   Light::power             src/kernels/lib/light.cuh:91-106
//   luminosity               src/util/math.h:248-250
//   getBoundingSphereRadius  src/kernels/lib/common.cuh:74-77
//   cudaCreateLightsCDF      src/kernels/light.cu:20-60
//   cudaCreateEmitters       src/kernels/emitter.cu:20-119
//   Primitive3::area         src/kernels/lib/primitive.cuh:74-85
// Per-light powers, areas and pdfs are float expressions in the reference's
// order.  The reference's float inclusive scans (Thrust, order unspecified)
// are restated as sequential prefix sums in double, each value divided by
// the double total and rounded once to float.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle_math.h"

namespace {
using namespace orc;

struct Light {  // layout of the C ABI's aur_light (80 B, src/util/light.h:24-35)
  int32_t type;
  float intensity, ambient, area;
  int16_t samples, pad;
  float position[3], color[3], direction[3], e1[3], e2[3];
};
struct Shader {  // aur_shader (80 B, src/util/shader.h:23-37)
  int32_t type;
  float diffuseColor[3], emissionColor[3], specularColor[3], reflectionColor[3];
  float diffuse, emission, exponent, reflectivity, translucence, refractiveIndex;
  uint32_t texture;
};
struct Emitter {  // aur_emitter (28 B, src/util/emitter.h:10-22)
  float area, pdf, cdf;
  float power[3];
  uint32_t triangleID;
};
static_assert(sizeof(Light) == 80 && sizeof(Shader) == 80 && sizeof(Emitter) == 28, "layouts");

inline float luminosity(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

float light_power(const Light& L, f3 bmin, f3 bmax) {
  switch (L.type) {
    case 0:  // ambient
      return 4.0f * kPi * luminosity(L.color) * L.intensity * (1.0f - L.ambient);
    case 1:  // point
      return 4.0f * kPi * luminosity(L.color) * L.intensity;
    case 2: {  // directional
      const float R = 0.5f * length(bmin - bmax);
      return luminosity(L.color) * L.intensity * kPi * R * R;
    }
    default:  // area
      return luminosity(L.color) * L.intensity * L.area * kPi;
  }
}

// Primitive3::area with readPoints' vertex order (points layout)
float triangle_area(const float* v) {
  const f3 v1 = mk3(v[0], v[3], v[6]), v2 = mk3(v[1], v[4], v[7]), v3 = mk3(v[2], v[5], v[8]);
  const f3 va = v1 - v2, vb = v1 - v3, vc = v2 - v3;
  const float a = dot(va, va), b = dot(vb, vb), c = dot(vc, vc);
  return sqrtf((2 * a * b + 2 * b * c + 2 * c * a - a * a - b * b - c * c) / 16.0f);
}
}  // namespace

extern "C" {

// bmin/bmax: per-axis min and max over the 3 vertices of every triangle.
void orc_scene_bounds(uint32_t count, const float* verts, float* bmin, float* bmax) {
  for (int a = 0; a < 3; ++a) {
    float lo = verts[3 * a], hi = verts[3 * a];
    for (uint32_t t = 0; t < count; ++t)
      for (int k = 0; k < 3; ++k) {
        const float x = verts[(size_t)t * 9 + 3 * a + k];
        lo = fmin_nn(lo, x);
        hi = fmax_nn(hi, x);
      }
    bmin[a] = lo;
    bmax[a] = hi;
  }
}

// Light::power of each light (powers: n floats).
void orc_light_powers(uint32_t n, const void* lights, const float* bmin, const float* bmax, float* powers) {
  const Light* L = static_cast<const Light*>(lights);
  const f3 lo = mk3(bmin[0], bmin[1], bmin[2]), hi = mk3(bmax[0], bmax[1], bmax[2]);
  for (uint32_t i = 0; i < n; ++i) powers[i] = light_power(L[i], lo, hi);
}

// cudaCreateLightsCDF: cdf[i] = prefix(i) / total; uniform when total == 0.
uint32_t orc_lights_cdf(uint32_t n, const void* lights, const float* bmin, const float* bmax, float* cdf) {
  const Light* L = static_cast<const Light*>(lights);
  const f3 lo = mk3(bmin[0], bmin[1], bmin[2]), hi = mk3(bmax[0], bmax[1], bmax[2]);
  double total = 0.0;
  for (uint32_t i = 0; i < n; ++i) total += (double)light_power(L[i], lo, hi);
  double run = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    run += (double)light_power(L[i], lo, hi);
    cdf[i] = total == 0.0 ? (float)((double)(i + 1) / (double)n) : (float)(run / total);
  }
  return n;
}

// cudaCreateEmitters over tree-order geometry: stable compaction of the
// triangles whose shader emission is > 0 (copy_if, emitter.cu:100), area and
// pdf per emitter (:59-65), inclusive scan of the emissions (:110) and
// normalisation by the last value (:113-117).  Returns the emitter count;
// out holds at least that many records.
uint32_t orc_emitters(uint32_t count, const float* verts, const uint16_t* shaderIds, uint32_t nShaders,
                      const void* shaders, void* out) {
  const Shader* S = static_cast<const Shader*>(shaders);
  Emitter* E = static_cast<Emitter*>(out);
  uint32_t n = 0;
  double total = 0.0;
  for (uint32_t t = 0; t < count; ++t) {
    uint32_t sid = shaderIds[t] > 0 ? shaderIds[t] - 1u : 0u;  // getSafeID
    if (sid >= nShaders) sid = 0;
    if (!(S[sid].emission > 0.0f)) continue;
    total += (double)S[sid].emission;
  }
  double run = 0.0;
  for (uint32_t t = 0; t < count; ++t) {
    uint32_t sid = shaderIds[t] > 0 ? shaderIds[t] - 1u : 0u;
    if (sid >= nShaders) sid = 0;
    const Shader& sh = S[sid];
    if (!(sh.emission > 0.0f)) continue;
    Emitter& e = E[n++];
    e.area = triangle_area(verts + (size_t)t * 9);
    e.pdf = 1.0f / e.area;
    run += (double)sh.emission;
    e.cdf = (float)(run / total);
    std::memcpy(e.power, sh.emissionColor, sizeof(e.power));
    e.triangleID = t;
  }
  return n;
}

}  // extern "C"
