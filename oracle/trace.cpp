// CPU oracle: NMH traversal and ray/triangle intersection, restated.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.h).  Follows
//   intersect()     src/kernels/lib/intersect.cuh:95-152  (closest hit)
//   intersectAny()  src/kernels/lib/intersect.cuh:36-92   (first hit)
//   min/max helpers src/kernels/lib/intersect.cuh:25-33
//   Ray::intersect(slab)     src/kernels/lib/ray.cuh:49-61
//   Ray::intersect(triangle) src/kernels/lib/ray.cuh:22-47 (one-sided)
//   Stack<TraversalState,32> src/kernels/lib/stack.cuh:8-26, src/config.h:19
// Geometry is in tree order ("points" layout, 9 floats per triangle).
// Counters per ray: popped = states popped with index < count (P);
// passed = states whose slab test passed (S).
#include <cstdint>
#include <cstring>

#include "oracle_math.h"

namespace {

using namespace orc;

struct Ray {
  f3 pos, dir;
  float t;
};

struct TState {
  uint32_t index;
  int axis;
  float rx, ry;
};

inline float vmin3(const float* v, int axis) {
  return fmin_nn(fmin_nn(v[3 * axis], v[3 * axis + 1]), v[3 * axis + 2]);
}
inline float vmax3(const float* v, int axis) {
  return fmax_nn(fmax_nn(v[3 * axis], v[3 * axis + 1]), v[3 * axis + 2]);
}

inline bool ray_slab(const Ray& r, float slo, float shi, int axis, float& rx, float& ry) {
  const float p = comp(r.pos, axis), d = comp(r.dir, axis);
  const float d0 = (slo - p) / d;
  const float d1 = (shi - p) / d;
  rx = fmax_nn(fmin_nn(d0, d1), rx);
  ry = fmin_nn(fmax_nn(d0, d1), ry);
  return (rx <= ry) && (rx <= r.t);
}

inline bool ray_tri(const Ray& r, const float* v, float& u, float& vv, float& t) {
  const f3 v1 = mk3(v[0], v[3], v[6]);
  const f3 v2 = mk3(v[1], v[4], v[7]);
  const f3 v3 = mk3(v[2], v[5], v[8]);
  const f3 e1 = v2 - v1;
  const f3 e2 = v3 - v1;
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

// Returns hit flag; for closest hit fills id/t/u/v.
template <bool kAny>
bool traverse(uint32_t count, const float* verts, Ray& ray, int32_t& id, float& hu, float& hv,
              uint32_t& popped, uint32_t& passed) {
  bool hit = false;
  TState stack[32];
  uint32_t size = 0;
  stack[size++] = TState{0, 0, kEpsilon, kInfinity};
  while (size > 0) {
    TState s = stack[--size];
    if (s.index >= count) continue;
    ++popped;
    const float* a = verts + size_t(s.index) * 9;
    const float* b = verts + size_t(s.index + 1) * 9;
    const float slo = fmin_nn(vmin3(a, s.axis), vmin3(b, s.axis));
    const float shi = fmax_nn(vmax3(a, s.axis), vmax3(b, s.axis));
    if (!ray_slab(ray, slo, shi, s.axis, s.rx, s.ry)) continue;
    ++passed;
    float u, v, t;
    if (ray_tri(ray, a, u, v, t) && t > kEpsilon && t < ray.t) {
      if (kAny) {
        id = (int32_t)s.index;
        return true;
      }
      hit = true;
      ray.t = t;
      hu = u;
      hv = v;
      id = (int32_t)s.index;
    }
    if (ray_tri(ray, b, u, v, t) && t > kEpsilon && t < ray.t) {
      if (kAny) {
        id = (int32_t)s.index + 1;
        return true;
      }
      hit = true;
      ray.t = t;
      hu = u;
      hv = v;
      id = (int32_t)s.index + 1;
    }
    const int caxis = (s.axis + 1) % 3;
    uint32_t L = 2 * s.index + 2;
    uint32_t R = 2 * s.index + 4;
    if (comp(ray.dir, caxis) < 0.0f) {
      const uint32_t tmp = L;
      L = R;
      R = tmp;
    }
    if (size + 2 > 32) return hit;  // the reference does not check its 32-entry stack; trees here stay < 32 deep
    stack[size++] = TState{R, caxis, s.rx, s.ry};
    stack[size++] = TState{L, caxis, s.rx, s.ry};
  }
  return hit;
}

}  // namespace

extern "C" {

// rays: n records of 8 floats (ox oy oz tmax dx dy dz unused).
// hits: n records of (int32 id, float t, float u, float v); id -1 on miss,
// t = tmax on miss.  counters (optional): 2 uint32 per ray (popped, passed).
// any != 0 runs intersectAny: id is the first accepted triangle, t/u/v unset.
int orc_trace(uint32_t count, const float* verts, uint64_t n, const float* rays, int any,
              void* hits, uint32_t* counters, int threads) {
  if (count < 2 || (count & 1u)) return -1;
  (void)threads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float* r = rays + size_t(i) * 8;
    Ray ray{mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3]};
    int32_t id = -1;
    float hu = 0.0f, hv = 0.0f;
    uint32_t popped = 0, passed = 0;
    if (any) {
      traverse<true>(count, verts, ray, id, hu, hv, popped, passed);
    } else {
      traverse<false>(count, verts, ray, id, hu, hv, popped, passed);
    }
    uint32_t* h = (uint32_t*)hits + size_t(i) * 4;
    float ft = any ? r[3] : ray.t;
    std::memcpy(&h[0], &id, 4);
    std::memcpy(&h[1], &ft, 4);
    std::memcpy(&h[2], &hu, 4);
    std::memcpy(&h[3], &hv, 4);
    if (counters) {
      counters[2 * i] = popped;
      counters[2 * i + 1] = passed;
    }
  }
  return 0;
}

// Ray::intersect(triangle) on one triangle for the brute-force tests.
int orc_ray_triangle(const float* ray8, const float* tri9, float* out_uvt) {
  Ray ray{mk3(ray8[0], ray8[1], ray8[2]), mk3(ray8[4], ray8[5], ray8[6]), ray8[3]};
  float u = 0, v = 0, t = 0;
  const bool ok = ray_tri(ray, tri9, u, v, t);
  out_uvt[0] = u;
  out_uvt[1] = v;
  out_uvt[2] = t;
  return ok ? 1 : 0;
}

}  // extern "C"
