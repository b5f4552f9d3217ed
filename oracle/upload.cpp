// CPU oracle: object-space upload -- padToEven and cudaTransform.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.h).
//
// Restates, in order:
//   Geometry::padToEven   src/data/geometry.cpp:164-176 (intent: the pad
//                         triangle is a copy of the last one; the reference
//                         indexes floats instead of triangles)
//   cudaTransformKernel   src/kernels/utility.cu:96-125: triangle t takes
//                         the last transform, in array order, with
//                         offset <= t; no transform leaves it unwritten in
//                         the reference and unchanged here
//   Transform::getPoint / getNormal  src/kernels/lib/transform.cuh:8-24
//   Primitive3::readPoints / readValues  src/kernels/lib/primitive.cuh:46-72
// -0.0 results are folded to +0.0 (the upload's key order).
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Transform {  // src/util/transform.h:14-27
  uint32_t offset, size;
  float world[4][4];
  float normal[4][4];
};
static_assert(sizeof(Transform) == 136, "transform layout");

}  // namespace

extern "C" {

// out_v, out_n: (count + count % 2) * 9 floats each.
int orc_transform_objects(uint32_t count, const float* vertices, const float* normals, uint32_t nobj,
                          const void* transforms, float* out_v, float* out_n) {
  if (count == 0) return -1;
  const Transform* xf = (const Transform*)transforms;
  const uint32_t N = count + (count & 1u);
  std::vector<float> v(size_t(N) * 9), n(size_t(N) * 9, 0.0f);
  std::memcpy(v.data(), vertices, size_t(count) * 9 * sizeof(float));
  if (normals) std::memcpy(n.data(), normals, size_t(count) * 9 * sizeof(float));
  if (N != count) {
    std::memcpy(&v[size_t(count) * 9], &v[size_t(count - 1) * 9], 9 * sizeof(float));
    std::memcpy(&n[size_t(count) * 9], &n[size_t(count - 1) * 9], 9 * sizeof(float));
  }
  for (uint32_t t = 0; t < N; ++t) {
    const Transform* T = nullptr;
    for (uint32_t i = 0; i < nobj; ++i)
      if (xf[i].offset <= t) T = &xf[i];
    const float* sv = &v[size_t(t) * 9];
    const float* sn = &n[size_t(t) * 9];
    float* dv = out_v + size_t(t) * 9;
    float* dn = out_n + size_t(t) * 9;
    if (!T) {
      std::memcpy(dv, sv, 9 * sizeof(float));
      std::memcpy(dn, sn, 9 * sizeof(float));
    } else {
      for (int k = 0; k < 3; ++k) {
        const float px = sv[k], py = sv[3 + k], pz = sv[6 + k];
        dv[k] = T->world[0][0] * px + T->world[1][0] * py + T->world[2][0] * pz + T->world[3][0];
        dv[3 + k] = T->world[0][1] * px + T->world[1][1] * py + T->world[2][1] * pz + T->world[3][1];
        dv[6 + k] = T->world[0][2] * px + T->world[1][2] * py + T->world[2][2] * pz + T->world[3][2];
        const float nx = sn[3 * k], ny = sn[3 * k + 1], nz = sn[3 * k + 2];
        dn[3 * k] = T->normal[0][0] * nx + T->normal[1][0] * ny + T->normal[2][0] * nz;
        dn[3 * k + 1] = T->normal[0][1] * nx + T->normal[1][1] * ny + T->normal[2][1] * nz;
        dn[3 * k + 2] = T->normal[0][2] * nx + T->normal[1][2] * ny + T->normal[2][2] * nz;
      }
    }
    for (int i = 0; i < 9; ++i) {
      dv[i] = dv[i] + 0.0f;
      dn[i] = dn[i] + 0.0f;
    }
  }
  return 0;
}

}  // extern "C"
