// Object-space upload: padding to an even count and the per-object transform
// (Scene::updateMeshes staging, src/core/scene.cpp:85-240; padToEven,
// src/data/geometry.cpp:164-176; cudaTransform, src/kernels/utility.cu:96-134).
//
// One thread per triangle finds its transform (the last one in array order
// whose offset <= t: a binary search when the offsets ascend, the
// reference's linear scan otherwise), maps the three positions by the
// row-vector world matrix and the three normals by the normal matrix, in the
// reference's operation order without contraction, folds -0.0 to +0.0 (the
// build's key order) and flags non-finite results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"

static_assert(sizeof(aur_transform) == 136, "aur_transform layout");

namespace aur {
namespace {

__device__ inline float3 xf_point(const aur_transform& T, float x, float y, float z) {
  return make_float3(T.world[0][0] * x + T.world[1][0] * y + T.world[2][0] * z + T.world[3][0],
                     T.world[0][1] * x + T.world[1][1] * y + T.world[2][1] * z + T.world[3][1],
                     T.world[0][2] * x + T.world[1][2] * y + T.world[2][2] * z + T.world[3][2]);
}

__device__ inline float3 xf_normal(const aur_transform& T, float x, float y, float z) {
  return make_float3(T.normal[0][0] * x + T.normal[1][0] * y + T.normal[2][0] * z,
                     T.normal[0][1] * x + T.normal[1][1] * y + T.normal[2][1] * z,
                     T.normal[0][2] * x + T.normal[1][2] * y + T.normal[2][2] * z);
}

__device__ inline float canon(float v) { return v + 0.0f; }

__global__ void k_transform(const float* __restrict__ sv, const float* __restrict__ sn, uint32_t count,
                            const aur_transform* __restrict__ xf, uint32_t nobj, int ascending,
                            float* __restrict__ dv, float* __restrict__ dn, uint32_t* __restrict__ bad) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  int found = -1;
  if (ascending) {
    uint32_t lo = 0, hi = nobj;  // first index with offset > t
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (xf[mid].offset <= t) lo = mid + 1; else hi = mid;
    }
    found = (int)lo - 1;
  } else {
    for (uint32_t i = 0; i < nobj; ++i)
      if (xf[i].offset <= t) found = (int)i;
  }
  const float* v = sv + (size_t)t * 9;
  const float* n = sn + (size_t)t * 9;
  float ov[9], on[9];
  if (found < 0) {
    for (int i = 0; i < 9; ++i) {
      ov[i] = v[i];
      on[i] = n[i];
    }
  } else {
    const aur_transform& T = xf[found];
    for (int k = 0; k < 3; ++k) {
      // points layout: x1 x2 x3 y1 y2 y3 z1 z2 z3
      const float3 p = xf_point(T, v[k], v[3 + k], v[6 + k]);
      ov[k] = p.x;
      ov[3 + k] = p.y;
      ov[6 + k] = p.z;
      // values layout: v1xyz v2xyz v3xyz
      const float3 q = xf_normal(T, n[3 * k], n[3 * k + 1], n[3 * k + 2]);
      on[3 * k] = q.x;
      on[3 * k + 1] = q.y;
      on[3 * k + 2] = q.z;
    }
  }
  bool finite = true;
  for (int i = 0; i < 9; ++i) {
    ov[i] = canon(ov[i]);
    on[i] = canon(on[i]);
    finite = finite && isfinite(ov[i]);
    dv[(size_t)t * 9 + i] = ov[i];
    dn[(size_t)t * 9 + i] = on[i];
  }
  if (!finite) atomicOr(bad, 1u);
}

}  // namespace
}  // namespace aur

extern "C" aur_status aur_geometry_upload_objects(aur_context* ctx, uint32_t count, const float* vertices,
                                                  const float* normals, const float* texcoords,
                                                  const uint16_t* shader_ids, uint32_t num_objects,
                                                  const aur_transform* transforms) {
  using aur::DevBuf;
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (count == 0) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: no triangles");
  if (count >= (1u << 30)) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: count exceeds 2^30");
  if (!vertices) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: null vertices");
  if (num_objects && !transforms) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: null transforms");
  if (num_objects > (1u << 24)) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: more than 2^24 objects");
  const size_t nv = (size_t)count * 9;
  for (size_t i = 0; i < nv; ++i)
    if (!std::isfinite(vertices[i])) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: non-finite vertex");
  int ascending = 1;
  for (uint32_t i = 0; i < num_objects; ++i) {
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c)
        if (!std::isfinite(transforms[i].world[r][c]) || !std::isfinite(transforms[i].normal[r][c]))
          return ctx->fail(AUR_ERR_INVALID_ARG, "objects: non-finite transform");
    if (i && transforms[i].offset < transforms[i - 1].offset) ascending = 0;
  }
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  // pad to an even count by duplicating the last triangle (triangle-wise)
  const uint32_t N = count + (count & 1u);
  std::vector<float> v((size_t)N * 9), nrm((size_t)N * 9, 0.0f), uv((size_t)N * 6, 0.0f);
  std::vector<uint16_t> ids(N, 0);
  std::memcpy(v.data(), vertices, nv * sizeof(float));
  if (normals) std::memcpy(nrm.data(), normals, nv * sizeof(float));
  if (texcoords) std::memcpy(uv.data(), texcoords, (size_t)count * 6 * sizeof(float));
  if (shader_ids) std::memcpy(ids.data(), shader_ids, count * sizeof(uint16_t));
  if (N != count) {
    std::memcpy(&v[(size_t)count * 9], &v[(size_t)(count - 1) * 9], 9 * sizeof(float));
    std::memcpy(&nrm[(size_t)count * 9], &nrm[(size_t)(count - 1) * 9], 9 * sizeof(float));
    std::memcpy(&uv[(size_t)count * 6], &uv[(size_t)(count - 1) * 6], 6 * sizeof(float));
    ids[count] = ids[count - 1];
  }
  // transform into staging buffers; the context's geometry (and a built
  // tree) changes only once the transformed vertices are known to be finite
  DevBuf<float> sv, sn, tv, tn, tuv;
  DevBuf<uint16_t> tids;
  DevBuf<aur_transform> dxf;
  DevBuf<uint32_t> bad;
  AUR_HIP(ctx, sv.reserve((size_t)N * 9));
  AUR_HIP(ctx, sn.reserve((size_t)N * 9));
  AUR_HIP(ctx, bad.reserve(1));
  AUR_HIP(ctx, dxf.reserve(std::max<uint32_t>(num_objects, 1)));
  AUR_HIP(ctx, tv.reserve((size_t)N * 9));
  AUR_HIP(ctx, tn.reserve((size_t)N * 9));
  AUR_HIP(ctx, tuv.reserve((size_t)N * 6));
  AUR_HIP(ctx, tids.reserve(N));
  AUR_HIP(ctx, hipMemcpy(sv.p, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(sn.p, nrm.data(), nrm.size() * sizeof(float), hipMemcpyHostToDevice));
  if (num_objects)
    AUR_HIP(ctx, hipMemcpy(dxf.p, transforms, num_objects * sizeof(aur_transform), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(tuv.p, uv.data(), uv.size() * sizeof(float), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(tids.p, ids.data(), N * sizeof(uint16_t), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemsetAsync(bad.p, 0, sizeof(uint32_t), ctx->stream));
  aur::k_transform<<<(N + 255) / 256, 256, 0, ctx->stream>>>(sv.p, sn.p, N, dxf.p, num_objects, ascending,
                                                              tv.p, tn.p, bad.p);
  AUR_LAUNCH_CHECK(ctx, "k_transform");
  uint32_t hbad = 0;
  AUR_HIP(ctx, hipMemcpyAsync(&hbad, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hbad) return ctx->fail(AUR_ERR_INVALID_ARG, "objects: a transformed vertex is not finite");
  ctx->in_vertices.swap(tv);
  ctx->in_normals.swap(tn);
  ctx->in_texcoords.swap(tuv);
  ctx->in_shader_ids.swap(tids);
  ctx->built = false;
  ctx->has_tb = false;
  ctx->count = N;
  return AUR_OK;
}
