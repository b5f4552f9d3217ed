// C++17 consumer of include/aurora_amd.h and tests/abi/aurora_renderer.hpp.
//
// 1. Mirrors of the reference's records with the CUDA layout (float3: 12 B,
//    align 4; float2: 8 B, align 8) and the reference's field names
//    (src/util/shader.h:23-37, src/util/light.h:24-35,
//    src/util/transform.h:22-25, src/util/camera.h:30-35,
//    src/util/emitter.h:13-18) must have the C ABI records' offsets, so a
//    maintainer can pass the reference's arrays straight through.  The one
//    size that differs: Camera is 64 B (tail padding after aspect, from the
//    8-byte float2), aur_camera 60 B; offsets agree, and the library reads
//    60 B, so a Camera* may be passed as an aur_camera*.
// 2. With a device, renders a fixed two-triangle floor under an area light
//    (8 x 8 pixels, 4 spp) through MonteCarloRaytracer in compat and PBR
//    mode, final and IPR (progressive) render, and writes each float RGBA
//    framebuffer to argv[1] (4 images of 256 floats, in that order).
//    Without a device prints "cpp-consumer no-device <status>".
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "aurora_renderer.hpp"

namespace ref {
struct float3 {
  float x, y, z;
};
struct alignas(8) float2 {
  float x, y;
};
struct Shader {
  int type;
  float3 diffuseColor, emissionColor, specularColor, reflectionColor;
  float diffuse, emission, exponent, reflectivity, translucence, refractiveIndex;
  unsigned int texture;
};
struct Light {
  int type;
  float intensity, ambient, area;
  short samples;
  float3 position, color, direction, e1, e2;
};
struct Transform {
  unsigned int offset, size;
  float worldMatrix[4][4], normalMatrix[4][4];
};
struct Camera {
  float3 position, forward, right, up;
  float2 tanfov;
  float aspect;
};
struct Emitter {
  float area, pdf, cdf;
  float3 power;
  unsigned int triangleID;
};
}  // namespace ref

#define SAME(R, A, rf, af) static_assert(offsetof(ref::R, rf) == offsetof(A, af), #R "." #rf)
static_assert(sizeof(ref::Shader) == sizeof(aur_shader) && sizeof(aur_shader) == 80, "Shader");
SAME(Shader, aur_shader, diffuseColor, diffuse_color);
SAME(Shader, aur_shader, emissionColor, emission_color);
SAME(Shader, aur_shader, specularColor, specular_color);
SAME(Shader, aur_shader, reflectionColor, reflection_color);
SAME(Shader, aur_shader, diffuse, diffuse);
SAME(Shader, aur_shader, emission, emission);
SAME(Shader, aur_shader, exponent, exponent);
SAME(Shader, aur_shader, reflectivity, reflectivity);
SAME(Shader, aur_shader, translucence, translucence);
SAME(Shader, aur_shader, refractiveIndex, refractive_index);
SAME(Shader, aur_shader, texture, texture);
static_assert(sizeof(ref::Light) == sizeof(aur_light) && sizeof(aur_light) == 80, "Light");
SAME(Light, aur_light, intensity, intensity);
SAME(Light, aur_light, ambient, ambient);
SAME(Light, aur_light, area, area);
SAME(Light, aur_light, samples, samples);
SAME(Light, aur_light, position, position);
SAME(Light, aur_light, color, color);
SAME(Light, aur_light, direction, direction);
SAME(Light, aur_light, e1, e1);
SAME(Light, aur_light, e2, e2);
static_assert(sizeof(ref::Transform) == sizeof(aur_transform) && sizeof(aur_transform) == 136, "Transform");
SAME(Transform, aur_transform, size, size);
SAME(Transform, aur_transform, worldMatrix, world);
SAME(Transform, aur_transform, normalMatrix, normal);
static_assert(sizeof(ref::Camera) == 64 && sizeof(aur_camera) == 60, "Camera: 60 B read of a 64 B record");
SAME(Camera, aur_camera, position, position);
SAME(Camera, aur_camera, forward, forward);
SAME(Camera, aur_camera, right, right);
SAME(Camera, aur_camera, up, up);
SAME(Camera, aur_camera, tanfov, tanfov);
SAME(Camera, aur_camera, aspect, aspect);
static_assert(sizeof(ref::Emitter) == sizeof(aur_emitter) && sizeof(aur_emitter) == 28, "Emitter");
SAME(Emitter, aur_emitter, pdf, pdf);
SAME(Emitter, aur_emitter, cdf, cdf);
SAME(Emitter, aur_emitter, power, power);
SAME(Emitter, aur_emitter, triangleID, triangle);

using aurora_amd::MonteCarloRaytracer;
using aurora_amd::Status;

static aurora_amd::Scene floor_scene() {
  aurora_amd::Scene s;
  // (-5,0,-5) (-5,0,5) (5,0,5) and (-5,0,-5) (5,0,5) (5,0,-5), facing +y, points layout
  s.vertices = {-5, -5, 5, 0, 0, 0, -5, 5, 5,  //
                -5, 5, 5, 0, 0, 0, -5, 5, -5};
  s.normals.assign(18, 0.0f);
  for (int v = 0; v < 6; ++v) s.normals[3 * v + 1] = 1.0f;
  s.shader_ids = {1, 1};
  aur_shader sh{};
  sh.type = 0;
  sh.diffuse_color[0] = 0.5f;
  sh.diffuse_color[1] = 0.4f;
  sh.diffuse_color[2] = 0.3f;
  sh.diffuse = 1.0f;
  sh.refractive_index = 1.0f;
  s.shaders.push_back(sh);
  aur_light L{};
  L.type = AUR_LIGHT_AREA;
  L.intensity = 1.0f;
  L.area = 1.0f;
  L.samples = 1;
  const float pos[3] = {-0.5f, 2.0f, -0.5f}, col[3] = {10, 10, 10}, dir[3] = {0, -1, 0}, e1[3] = {1, 0, 0},
              e2[3] = {0, 0, 1};
  std::memcpy(L.position, pos, sizeof pos);
  std::memcpy(L.color, col, sizeof col);
  std::memcpy(L.direction, dir, sizeof dir);
  std::memcpy(L.e1, e1, sizeof e1);
  std::memcpy(L.e2, e2, sizeof e2);
  s.lights.push_back(L);
  return s;
}

int main(int argc, char** argv) {
  aur_context* ctx = nullptr;
  const aur_status st = aur_create(0, &ctx);
  if (st != AUR_OK) {
    std::printf("cpp-consumer no-device %d\n", (int)st);
    return st == AUR_ERR_DEVICE ? 0 : 1;
  }
  if (argc < 2) {
    std::printf("usage: cpp_consumer out.f32\n");
    aur_destroy(ctx);
    return 2;
  }
  aurora_amd::Scene scene = floor_scene();
  const aur_camera cam{{0.0f, 3.0f, -4.0f}, {0.0f, -0.6f, 0.8f}, {-0.8f, 0.0f, 0.0f},
                       {0.0f, 0.64f, 0.48f}, {0.6f, 0.6f}, 1.0f};
  std::FILE* f = std::fopen(argv[1], "wb");
  int bad = f ? 0 : 1;
  for (uint32_t mode : {(uint32_t)AUR_MODE_COMPAT, (uint32_t)AUR_MODE_PBR}) {
    for (bool ipr : {false, true}) {
      MonteCarloRaytracer r(ctx, mode, 5);
      if (r.createFrame(8, 8, 4, &scene, cam) != Status::kSuccess || r.render(ipr) != Status::kSuccess) {
        std::printf("cpp-consumer render failed: %s\n", aur_last_error(ctx));
        bad = 1;
        continue;
      }
      const float* px = r.framebuffer();
      if (!px || !f || std::fwrite(px, sizeof(float), 256, f) != 256) bad = 1;
      if (r.destroyFrame() != Status::kSuccess) bad = 1;
    }
  }
  if (f) std::fclose(f);
  aur_destroy(ctx);
  std::printf("cpp-consumer %s\n", bad ? "failed" : "rendered");
  return bad;
}
