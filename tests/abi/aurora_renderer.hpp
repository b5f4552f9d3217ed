// The reference's renderer interface over the C ABI: what a maintainer puts
// in place of MonteCarloRaytracer (src/render/monteCarloRaytracer.h:14-39,
// src/render/monteCarloRaytracer.cpp:23-127) and the upload half of
// Scene::update (src/core/scene.cpp:398-449).  Header-only C++17 over
// include/aurora_amd.h; no HIP or Maya types.  Statuses map to the three
// MStatus codes the reference returns (kSuccess, kFailure,
// kInsufficientMemory; monteCarloRaytracer.cpp:28-46).
#pragma once

#include <cstdint>
#include <vector>

#include "aurora_amd.h"

namespace aurora_amd {

enum class Status { kSuccess, kFailure, kInsufficientMemory };

inline Status to_status(aur_status s) {
  return s == AUR_OK ? Status::kSuccess : s == AUR_ERR_OOM ? Status::kInsufficientMemory : Status::kFailure;
}

// Rect of src/util/math.h: inclusive pixel bounds.
struct Rect {
  uint32_t left, right, bottom, top;
};

// Host arrays in the reference's upload layout (src/data/geometry.h:13-55).
struct Scene {
  std::vector<float> vertices, normals, texcoords;  // 9, 9 and 6 floats per triangle
  std::vector<uint16_t> shader_ids;                 // 1-based, 0 = default
  std::vector<aur_shader> shaders;
  std::vector<aur_light> lights;
  uint32_t count() const { return (uint32_t)shader_ids.size(); }
};

// The abstract Renderer of src/core/renderer.h:21-42 without Maya types:
// a camera record instead of an MDagPath, a float RGBA framebuffer (row 0 at
// the bottom) instead of RV_PIXEL.
class Renderer {
 public:
  virtual ~Renderer() = default;
  virtual Status createFrame(uint32_t width, uint32_t height, uint16_t samples, Scene* scene,
                             const aur_camera& camera) = 0;
  virtual Status destroyFrame() = 0;
  virtual Status setRegion(const Rect& region) = 0;
  virtual Status update() { return Status::kSuccess; }
  virtual Status render(bool ipr) = 0;
  virtual const float* framebuffer() = 0;
};

class MonteCarloRaytracer : public Renderer {
 public:
  explicit MonteCarloRaytracer(aur_context* ctx, uint32_t mode = AUR_MODE_COMPAT, uint32_t seed = 1)
      : ctx_(ctx), mode_(mode), seed_(seed) {}

  // Scene::update's uploads followed by cudaRebuildNMH (scene.cpp:398-449).
  Status update() override {
    if (!scene_) return Status::kFailure;
    const Scene& s = *scene_;
    aur_status st = aur_geometry_upload(ctx_, s.count(), s.vertices.data(), s.normals.data(),
                                        s.texcoords.empty() ? nullptr : s.texcoords.data(), s.shader_ids.data());
    if (st == AUR_OK) st = aur_shaders_upload(ctx_, (uint32_t)s.shaders.size(), s.shaders.data());
    if (st == AUR_OK)
      st = aur_lights_upload(ctx_, (uint32_t)s.lights.size(), s.lights.empty() ? nullptr : s.lights.data());
    if (st == AUR_OK) st = aur_rebuild_nmh(ctx_, nullptr);
    return to_status(st);
  }

  Status createFrame(uint32_t width, uint32_t height, uint16_t samples, Scene* scene,
                     const aur_camera& camera) override {
    scene_ = scene;
    camera_ = camera;
    samples_ = samples;
    width_ = width;
    height_ = height;
    Status st = update();
    if (st != Status::kSuccess) return st;
    framebuffer_.assign((size_t)width * height * 4, 0.0f);
    return to_status(aur_frame_create(ctx_, width, height, samples));
  }

  Status destroyFrame() override {
    framebuffer_.clear();
    return to_status(aur_frame_destroy(ctx_));
  }

  Status setRegion(const Rect& r) override {
    return to_status(aur_frame_set_region(ctx_, r.left, r.right, r.bottom, r.top));
  }

  // Final render: every sample; IPR: progressive passes, stopped by aur_cancel.
  Status render(bool ipr) override {
    aur_render_opts o{AUR_ABI_VERSION, mode_, samples_, 8, 3, seed_};
    if (!ipr) return to_status(aur_frame_render(ctx_, &camera_, &o));
    aur_status st = aur_frame_clear(ctx_);
    for (uint32_t s = 0; st == AUR_OK && s < samples_; ++s)
      st = aur_frame_render_samples(ctx_, &camera_, &o, s, s + 1);
    return to_status(st);
  }

  const float* framebuffer() override {
    if (aur_frame_read(ctx_, framebuffer_.data()) != AUR_OK) return nullptr;
    return framebuffer_.data();
  }

 private:
  aur_context* ctx_;
  uint32_t mode_, seed_;
  Scene* scene_ = nullptr;
  aur_camera camera_{};
  uint32_t samples_ = 1, width_ = 0, height_ = 0;
  std::vector<float> framebuffer_;
};

}  // namespace aurora_amd
