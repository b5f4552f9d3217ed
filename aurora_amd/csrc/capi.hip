// C ABI entry points (include/aurora_amd.h): argument checking, uploads,
// read-back and the context lifecycle.  Nothing here throws across the ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "context.h"

using aur::DevBuf;

namespace {

__global__ void k_clear_f4(float4* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

bool finite_all(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

}  // namespace

namespace aur {
namespace {
std::mutex g_hook_mutex;
aur_alloc_hook g_hook{nullptr, nullptr, nullptr};
}  // namespace

aur_alloc_hook current_alloc_hook() {
  std::lock_guard<std::mutex> lock(g_hook_mutex);
  return g_hook;
}
}  // namespace aur

extern "C" {

uint32_t aur_abi_version(void) { return AUR_ABI_VERSION; }

aur_status aur_create(int32_t device, aur_context** out) {
  if (!out) return AUR_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return AUR_ERR_DEVICE;
  if (device < 0 || device >= ndev) return AUR_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return AUR_ERR_DEVICE;
  aur_context* ctx = new (std::nothrow) aur_context();
  if (!ctx) return AUR_ERR_OOM;
  ctx->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->num_cus = (uint32_t)cus;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return AUR_ERR_DEVICE;
  }
  // device address of the allocation-free entry points' scratch on this
  // device (once per device); on failure only those entry points refuse
  (void)aur::scratch_setup(device);
  *out = ctx;
  return AUR_OK;
}

aur_status aur_set_trace_params(aur_context* ctx, const aur_trace_params* p) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!p) return ctx->fail(AUR_ERR_INVALID_ARG, "trace params: null");
  if (p->block != 64 && p->block != 128 && p->block != 256)
    return ctx->fail(AUR_ERR_INVALID_ARG, "trace params: block must be 64, 128 or 256");
  if (p->persistent && (p->refill < 1 || p->refill > 64))
    return ctx->fail(AUR_ERR_INVALID_ARG, "trace params: refill must be in 1..64");
  if (p->waves_per_cu > 64) return ctx->fail(AUR_ERR_INVALID_ARG, "trace params: waves_per_cu > 64");
  if (p->stack_lds_rows > 32) return ctx->fail(AUR_ERR_INVALID_ARG, "trace params: stack_lds_rows > 32");
  ctx->trace_params = *p;
  return AUR_OK;
}

aur_status aur_get_trace_params(aur_context* ctx, aur_trace_params* p) {
  if (!ctx || !p) return AUR_ERR_INVALID_ARG;
  *p = ctx->trace_params;
  return AUR_OK;
}

aur_status aur_set_render_params(aur_context* ctx, const aur_render_params* p) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!p) return ctx->fail(AUR_ERR_INVALID_ARG, "render params: null");
  if (p->max_paths > (1u << 26)) return ctx->fail(AUR_ERR_INVALID_ARG, "render params: max_paths > 2^26");
  if (p->graph > 1) return ctx->fail(AUR_ERR_INVALID_ARG, "render params: graph must be 0 or 1");
  if (p->refill > 64) return ctx->fail(AUR_ERR_INVALID_ARG, "render params: refill must be in 0..64");
  if (p->stack_lds_rows > 64)
    return ctx->fail(AUR_ERR_INVALID_ARG, "render params: stack_lds_rows must be in 0..64");
  ctx->render_params.max_paths = p->max_paths ? p->max_paths : aur::kDefaultMaxPaths;
  ctx->render_params.graph = p->graph;
  ctx->render_params.refill = p->refill ? p->refill : aur::kDefaultRenderRefill;
  ctx->render_params.stack_lds_rows = p->stack_lds_rows ? p->stack_lds_rows : aur::kDefaultRenderStackRows;
  return AUR_OK;
}

aur_status aur_get_render_params(aur_context* ctx, aur_render_params* p) {
  if (!ctx || !p) return AUR_ERR_INVALID_ARG;
  *p = ctx->render_params;
  return AUR_OK;
}

void aur_destroy(aur_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  hipStream_t s = ctx->stream;
  delete ctx;  // frees every DevBuf
  (void)hipStreamDestroy(s);
}

const char* aur_last_error(const aur_context* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

void* aur_stream(aur_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

aur_status aur_synchronize(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

aur_status aur_set_alloc_hook(const aur_alloc_hook* hook) {
  if (hook && (!hook->alloc || !hook->free)) return AUR_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lock(aur::g_hook_mutex);
  aur::g_hook = hook ? *hook : aur_alloc_hook{nullptr, nullptr, nullptr};
  return AUR_OK;
}

aur_status aur_cancel(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  ctx->cancel.store(1u, std::memory_order_release);
  return AUR_OK;
}

aur_status aur_geometry_upload(aur_context* ctx, uint32_t count, const float* vertices,
                               const float* normals, const float* texcoords,
                               const uint16_t* shader_ids) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (count < 2 || (count & 1u)) return ctx->fail(AUR_ERR_INVALID_ARG, "geometry: count must be even and >= 2");
  if (count > (1u << 30)) return ctx->fail(AUR_ERR_INVALID_ARG, "geometry: count exceeds 2^30");
  if (!vertices) return ctx->fail(AUR_ERR_INVALID_ARG, "geometry: null vertices");
  const size_t nv = (size_t)count * 9;
  if (!finite_all(vertices, nv)) return ctx->fail(AUR_ERR_INVALID_ARG, "geometry: non-finite vertex");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  // -0.0 -> +0.0 so that the float key order equals the radix order
  std::vector<float> v(vertices, vertices + nv);
  for (float& x : v) x = x + 0.0f;
  std::vector<float> nrm(nv, 0.0f), uv((size_t)count * 6, 0.0f);
  if (normals) {
    std::memcpy(nrm.data(), normals, nv * sizeof(float));
    for (float& x : nrm) x = x + 0.0f;
  }
  if (texcoords) std::memcpy(uv.data(), texcoords, uv.size() * sizeof(float));
  std::vector<uint16_t> ids(count, 0);
  if (shader_ids) std::memcpy(ids.data(), shader_ids, count * sizeof(uint16_t));
  ctx->built = false;
  ctx->has_tb = false;
  ctx->count = 0;
  AUR_HIP(ctx, ctx->in_vertices.reserve(nv));
  AUR_HIP(ctx, ctx->in_normals.reserve(nv));
  AUR_HIP(ctx, ctx->in_texcoords.reserve((size_t)count * 6));
  AUR_HIP(ctx, ctx->in_shader_ids.reserve(count));
  AUR_HIP(ctx, hipMemcpy(ctx->in_vertices.p, v.data(), nv * sizeof(float), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(ctx->in_normals.p, nrm.data(), nv * sizeof(float), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(ctx->in_texcoords.p, uv.data(), uv.size() * sizeof(float), hipMemcpyHostToDevice));
  AUR_HIP(ctx, hipMemcpy(ctx->in_shader_ids.p, ids.data(), count * sizeof(uint16_t), hipMemcpyHostToDevice));
  ctx->count = count;
  return AUR_OK;
}

aur_status aur_shaders_upload(aur_context* ctx, uint32_t count, const aur_shader* shaders) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (count == 0 || !shaders) return ctx->fail(AUR_ERR_INVALID_ARG, "shaders: empty");
  bool textured = false, branching = false, children = false;
  for (uint32_t i = 0; i < count; ++i) {
    if (shaders[i].type != 0 && shaders[i].type != 1)
      return ctx->fail(AUR_ERR_INVALID_ARG, "shaders: unknown type");
    textured = textured || shaders[i].texture != 0;
    branching = branching || (shaders[i].reflectivity > 0.0f && shaders[i].translucence > 0.0f);
    children = children || shaders[i].reflectivity > 0.0f || shaders[i].translucence > 0.0f;
  }
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, ctx->shaders.reserve(count));
  AUR_HIP(ctx, hipMemcpy(ctx->shaders.p, shaders, count * sizeof(aur_shader), hipMemcpyHostToDevice));
  ctx->num_shaders = count;
  ctx->textured_shaders = textured;
  ctx->branching_shaders = branching;
  ctx->child_shaders = children;
  return AUR_OK;
}

aur_status aur_textures_upload(aur_context* ctx, uint32_t count, const aur_texture* textures,
                               const uint8_t* texels, uint64_t texel_bytes) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (count && (!textures || !texels)) return ctx->fail(AUR_ERR_INVALID_ARG, "textures: null");
  std::vector<aur::TexDesc> desc(count);
  for (uint32_t i = 0; i < count; ++i) {
    const aur_texture& t = textures[i];
    if (t.width == 0 || t.height == 0 || t.width > 65536 || t.height > 65536)
      return ctx->fail(AUR_ERR_INVALID_ARG, "textures: size must be 1..65536 per side");
    const uint64_t bytes = 4ull * t.width * t.height;
    if (t.offset > texel_bytes || bytes > texel_bytes - t.offset)
      return ctx->fail(AUR_ERR_INVALID_ARG, "textures: texels outside the buffer");
    desc[i] = aur::TexDesc{t.offset, t.width, t.height};
  }
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->num_textures = 0;
  if (count) {
    AUR_HIP(ctx, ctx->textures.reserve(count));
    AUR_HIP(ctx, ctx->texels.reserve(texel_bytes));
    AUR_HIP(ctx, hipMemcpy(ctx->textures.p, desc.data(), count * sizeof(aur::TexDesc), hipMemcpyHostToDevice));
    AUR_HIP(ctx, hipMemcpy(ctx->texels.p, texels, texel_bytes, hipMemcpyHostToDevice));
  }
  ctx->num_textures = count;
  return AUR_OK;
}

aur_status aur_lights_upload(aur_context* ctx, uint32_t count, const aur_light* lights) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (count && !lights) return ctx->fail(AUR_ERR_INVALID_ARG, "lights: null");
  for (uint32_t i = 0; i < count; ++i) {
    if (lights[i].type < 0 || lights[i].type > 3) return ctx->fail(AUR_ERR_INVALID_ARG, "lights: unknown type");
    if (lights[i].type == AUR_LIGHT_AREA && lights[i].samples < 0)
      return ctx->fail(AUR_ERR_INVALID_ARG, "lights: negative sample count");
  }
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  if (count) {
    AUR_HIP(ctx, ctx->lights.reserve(count));
    AUR_HIP(ctx, hipMemcpy(ctx->lights.p, lights, count * sizeof(aur_light), hipMemcpyHostToDevice));
  }
  ctx->num_lights = count;
  ctx->num_slots = aur::light_set_slots(count, lights);
  uint64_t cslots = 0;  // compat: the estimator's shadow rays per hit
  for (uint32_t i = 0; i < count; ++i) cslots += lights[i].type == AUR_LIGHT_AREA ? (uint64_t)lights[i].samples : 1u;
  ctx->compat_slots = cslots > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)cslots;
  return AUR_OK;
}

aur_status aur_rebuild_nmh(aur_context* ctx, aur_build_stats* stats) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  aur_status st = aur::nmh_build(ctx, stats);
  if (st != AUR_OK) return st;
  st = aur::generate_tb(ctx);
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

aur_status aur_generate_tb(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  aur_status st = aur::generate_tb(ctx);
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

uint32_t aur_geometry_count(const aur_context* ctx) { return ctx ? ctx->count : 0; }

aur_status aur_read_permutation(aur_context* ctx, uint32_t* perm) {
  if (!ctx || !perm) return AUR_ERR_INVALID_ARG;
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "read: geometry not built");
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  AUR_HIP(ctx, hipMemcpy(perm, ctx->perm.p, ctx->count * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return AUR_OK;
}

aur_status aur_read_geometry(aur_context* ctx, float* vertices, float* normals, float* texcoords,
                             uint16_t* shader_ids, float* tangents, float* bitangents) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "read: geometry not built");
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = ctx->count;
  if (vertices) AUR_HIP(ctx, hipMemcpy(vertices, ctx->vertices.p, n * 36, hipMemcpyDeviceToHost));
  if (normals) AUR_HIP(ctx, hipMemcpy(normals, ctx->normals.p, n * 36, hipMemcpyDeviceToHost));
  if (texcoords) AUR_HIP(ctx, hipMemcpy(texcoords, ctx->texcoords.p, n * 24, hipMemcpyDeviceToHost));
  if (shader_ids) AUR_HIP(ctx, hipMemcpy(shader_ids, ctx->shader_ids.p, n * 2, hipMemcpyDeviceToHost));
  if ((tangents || bitangents) && !ctx->has_tb) return ctx->fail(AUR_ERR_STATE, "read: tangents not generated");
  if (tangents) AUR_HIP(ctx, hipMemcpy(tangents, ctx->tangents.p, n * 36, hipMemcpyDeviceToHost));
  if (bitangents) AUR_HIP(ctx, hipMemcpy(bitangents, ctx->bitangents.p, n * 36, hipMemcpyDeviceToHost));
  return AUR_OK;
}

aur_status aur_trace_closest(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  return aur::trace_closest(ctx, n, rays, hits);
}

aur_status aur_trace_closest_stats(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                                   uint64_t* popped, uint64_t* passed) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  return aur::trace_closest_count(ctx, n, rays, hits, popped, passed);
}

aur_status aur_check_division(aur_context* ctx, uint64_t samples, uint64_t seed,
                              uint64_t* recip_checked, uint64_t* recip_bad, uint64_t* quot_bad) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  return aur::division_check(ctx, samples, seed, recip_checked, recip_bad, quot_bad);
}

aur_status aur_trace_any(aur_context* ctx, uint64_t n, const aur_ray* rays, int32_t* hit_ids) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  return aur::trace_any(ctx, n, rays, hit_ids);
}

aur_status aur_frame_create(aur_context* ctx, uint32_t width, uint32_t height, uint32_t spp) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (width == 0 || height == 0 || spp == 0) return ctx->fail(AUR_ERR_INVALID_ARG, "frame: zero size");
  if ((uint64_t)width * height > (1ull << 28)) return ctx->fail(AUR_ERR_INVALID_ARG, "frame: too large");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  aur::FrameState& fr = ctx->frame;
  fr.width = fr.height = fr.spp = 0;
  AUR_HIP(ctx, fr.pixels.reserve((size_t)width * height));
  fr.width = width;
  fr.height = height;
  fr.spp = spp;
  fr.left = fr.bottom = 0;
  fr.right = width - 1;
  fr.top = height - 1;
  return aur_frame_clear(ctx);
}

aur_status aur_frame_set_region(aur_context* ctx, uint32_t left, uint32_t right, uint32_t bottom, uint32_t top) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "region: no frame");
  if (left > right || bottom > top || right >= fr.width || top >= fr.height)
    return ctx->fail(AUR_ERR_INVALID_ARG, "region: outside the frame or empty");
  fr.left = left;
  fr.right = right;
  fr.bottom = bottom;
  fr.top = top;
  return AUR_OK;
}

aur_status aur_frame_get_region(aur_context* ctx, uint32_t* left, uint32_t* right, uint32_t* bottom, uint32_t* top) {
  if (!ctx || !left || !right || !bottom || !top) return AUR_ERR_INVALID_ARG;
  const aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "region: no frame");
  *left = fr.left;
  *right = fr.right;
  *bottom = fr.bottom;
  *top = fr.top;
  return AUR_OK;
}

aur_status aur_frame_filter(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  return aur::frame_filter(ctx);
}

aur_status aur_frame_write(aur_context* ctx, const float* rgba) {
  if (!ctx || !rgba) return AUR_ERR_INVALID_ARG;
  aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "frame: none");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  AUR_HIP(ctx, hipMemcpy(fr.pixels.p, rgba, sizeof(float4) * fr.width * fr.height, hipMemcpyHostToDevice));
  return AUR_OK;
}

aur_status aur_frame_destroy(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->frame.~FrameState();
  new (&ctx->frame) aur::FrameState();
  return AUR_OK;
}

aur_status aur_frame_clear(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "frame: none");
  const uint32_t n = fr.width * fr.height;
  k_clear_f4<<<(n + 255) / 256, 256, 0, ctx->stream>>>(fr.pixels.p, n);
  AUR_LAUNCH_CHECK(ctx, "k_clear_f4");
  return AUR_OK;
}

aur_status aur_frame_render(aur_context* ctx, const aur_camera* camera, const aur_render_opts* opts) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!camera || !opts) return ctx->fail(AUR_ERR_INVALID_ARG, "render: null argument");
  if (opts->abi_version != AUR_ABI_VERSION) return ctx->fail(AUR_ERR_UNSUPPORTED, "render: ABI version");
  aur_status st = aur::frame_render_check(ctx, opts, 0, opts->spp);  // a refused render keeps the frame
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  st = aur_frame_clear(ctx);
  if (st != AUR_OK) return st;
  return aur::frame_render(ctx, camera, opts, 0, opts->spp);
}

aur_status aur_frame_render_samples(aur_context* ctx, const aur_camera* camera,
                                    const aur_render_opts* opts, uint32_t sample_begin,
                                    uint32_t sample_end) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!camera || !opts) return ctx->fail(AUR_ERR_INVALID_ARG, "render: null argument");
  if (opts->abi_version != AUR_ABI_VERSION) return ctx->fail(AUR_ERR_UNSUPPORTED, "render: ABI version");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  return aur::frame_render(ctx, camera, opts, sample_begin, sample_end);
}

aur_status aur_frame_read(aur_context* ctx, float* rgba) {
  if (!ctx || !rgba) return AUR_ERR_INVALID_ARG;
  aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "frame: none");
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  AUR_HIP(ctx, hipMemcpy(rgba, fr.pixels.p, sizeof(float4) * fr.width * fr.height, hipMemcpyDeviceToHost));
  return AUR_OK;
}

aur_status aur_frame_device_buffer(aur_context* ctx, float** rgba) {
  if (!ctx || !rgba) return AUR_ERR_INVALID_ARG;
  if (ctx->frame.width == 0) return ctx->fail(AUR_ERR_STATE, "frame: none");
  *rgba = reinterpret_cast<float*>(ctx->frame.pixels.p);
  return AUR_OK;
}

aur_status aur_frame_copy(aur_context* ctx, float* rgba) {
  if (!ctx || !rgba) return AUR_ERR_INVALID_ARG;
  aur::FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "frame: none");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, hipMemcpyAsync(rgba, fr.pixels.p, sizeof(float4) * fr.width * fr.height,
                              hipMemcpyDeviceToDevice, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

uint64_t aur_frame_rays_traced(const aur_context* ctx) { return ctx ? ctx->frame.rays_traced : 0; }

aur_status aur_frame_ray_counts(const aur_context* ctx, uint64_t* camera_and_extension, uint64_t* shadow) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (camera_and_extension) *camera_and_extension = ctx->frame.rays_extension;
  if (shadow) *shadow = ctx->frame.rays_shadow;
  return AUR_OK;
}

aur_status aur_frame_adjoint(aur_context* ctx, const aur_camera* camera, const aur_render_opts* opts,
                             const float* d_image, float* grad_albedo, float* grad_radiance) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!camera || !opts) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: null argument");
  if (opts->abi_version != AUR_ABI_VERSION) return ctx->fail(AUR_ERR_UNSUPPORTED, "adjoint: ABI version");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  return aur::frame_adjoint(ctx, camera, opts, d_image, grad_albedo, grad_radiance);
}

}  // extern "C"
