// Batched rendering: many jobs (camera + scene + resolution + spp) share
// every pass of the PBR wavefront.  No reference counterpart: the reference
// renders one frame of one scene per MonteCarloRaytracer::render call
// (src/render/monteCarloRaytracer.cpp:81-120).
//
// Scenes are built once with the context's NMH builder and kept (tree-order
// nodes, normals, shader ids, shaders, lights).  A batch numbers every job's
// paths globally (jobs in ascending order of path count, so that short jobs
// finish in the first passes) and runs the passes of pbr_run over all of
// them; each job's image equals its single-frame render bit for bit, because
// a path's random numbers depend only on (job seed, pixel, sample) and each
// pixel accumulates its samples in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "context.h"

static_assert(sizeof(aur_job) == 104, "aur_job layout");

extern "C" {

aur_status aur_scene_add(aur_context* ctx, uint32_t count, const float* vertices, const float* normals,
                         const uint16_t* shader_ids, uint32_t num_shaders, const aur_shader* shaders,
                         uint32_t num_lights, const aur_light* lights, uint32_t* scene_id) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!scene_id) return ctx->fail(AUR_ERR_INVALID_ARG, "scene: null id");
  if (ctx->scenes.size() >= 65536) return ctx->fail(AUR_ERR_INVALID_ARG, "scene: more than 65536 scenes");
  for (uint32_t i = 0; i < num_shaders && shaders; ++i)
    if (shaders[i].texture != 0)
      return ctx->fail(AUR_ERR_UNSUPPORTED, "scene: texture-mapped shaders are not kept for batches");
  // geometry, shaders and lights are checked by the single-scene entry points
  aur_status st = aur_geometry_upload(ctx, count, vertices, normals, nullptr, shader_ids);
  if (st != AUR_OK) return st;
  st = aur_shaders_upload(ctx, num_shaders, shaders);
  if (st != AUR_OK) return st;
  st = aur_lights_upload(ctx, num_lights, lights);
  if (st != AUR_OK) return st;
  st = aur::nmh_build(ctx, nullptr);
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto sc = std::make_unique<aur::SceneStore>();
  sc->count = ctx->count;
  sc->levels = ctx->levels;
  sc->num_shaders = ctx->num_shaders;
  sc->num_lights = ctx->num_lights;
  sc->nslots = ctx->num_slots;
  // take the built arrays; the context's own scene is left unbuilt
  sc->nodes.swap(ctx->nodes);
  sc->root_slab = ctx->root_slab;
  sc->rec_bytes = ctx->rec_bytes;
  sc->normals.swap(ctx->normals);
  sc->shader_ids.swap(ctx->shader_ids);
  sc->shaders.swap(ctx->shaders);
  sc->lights.swap(ctx->lights);
  ctx->built = false;
  ctx->has_tb = false;
  ctx->num_shaders = 0;
  ctx->num_lights = 0;
  ctx->num_slots = 0;
  ctx->compat_slots = 0;
  *scene_id = (uint32_t)ctx->scenes.size();
  ctx->scenes.push_back(std::move(sc));
  return AUR_OK;
}

aur_status aur_scene_clear(aur_context* ctx) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->scenes.clear();
  return AUR_OK;
}

uint32_t aur_scene_count(const aur_context* ctx) { return ctx ? (uint32_t)ctx->scenes.size() : 0u; }

aur_status aur_batch_render(aur_context* ctx, uint32_t njobs, const aur_job* jobs, float* rgba,
                            float* job_ms) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (njobs == 0) return AUR_OK;
  if (!jobs || !rgba) return ctx->fail(AUR_ERR_INVALID_ARG, "batch: null argument");
  if (njobs > (1u << 20)) return ctx->fail(AUR_ERR_INVALID_ARG, "batch: more than 2^20 jobs");
  for (uint32_t i = 0; i < njobs; ++i) {
    const aur_job& j = jobs[i];
    if (j.abi_version != AUR_ABI_VERSION) return ctx->fail(AUR_ERR_UNSUPPORTED, "batch: ABI version");
    if (j.scene >= ctx->scenes.size()) return ctx->fail(AUR_ERR_INVALID_ARG, "batch: unknown scene");
    if (j.width == 0 || j.height == 0 || j.spp == 0 || j.max_depth == 0)
      return ctx->fail(AUR_ERR_INVALID_ARG, "batch: zero width, height, spp or max_depth");
    if ((uint64_t)j.width * j.height * j.spp > 0x7FFFFFFFull)
      return ctx->fail(AUR_ERR_INVALID_ARG, "batch: a job's pixels * spp exceeds 2^31");
    if (ctx->scenes[j.scene]->num_shaders == 0) return ctx->fail(AUR_ERR_STATE, "batch: scene without shaders");
    if (ctx->scenes[j.scene]->nslots > aur::kMaxSlots)
      return ctx->fail(AUR_ERR_UNSUPPORTED, "batch: a scene has more than 64 next-event slots (light samples) per vertex");
  }
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  // shortest jobs first
  std::vector<uint32_t> order(njobs);
  std::iota(order.begin(), order.end(), 0u);
  auto paths_of = [&](uint32_t i) { return (uint64_t)jobs[i].width * jobs[i].height * jobs[i].spp; };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return paths_of(x) < paths_of(y); });
  std::vector<aur::JobDev> dev(njobs);
  aur::PbrRun run;
  unsigned long long start = 0;
  for (uint32_t k = 0; k < njobs; ++k) {
    const aur_job& j = jobs[order[k]];
    const aur::SceneStore& sc = *ctx->scenes[j.scene];
    aur::JobDev& J = dev[k];
    J = aur::JobDev{};
    J.nodes = sc.nodes.p;
    J.normals = sc.normals.p;
    J.shader_ids = sc.shader_ids.p;
    J.shaders = sc.shaders.p;
    J.lights = sc.lights.p;
    J.pixels = reinterpret_cast<float4*>(rgba) + j.out_offset;
    J.path_start = start;
    J.n_nodes = sc.count / 2;
    J.root = sc.root_slab;
    J.rec_bytes = sc.rec_bytes;
    J.num_shaders = sc.num_shaders;
    J.num_lights = sc.num_lights;
    J.nslots = sc.nslots;
    J.w = j.width;
    J.h = j.height;
    J.npix = j.width * j.height;
    J.spp = j.spp;
    J.max_depth = j.max_depth;
    J.rr_start = j.rr_start;
    J.seed = j.seed;
    J.paths = J.npix * j.spp;
    J.rx0 = 0;
    J.ry0 = 0;
    J.rw = j.width;
    J.cam = j.camera;
    start += J.paths;
    run.levels = std::max(run.levels, sc.levels);
    run.nslots = std::max(run.nslots, sc.nslots);
    run.max_depth = std::max(run.max_depth, j.max_depth);
  }
  run.jobs = dev.data();
  run.njobs = njobs;
  run.clear_images = true;
  run.begin = 0;
  run.end = start;
  run.cap = (uint32_t)std::min<unsigned long long>(std::max<uint32_t>(ctx->render_params.max_paths, 1u), start);
  std::vector<float> pass_ms;
  if (job_ms) run.pass_end_ms = &pass_ms;
  aur::FrameState& fr = ctx->frame;
  AUR_HIP(ctx, fr.counters.reserve(aur::kCounterWords));
  unsigned long long* stats = reinterpret_cast<unsigned long long*>(fr.counters.p + aur::kStatsOffset);
  AUR_HIP(ctx, hipMemsetAsync(stats, 0, 2 * sizeof(unsigned long long), ctx->stream));
  aur_status st = aur::pbr_run(ctx, run, stats);
  if (st != AUR_OK) return st;
  unsigned long long hs[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  fr.rays_traced = hs[0] + hs[1];
  fr.rays_extension = hs[0];
  fr.rays_shadow = hs[1];
  if (job_ms) {
    // a job is complete at the end of the pass holding its last path
    for (uint32_t k = 0; k < njobs; ++k) {
      const unsigned long long last = dev[k].path_start + dev[k].paths - 1;
      job_ms[order[k]] = pass_ms[last / run.cap];
    }
  }
  return AUR_OK;
}

}  // extern "C"
