/* aurora_amd: MI355X (gfx950) render core of the Aurora renderer, C ABI.
 *
 * This is the drop-in boundary below the reference's renderer layer.  The
 * reference calls C++ free functions declared in src/kernels/kernels.h:25-64
 * from its Renderer/Scene/Geometry classes; each entry point below names the
 * reference interface it replaces.  Conventions:
 *   - every call returns aur_status (0 = AUR_OK); nothing throws across the
 *     ABI; aur_last_error() describes the last failure of a context;
 *   - a context owns one HIP device stream and all device memory it allocates;
 *     a context is used by one thread at a time, distinct contexts are
 *     independent;
 *   - "host" pointers are ordinary CPU memory and the call is synchronous;
 *     "device" pointers are HIP device memory and the work is queued on the
 *     context's stream (aur_stream / aur_synchronize).
 *
 * Geometry layout (src/data/geometry.h:13-55): per triangle, vertices in the
 * "points" layout x1 x2 x3 y1 y2 y3 z1 z2 z3 (9 floats), normals in the
 * "values" layout v1xyz v2xyz v3xyz (9 floats), 6 texture coordinates and a
 * uint16 shader id (1-based, 0 = default, src/util/array.h:12-18).
 */
#ifndef AURORA_AMD_H
#define AURORA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AUR_ABI_VERSION 3u

typedef int32_t aur_status;
enum {
  AUR_OK = 0,
  AUR_ERR_OOM = 1,
  AUR_ERR_INVALID_ARG = 2,
  AUR_ERR_DEVICE = 3,
  AUR_ERR_CANCELLED = 4,
  AUR_ERR_UNSUPPORTED = 5,
  AUR_ERR_STATE = 6
};

typedef struct aur_context aur_context;

/* Shader record, field for field the reference's Shader (80 B,
 * src/util/shader.h:15-43; attribute mapping src/core/scene.cpp:242-291). */
typedef struct aur_shader {
  int32_t type; /* 0 Lambert, 1 Phong */
  float diffuse_color[3];
  float emission_color[3];
  float specular_color[3];
  float reflection_color[3];
  float diffuse, emission, exponent, reflectivity, translucence, refractive_index;
  uint32_t texture; /* 1-based; 0 = none */
} aur_shader;

/* Light record, field for field the reference's Light (80 B,
 * src/util/light.h:14-51; built at src/core/scene.cpp:328-396).  Radiance is
 * color * intensity.  Point: position, falling off as 1/d^2.  Directional:
 * direction = the way the light travels.  Ambient: the fraction `ambient`
 * lights every point unshadowed, the rest comes from position with a shadow
 * ray and no falloff.  Area: corner position, edges e1 and e2, emitting
 * along direction, area = |e1| |e2|, `samples` shadow rays per hit.
 * Renders refuse (AUR_ERR_UNSUPPORTED) light sets with more than 64 shadow
 * rays per path vertex in PBR mode (an ambient light takes two) or more than
 * 4096 per hit in compat mode. */
enum { AUR_LIGHT_AMBIENT = 0, AUR_LIGHT_POINT = 1, AUR_LIGHT_DIRECTIONAL = 2, AUR_LIGHT_AREA = 3 };
typedef struct aur_light {
  int32_t type;
  float intensity, ambient, area;
  int16_t samples, pad;
  float position[3], color[3], direction[3], e1[3], e2[3];
} aur_light;

/* Pinhole camera as built by the reference host code (src/util/camera.h:12-36). */
typedef struct aur_camera {
  float position[3], forward[3], right[3], up[3], tanfov[2], aspect;
} aur_camera;

/* Ray record for the trace entry points: origin, tmax (the reference's
 * Ray::t, src/util/ray.h:44-48), direction.  The slab range always starts at
 * Epsilon = 1e-7 as in src/kernels/lib/intersect.cuh:40,100. */
typedef struct aur_ray {
  float org[3];
  float tmax;
  float dir[3];
  float reserved;
} aur_ray;

/* Hit record: tree-order triangle index (-1 on miss), t (tmax on miss) and
 * barycentrics; the reference's HitPoint fields triangleID/u/v
 * (src/util/hitpoint.h:10-24) plus the accepted ray.t. */
typedef struct aur_hit {
  int32_t id;
  float t, u, v;
} aur_hit;

enum { AUR_MODE_COMPAT = 0, AUR_MODE_PBR = 1 };
typedef struct aur_render_opts {
  uint32_t abi_version; /* AUR_ABI_VERSION */
  uint32_t mode;        /* AUR_MODE_COMPAT: the reference's MC ray tracer; AUR_MODE_PBR: path loop */
  uint32_t spp;         /* samples per pixel (the reference's Dim::depth) */
  uint32_t max_depth;   /* PBR: bounces per path (compat mode ignores it: the reference's
                           fixed MAX_DEPTH 8 of mirror and refraction rays) */
  uint32_t rr_start;    /* PBR: first bounce that applies Russian roulette (ignored in compat) */
  uint32_t seed;        /* compat: Philox key; PBR: scrambling seed of the Sobol sample points */
} aur_render_opts;

typedef struct aur_build_stats {
  uint32_t count;   /* triangles */
  uint32_t levels;  /* log2i(count/2)+1 */
  float build_ms;   /* device time of the build (bounds .. gather) */
} aur_build_stats;

/* ---- context (replaces Engine's device setup, src/core/engine.cpp:33-63) ---- */
aur_status aur_create(int32_t device, aur_context** out);
void aur_destroy(aur_context* ctx);
const char* aur_last_error(const aur_context* ctx);
uint32_t aur_abi_version(void);
void* aur_stream(aur_context* ctx); /* hipStream_t of the context */
aur_status aur_synchronize(aur_context* ctx);
/* Stops the render running on ctx at its next pass boundary (or, with none
 * running, the next render at its first): that call returns
 * AUR_ERR_CANCELLED, its frame holding the passes that completed, and the
 * request is consumed.  The one call that may run concurrently with another
 * call on the same context (from another thread).  Covers aur_frame_render,
 * aur_frame_render_samples, aur_frame_adjoint and aur_batch_render; passes
 * are queued at most two ahead of the GPU so that a cancel takes effect
 * within about two passes (one sample per compat pass).  Replaces the
 * reference's stop-on-interrupt of the IPR loop (src/core/engine.cpp:258-317,
 * m_pause / MComputation::isInterruptRequested). */
aur_status aur_cancel(aur_context* ctx);
/* Device memory allocator for every buffer the library allocates from now
 * on, in any context (no reference counterpart: the reference calls
 * cudaMalloc / cudaFree directly, src/util/array.h:22-32).  alloc returns
 * device memory of the current device, aligned to at least 256 bytes, or
 * NULL when out of memory (the call then fails with AUR_ERR_OOM); free
 * receives exactly the pointers alloc returned, after the library has
 * synchronised the device (as hipFree does), so no queued work still uses
 * the memory.  A buffer is freed by the
 * allocator that made it, so the hook may change while contexts live.  NULL
 * restores hipMalloc / hipFree.  Set it while no other library call runs. */
typedef struct aur_alloc_hook {
  void* (*alloc)(size_t bytes, void* user);
  void (*free)(void* ptr, void* user);
  void* user;
} aur_alloc_hook;
aur_status aur_set_alloc_hook(const aur_alloc_hook* hook);

/* ---- scene upload (replaces Geometry::resize/copyToDevice and
 * Scene::updateShaders/updateLights uploads: src/data/geometry.cpp:88-151,
 * src/core/scene.cpp:242-396).  Host arrays; the caller keeps ownership.
 * count must be even (Geometry::resize pads to even, geometry.cpp:90); -0.0
 * is stored as +0.0; NaN/Inf vertices are rejected. ---- */
aur_status aur_geometry_upload(aur_context* ctx, uint32_t count, const float* vertices,
                               const float* normals, const float* texcoords,
                               const uint16_t* shader_ids);
aur_status aur_shaders_upload(aur_context* ctx, uint32_t count, const aur_shader* shaders);
aur_status aur_lights_upload(aur_context* ctx, uint32_t count, const aur_light* lights);
/* RGBA8 textures (replaces Scene::updateTextures / Texture::load,
 * src/core/scene.cpp:298-326, src/data/texture.cpp:15-30; sampled as
 * texfetch, src/kernels/lib/texunit.cuh:54-94).  Texture i (shader.texture =
 * i + 1) is width x height texels of 4 bytes, row-major, at byte `offset` of
 * `texels`.  Bilinear with wrap addressing on normalised coordinates, value
 * / 255; a shader whose texture index exceeds the count samples white.
 * Replaces all textures of the context. */
typedef struct aur_texture {
  uint32_t width, height;
  uint64_t offset;
} aur_texture;
aur_status aur_textures_upload(aur_context* ctx, uint32_t count, const aur_texture* textures,
                               const uint8_t* texels, uint64_t texel_bytes);

/* ---- NMH build (replaces cudaRebuildNMH(Geometry&), src/kernels/rebuilder.cu:311-450,
 * declared at src/kernels/kernels.h:45).  Reorders the uploaded geometry into
 * tree order on the device and prepares the traversal records. ---- */
aur_status aur_rebuild_nmh(aur_context* ctx, aur_build_stats* stats);

/* ---- tangent frames (replaces cudaGenerateTB, src/kernels/utility.cu:169-180,
 * src/kernels/kernels.h:35).  Runs on the tree-order geometry. ---- */
aur_status aur_generate_tb(aur_context* ctx);

/* Read back (host) for parity checks.  perm[i] = original index of the
 * triangle at tree position i.  Any array pointer may be NULL. */
aur_status aur_read_permutation(aur_context* ctx, uint32_t* perm);
aur_status aur_read_geometry(aur_context* ctx, float* vertices, float* normals, float* texcoords,
                             uint16_t* shader_ids, float* tangents, float* bitangents);
uint32_t aur_geometry_count(const aur_context* ctx);

/* ---- traversal (the device functions intersect()/intersectAny(),
 * src/kernels/lib/intersect.cuh:36-152, as standalone launches).  Device
 * pointers; queued on the context stream. ---- */
aur_status aur_trace_closest(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits);
aur_status aur_trace_any(aur_context* ctx, uint64_t n, const aur_ray* rays, int32_t* hit_ids);
/* Launch shape of the trace entry points (no reference counterpart: the
 * reference launched one thread per ray in blocks of 64,
 * src/kernels/monteCarloRaytracer.cu:119-124).  persistent = 1: resident
 * waves each own a contiguous ray range and refill idle lanes once `refill`
 * (1..64) lanes are idle; waves_per_cu = 0 uses the occupancy
 * limit.  persistent = 0: one lane per ray.  block: 64, 128 or 256.
 * stack_lds_rows: how many rows of each lane's stack of deferred ranges
 * (a compact stack of at most levels - 1 entries) live in LDS (0 = all);
 * deeper entries spill to global memory, leaving LDS for more waves.
 * Results do not depend on these values. */
typedef struct aur_trace_params {
  uint32_t persistent, block, refill, waves_per_cu, stack_lds_rows;
} aur_trace_params;
aur_status aur_set_trace_params(aur_context* ctx, const aur_trace_params* params);
aur_status aur_get_trace_params(aur_context* ctx, aur_trace_params* params);
/* Closest hit plus the reference loop's traversal counters summed over the
 * rays: popped = states popped with index < count, passed = states whose slab
 * test passed.  Synchronous.  Used for the algorithmic-bytes model
 * 32 + 16 + 24*popped + 48*passed bytes per ray. */
aur_status aur_trace_closest_stats(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                                   uint64_t* popped, uint64_t* passed);
/* Device self-check of the traversal's division shortcuts (no reference
 * counterpart: the reference divides directly).  recip_bad counts floats x
 * with |x| in [2^-60, 2^64] (all recip_checked of them) whose shortcut
 * reciprocal differs from 1.0f / x; quot_bad counts, over `samples` hashed pairs (n, d) in the
 * shortcut's range, reciprocal quotients that differ from n / d.  Both must be
 * 0.  Synchronous. */
aur_status aur_check_division(aur_context* ctx, uint64_t samples, uint64_t seed,
                              uint64_t* recip_checked, uint64_t* recip_bad, uint64_t* quot_bad);

/* ---- frame (replaces MonteCarloRaytracer::createFrame/destroyFrame/
 * setRegion/render/framebuffer, src/render/monteCarloRaytracer.cpp:23-127,
 * with the ray generation / trace / draw kernels of src/kernels/kernels.h:28-30,39-42,51-53). ---- */
aur_status aur_frame_create(aur_context* ctx, uint32_t width, uint32_t height, uint32_t spp);
aur_status aur_frame_destroy(aur_context* ctx);
/* Clears the accumulation buffer and renders opts->spp samples. */
aur_status aur_frame_render(aur_context* ctx, const aur_camera* camera,
                            const aur_render_opts* opts);
/* Path-loop launch shape (no reference counterpart: the reference traced
 * 65,536-ray tiles one sample at a time, src/render/monteCarloRaytracer.cpp:13).
 * max_paths: paths in flight per wavefront pass; a frame pass traces
 * clamp(max_paths / pixels, 1, remaining) samples of every pixel, a batch
 * pass max_paths paths, a compat-mode round as many whole samples of every
 * pixel (or pixels of one sample) as fit max_paths entries.  0 selects the
 * default (2^24).
 * graph: 1 captures one PBR pass (camera, bounce loop, accumulation) in a
 * HIP graph and replays it for every pass; 0 launches the kernels directly.
 * refill: the render's persistent traversal launches refill idle lanes once
 * this many (1..64) are idle; 0 selects the default (16).
 * stack_lds_rows: how many rows of each lane's traversal stack live in LDS
 * in those launches; deeper rows go to global memory, leaving LDS for more
 * resident waves.  0 selects the default (12); a value of at least the
 * tree's levels - 1 keeps every row in LDS.
 * Images do not depend on any of them: samples are accumulated in sample
 * order. */
typedef struct aur_render_params {
  uint32_t max_paths;
  uint32_t graph;
  uint32_t refill;
  uint32_t stack_lds_rows;
} aur_render_params;
aur_status aur_set_render_params(aur_context* ctx, const aur_render_params* params);
aur_status aur_get_render_params(aur_context* ctx, aur_render_params* params);
/* Renders samples [sample_begin, sample_end) of opts->spp into the existing
 * accumulation (progressive rendering and resume; every random number is a
 * function of (seed, pixel, sample, draw), so a split render equals an
 * unsplit one). */
aur_status aur_frame_render_samples(aur_context* ctx, const aur_camera* camera,
                                    const aur_render_opts* opts, uint32_t sample_begin,
                                    uint32_t sample_end);
aur_status aur_frame_clear(aur_context* ctx);
/* Render region, inclusive pixel bounds (MonteCarloRaytracer::setRegion,
 * src/render/monteCarloRaytracer.cpp:75-79; Rect left/right/bottom/top).
 * Renders trace and accumulate only region pixels, each exactly as a full
 * render would (the random numbers are keyed by frame pixel).
 * aur_frame_create resets it to the whole frame. */
aur_status aur_frame_set_region(aur_context* ctx, uint32_t left, uint32_t right, uint32_t bottom,
                                uint32_t top);
aur_status aur_frame_get_region(aur_context* ctx, uint32_t* left, uint32_t* right, uint32_t* bottom,
                                uint32_t* top);
/* 3x3 tent filter (1/16 1/8 1/16; 1/8 1/4 1/8; 1/16 1/8 1/16) of the region
 * in place, taps clamped to the region (cudaFilterPixels,
 * src/kernels/utility.cu:231-268; the reference clamps the row above to
 * region.bottom, a bug, and never calls it).  Synchronous. */
aur_status aur_frame_filter(aur_context* ctx);
/* Loads host float RGBA (width*height*4) into the accumulation buffer: with
 * aur_frame_render_samples, resumes a progressive render from a saved
 * accumulation and sample count; the result equals an uninterrupted render. */
aur_status aur_frame_write(aur_context* ctx, const float* rgba);
/* Host float RGBA, width*height*4 floats, row 0 = bottom (framebuffer(),
 * src/render/monteCarloRaytracer.cpp:122-127). */
aur_status aur_frame_read(aur_context* ctx, float* rgba);
/* Copies the float4 accumulation buffer (width*height*4 floats, same layout
 * as aur_frame_read) to device memory at rgba; synchronous. */
aur_status aur_frame_copy(aur_context* ctx, float* rgba);
/* Device pointer of the float4 accumulation buffer. */
aur_status aur_frame_device_buffer(aur_context* ctx, float** rgba);
/* Number of ray segments (camera + extension + shadow) traced by the last
 * aur_frame_render / aur_frame_render_samples call. */
uint64_t aur_frame_rays_traced(const aur_context* ctx);
/* The same count split into camera + extension (closest hit) and shadow
 * (any hit) rays. */
aur_status aur_frame_ray_counts(const aur_context* ctx, uint64_t* camera_and_extension, uint64_t* shadow);

/* ---- adjoint (no reference counterpart: the reference renders only; this
 * is the differentiable form of the AUR_MODE_PBR path loop) ----
 * Gradient of a scalar loss of the frame image with respect to each shader's
 * albedo (diffuse_color * diffuse, RGB) and each light's radiance
 * (color * intensity, RGB), for the image aur_frame_render produces with the
 * same camera, opts, geometry, shaders and lights.  The paths are replayed
 * (same counter-based random numbers), recording each vertex's shader,
 * Russian-roulette probability and next-event weights in device memory for
 * one pass of paths at a time, then swept backwards.  Russian-roulette
 * probabilities are treated as constants.  Device pointers, queued on the
 * context stream and synchronised before return:
 *   d_image       float4 per pixel, dLoss/dI in the framebuffer layout (A unused)
 *   grad_albedo   float[3 * shader count], overwritten
 *   grad_radiance float[3 * light count], overwritten (may be NULL without lights)
 * Requires mode AUR_MODE_PBR, max_depth <= 16, shaders + lights <= 2048.
 * Deterministic: the same inputs give the same bits. */
aur_status aur_frame_adjoint(aur_context* ctx, const aur_camera* camera, const aur_render_opts* opts,
                             const float* d_image, float* grad_albedo, float* grad_radiance);

/* ---- object-space upload (replaces Scene::updateMeshes' staging,
 * Geometry::padToEven and cudaTransform: src/core/scene.cpp:85-240,
 * src/data/geometry.cpp:164-176, src/kernels/utility.cu:96-134) ----
 * Layout of the reference's Transform (src/util/transform.h:14-27). */
typedef struct aur_transform {
  uint32_t offset, size; /* first triangle and triangle count of the object */
  float world[4][4];     /* row-vector matrix, translation in row 3: p' = (p, 1) world */
  float normal[4][4];    /* inverse transpose of world; n' = n normal[0..2][0..2] */
} aur_transform;
/* Uploads object-space triangles (layouts as aur_geometry_upload; count may
 * be odd: the last triangle is duplicated to an even count, triangle-wise,
 * where the reference copied floats) and transforms them on the device:
 * triangle t takes the last transform, in array order, whose offset <= t
 * (the pad triangle included); positions by `world`, normals by the upper
 * 3x3 of `normal` (not renormalised); uvs and shader ids unchanged.
 * Triangles before every offset keep their object-space values (the
 * reference left them unwritten).  Transformed coordinates must be finite
 * (AUR_ERR_INVALID_ARG otherwise).  Then call aur_rebuild_nmh as after
 * aur_geometry_upload. */
aur_status aur_geometry_upload_objects(aur_context* ctx, uint32_t count, const float* vertices,
                                       const float* normals, const float* texcoords,
                                       const uint16_t* shader_ids, uint32_t num_objects,
                                       const aur_transform* transforms);

/* ---- batches (no reference counterpart: the reference renders one frame
 * of one scene per MonteCarloRaytracer::render call,
 * src/render/monteCarloRaytracer.cpp:81-120) ----
 * aur_scene_add builds a scene with the context's NMH builder and keeps it
 * for batches; it replaces the context's current geometry, shaders and
 * lights (upload again before aur_frame_render).  Arguments as for
 * aur_geometry_upload / aur_shaders_upload / aur_lights_upload. */
aur_status aur_scene_add(aur_context* ctx, uint32_t count, const float* vertices, const float* normals,
                         const uint16_t* shader_ids, uint32_t num_shaders, const aur_shader* shaders,
                         uint32_t num_lights, const aur_light* lights, uint32_t* scene_id);
aur_status aur_scene_clear(aur_context* ctx);
uint32_t aur_scene_count(const aur_context* ctx);

/* One job: a PBR render (AUR_MODE_PBR semantics) of a kept scene. */
typedef struct aur_job {
  uint32_t abi_version; /* AUR_ABI_VERSION */
  uint32_t scene;       /* id from aur_scene_add */
  uint32_t width, height, spp, max_depth, rr_start, seed;
  uint64_t out_offset;  /* first pixel of the job's image in rgba, in float4 units */
  aur_camera camera;
} aur_job;

/* Renders every job; all jobs share each wavefront pass (render params:
 * max_paths paths per pass, graph = 1 replays one captured pass).  rgba:
 * device buffer, job i's float4 image (width*height, row 0 = bottom) at
 * out_offset; images must not overlap.  Each image equals aur_frame_render
 * of the same scene, camera and options bit for bit.  job_ms (optional,
 * host, njobs floats): device time from the start of the batch to the end of
 * the pass that completed each job (jobs are scheduled shortest first).
 * Synchronous. */
aur_status aur_batch_render(aur_context* ctx, uint32_t njobs, const aur_job* jobs, float* rgba,
                            float* job_ms);

/* ---- light sampling tables (cudaCreateLightsCDF, src/kernels/light.cu:34-60;
 * cudaCreateEmitters, src/kernels/emitter.cu:79-119; declared at
 * src/kernels/kernels.h:56-57).  Both run on the built scene (tree order)
 * with the current shaders and lights, write into caller-provided device
 * memory, and are synchronous. ---- */

/* Scene bounds: per-axis minimum and maximum over every vertex of the built
 * scene (computeBounds, src/kernels/rebuilder.cu:291-309, which reads only
 * the first two triangles in tree order; here all of them).  Host arrays of
 * 3 floats. */
aur_status aur_scene_bounds(aur_context* ctx, float bmin[3], float bmax[3]);

/* Light-power CDF: cdf[i] = (p_0 + ... + p_i) / (p_0 + ... + p_{n-1}) over the
 * n uploaded lights, with Light::power (src/kernels/lib/light.cuh:91-106):
 *   ambient      4 pi lum(color) intensity (1 - ambient)
 *   point        4 pi lum(color) intensity
 *   directional  lum(color) intensity pi R^2, R = |bmax - bmin| / 2
 *   area         lum(color) intensity area pi
 * lum = 0.2126 r + 0.7152 g + 0.0722 b; each power in float as the reference
 * writes it, the prefix sums in double, each cdf[i] rounded once to float
 * (so cdf[n-1] = 1 exactly).  A set whose powers sum to 0 gets the uniform
 * CDF (i + 1) / n (the reference divides by 0).  cdf: device, n floats.
 * *count = n. */
aur_status aur_create_lights_cdf(aur_context* ctx, float* cdf, uint32_t* count);

/* Emitter record, field for field the reference's Emitter (28 B,
 * src/util/emitter.h:10-22). */
typedef struct aur_emitter {
  float area;       /* triangle area, Primitive3::area (primitive.cuh:74-85) */
  float pdf;        /* 1 / area */
  float cdf;        /* normalised prefix sum of the emitters' shader emission */
  float power[3];   /* the shader's emission colour */
  uint32_t triangle; /* tree-order triangle index */
} aur_emitter;
/* Emissive triangles: every triangle (tree order) whose shader's scalar
 * emission is > 0, in triangle order, with cdf the prefix sum of those
 * emissions (not area-weighted, as the reference) divided by their total:
 * prefix sums in double, each cdf rounded once to float.  emitters: device,
 * capacity records; *count = the number of emitters (0: none, nothing
 * written).  A count above capacity fails with AUR_ERR_INVALID_ARG before
 * anything is written.  Shader ids past the shader count use shader 0. */
aur_status aur_create_emitters(aur_context* ctx, aur_emitter* emitters, uint64_t capacity, uint32_t* count);
/* aur_create_emitters without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * and harnesses that forbid hipMalloc).  Same arguments and the same table
 * bit for bit; its per-tile scratch is a fixed device array of the library
 * (2 MiB on each GPU it runs on, held until the process ends), which limits
 * it to scenes of at most 2^27 triangles (AUR_ERR_INVALID_ARG beyond).  Calls
 * on one GPU share that array: calls that overlap on one GPU (from several
 * threads or contexts) race. */
aur_status aur_create_emitters_noalloc(aur_context* ctx, aur_emitter* emitters, uint64_t capacity,
                                       uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* AURORA_AMD_H */
