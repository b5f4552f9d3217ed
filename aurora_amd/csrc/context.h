// Internal state of an aur_context and the device-buffer helper.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/aurora_amd.h"
#include "common.h"

namespace aur {

// PBR mode carries a vertex's next-event slots in one 64-bit mask.
constexpr uint32_t kMaxSlots = 64;
// Compat mode queues every shadow ray of a hit (a round holds at most 2^26
// of them, so 4096 per hit still leaves 16,384 entries per round).
constexpr uint32_t kMaxCompatSlots = 4096;

// Next-event slots a light takes per path vertex in PBR mode: an area light
// one per sample, an ambient light two (its unshadowed ambient part and its
// shadowed positional part), point and directional lights one.
__host__ __device__ inline uint32_t light_slots(const aur_light& L) {
  if (L.type == AUR_LIGHT_AREA) return L.samples > 1 ? (uint32_t)L.samples : 1u;
  return L.type == AUR_LIGHT_AMBIENT ? 2u : 1u;
}

// Slots of a light set, summed in 64 bits and saturated at 2^32 - 1: any sum
// above kMaxSlots makes PBR renders refuse the set (compat mode counts its
// shadow rays per hit separately, up to kMaxCompatSlots).
inline uint32_t light_set_slots(uint32_t count, const aur_light* lights) {
  uint64_t sum = 0;
  for (uint32_t i = 0; i < count; ++i) sum += light_slots(lights[i]);
  return sum > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)sum;
}

// The device allocator in force (aur_set_alloc_hook); alloc == nullptr means
// hipMalloc / hipFree.
aur_alloc_hook current_alloc_hook();

// Owning device allocation, freed by the allocator that made it.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  aur_alloc_hook by{nullptr, nullptr, nullptr};
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(by, o.by);
  }
  void release() {
    if (p) {
      if (by.alloc) {
        (void)hipDeviceSynchronize();  // as hipFree does: no queued work still uses p
        by.free(p, by.user);
      } else {
        (void)hipFree(p);
      }
    }
    p = nullptr;
    n = 0;
  }
  // Ensures capacity for `count` elements; contents are not preserved.
  hipError_t reserve(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    if (count == 0) return hipSuccess;
    const aur_alloc_hook h = current_alloc_hook();
    void* q = nullptr;
    if (h.alloc) {
      q = h.alloc(count * sizeof(T), h.user);
      if (!q) return hipErrorOutOfMemory;
    } else {
      hipError_t e = hipMalloc(&q, count * sizeof(T));
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
      }
    }
    p = static_cast<T*>(q);
    n = count;
    by = h;
    return hipSuccess;
  }
};

// Default paths in flight per PBR wavefront pass (2^24 paths: about 2 GB of
// path state plus 56 B per next-event slot and path).
constexpr uint32_t kDefaultMaxPaths = 1u << 24;
constexpr uint32_t kDefaultRenderRefill = 16;
constexpr uint32_t kDefaultRenderStackRows = 12;
// Frame counters (u32): three 64-bit queue-count words of the wavefront
// passes, then two u64 ray counters (camera + extension, shadow) at
// kStatsOffset.
constexpr uint32_t kStatsOffset = 8;
constexpr uint32_t kCounterWords = 12;

// One render job of the PBR wavefront: a camera over one built scene.  A
// frame render is a single job; a batch is many.  Paths are numbered
// globally: the job owns [path_start, path_start + paths), sample-major, so
// global path g is region pixel (g - path_start) % npix of sample
// (g - path_start) / npix.  Region pixel l is frame pixel
// (ry0 + l / rw) * w + rx0 + l % rw (the whole frame: rx0 = ry0 = 0, rw = w).
struct JobDev {
  const NodeRec* nodes;
  const float* normals;  // tree order, 9 floats per triangle
  const uint16_t* shader_ids;
  const aur_shader* shaders;
  const aur_light* lights;
  float4* pixels;  // accumulation, npix float4
  const float* texcoords;  // tree order, 6 floats per triangle
  const TexDesc* textures;
  const uint8_t* texels;
  uint32_t num_textures;
  unsigned long long path_start;
  uint32_t n_nodes, num_shaders, num_lights, nslots;
  uint32_t w, h, npix, spp, max_depth, rr_start, seed, paths;
  uint32_t rx0, ry0, rw;
  float2 root;  // the tree's root slab (axis 0)
  uint32_t rec_bytes;  // distance between node records
  aur_camera cam;
};

// A built scene kept for batches: tree-order arrays as the frame render uses.
struct SceneStore {
  uint32_t count = 0, levels = 0, num_shaders = 0, num_lights = 0, nslots = 0;
  float2 root_slab = make_float2(0.0f, 0.0f);
  uint32_t rec_bytes = kNodeBytes;
  DevBuf<NodeRec> nodes;
  DevBuf<float> normals;
  DevBuf<uint16_t> shader_ids;
  DevBuf<aur_shader> shaders;
  DevBuf<aur_light> lights;
};

struct FrameState {
  uint32_t width = 0, height = 0, spp = 0;
  // render region, inclusive (Rect left/right/bottom/top, src/util/math.h)
  uint32_t left = 0, right = 0, bottom = 0, top = 0;
  DevBuf<float4> pixels;  // RGBA accumulation (framebuffer)
  // PBR wavefront state, indexed by path (pixel + pixels * sample-in-pass)
  DevBuf<float4> ray_o, ray_d, beta, radiance;
  DevBuf<aur_hit> hits;
  DevBuf<uint32_t> queue[2];
  DevBuf<uint32_t> counters;  // kCounterWords u32: queue counts, then ray counters
  DevBuf<float4> sh_o, sh_d, sh_c;  // shadow rays: origin+tmax, dir, contribution
  DevBuf<int32_t> sh_occ;
  DevBuf<uint32_t> sh_queue;  // slots holding a shadow ray, appended by shading
  DevBuf<float2> gstack;      // traversal stack rows beyond the LDS window
  DevBuf<float4> compat_pend; // compat mode: mirror rays waiting below a refraction ray
  // compat wavefront, per (pixel, sample) entry of a round: the stack's top
  // ray (origin + weight, direction + eta), colour + last hit, (depth, rays
  // waiting, random draws, frame pixel), final colour; per queue position:
  // emission + ray weight, albedo + flags; per (shadow slot, queue
  // position): contribution + cast flag
  DevBuf<float4> cw_o, cw_d, cw_c, cw_emw, cw_c1, cw_contrib, cw_fin;
  DevBuf<uint4> cw_i;
  DevBuf<uint32_t> sh_job;    // batches: job of each shadow-queue entry
  // adjoint replay: path-vertex records, per-block partial gradients,
  // running gradient sums and the largest |dLoss/dI|
  DevBuf<int32_t> rec_mat;
  DevBuf<float> rec_rr, rec_w;
  DevBuf<double> adj_partials, adj_total;
  DevBuf<uint32_t> adj_absmax;
  // batches: job table, job index per path of a pass, pass start
  DevBuf<JobDev> jobs;
  DevBuf<uint32_t> path_job;
  DevBuf<unsigned long long> pass_g0;
  size_t shadow_cap = 0;
  uint64_t rays_traced = 0;
  uint64_t rays_extension = 0, rays_shadow = 0;  // of the last render
};

}  // namespace aur

struct aur_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  uint32_t num_cus = 256;
  // persistent waves, 256-thread blocks, refill at 24 idle lanes, occupancy
  // limited residency: fastest shape measured at 1M triangles / 2^24 rays
  aur_trace_params trace_params{1, 256, 24, 0, 0};
  aur::DevBuf<float2> gstack;  // global rows of the persistent traversal stack
  float2 root_slab = make_float2(0.0f, 0.0f);  // host copy of the built tree's root slab
  aur_render_params render_params{aur::kDefaultMaxPaths, 0, aur::kDefaultRenderRefill,
                                  aur::kDefaultRenderStackRows};
  std::atomic<uint32_t> cancel{0};  // set by aur_cancel, consumed by the render it stops

  // uploaded geometry (input order)
  uint32_t count = 0;
  aur::DevBuf<float> in_vertices, in_normals, in_texcoords;
  aur::DevBuf<uint16_t> in_shader_ids;

  // tree-order geometry and traversal records (valid after the build)
  bool built = false;
  bool has_tb = false;
  uint32_t levels = 0;
  aur::DevBuf<float> vertices, normals, texcoords, tangents, bitangents;
  aur::DevBuf<uint16_t> shader_ids;
  aur::DevBuf<uint32_t> perm;
  aur::DevBuf<aur::NodeRec> nodes;
  uint32_t rec_bytes = aur::kNodeBytes;  // distance between node records
  float root_lo = 0.0f, root_hi = 0.0f;

  // materials and lights
  aur::DevBuf<aur_shader> shaders;
  uint32_t num_shaders = 0;
  aur::DevBuf<aur_light> lights;
  uint32_t num_lights = 0;
  uint32_t num_slots = 0;  // next-event slots per path vertex
  bool textured_shaders = false;  // a shader uses a texture map
  bool branching_shaders = false;  // a shader is both reflective and translucent (two compat child rays)
  bool child_shaders = false;      // a shader is reflective or translucent (compat child rays)
  uint32_t compat_slots = 0;       // compat shadow rays per hit: area-light samples, one per other light
  // RGBA8 textures (aur_textures_upload)
  aur::DevBuf<aur::TexDesc> textures;
  aur::DevBuf<uint8_t> texels;
  uint32_t num_textures = 0;

  aur::FrameState frame;

  // light sampling tables: scene bounds (min xyz, max xyz), per-block bounds
  // partials, and the emitter table's per-tile counts and emission sums
  aur::DevBuf<float> bounds, bounds_partial;
  aur::DevBuf<double> emit_scratch;

  // built scenes kept for batches (aur_scene_add)
  std::vector<std::unique_ptr<aur::SceneStore>> scenes;

  aur_status fail(aur_status code, const std::string& msg) {
    last_error = msg;
    return code;
  }
  // true once per aur_cancel request
  bool take_cancel() { return cancel.exchange(0u, std::memory_order_acq_rel) != 0u; }
  aur_status hip_fail(hipError_t e, const char* what) {
    last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? AUR_ERR_OOM : AUR_ERR_DEVICE;
  }
};

#define AUR_HIP(ctx, expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return (ctx)->hip_fail(_e, #expr); \
  } while (0)

// Keeps at most two passes (or samples) queued ahead of the GPU, so that a
// cancel request is seen within about two passes: wait(i) before queuing
// pass i, mark(i) after it.
struct PassThrottle {
  hipStream_t stream;
  hipEvent_t ev[2] = {nullptr, nullptr};
  explicit PassThrottle(hipStream_t s) : stream(s) {
    for (auto& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
  }
  ~PassThrottle() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  hipError_t wait(uint64_t i) { return (i >= 2 && ev[i & 1]) ? hipEventSynchronize(ev[i & 1]) : hipSuccess; }
  hipError_t mark(uint64_t i) { return ev[i & 1] ? hipEventRecord(ev[i & 1], stream) : hipSuccess; }
};

#define AUR_LAUNCH_CHECK(ctx, what)                          \
  do {                                                       \
    hipError_t _e = hipGetLastError();                       \
    if (_e != hipSuccess) return (ctx)->hip_fail(_e, what);  \
  } while (0)

// Entry points implemented across translation units.
namespace aur {
// Device address of the allocation-free entry points' scratch array
// (emitters.hip): taken once per device by scratch_setup, with that device
// current; scratch_ptr returns it (null before setup or on failure).
hipError_t scratch_setup(int device);
void* scratch_ptr(int device);

aur_status nmh_build(aur_context* ctx, aur_build_stats* stats);
aur_status generate_tb(aur_context* ctx);
aur_status trace_closest(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits);
aur_status trace_any(aur_context* ctx, uint64_t n, const aur_ray* rays, int32_t* ids);
aur_status division_check(aur_context* ctx, uint64_t samples, uint64_t seed, uint64_t* recip_checked,
                          uint64_t* recip_bad, uint64_t* quot_bad);
aur_status trace_closest_count(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                               uint64_t* popped, uint64_t* passed);
// Argument and state checks of frame_render; no work is queued.
aur_status frame_render_check(aur_context* ctx, const aur_render_opts* opts, uint32_t s0, uint32_t s1);
aur_status frame_render(aur_context* ctx, const aur_camera* cam, const aur_render_opts* opts,
                        uint32_t s0, uint32_t s1);
aur_status frame_adjoint(aur_context* ctx, const aur_camera* cam, const aur_render_opts* opts,
                         const float* d_image, float* g_albedo, float* g_light);
aur_status frame_filter(aur_context* ctx);

// PBR wavefront passes over the global path range [begin, end) of `jobs`
// (host array; one job renders with wave-uniform parameters, several share
// every pass).  cap: paths per pass; levels, nslots, max_depth: maxima over
// the jobs.  pass_end_ms (optional, host, one entry per pass): device time
// from the start to the end of each pass.
struct PbrRun {
  const JobDev* jobs = nullptr;
  uint32_t njobs = 0;
  unsigned long long begin = 0, end = 0;
  uint32_t cap = 0, levels = 0, nslots = 0, max_depth = 0;
  std::vector<float>* pass_end_ms = nullptr;
  bool clear_images = false;  // zero every job's image before the first pass (batches)
};
aur_status pbr_run(aur_context* ctx, const PbrRun& run, unsigned long long* stats);
}  // namespace aur
