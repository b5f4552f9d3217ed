// Light sampling tables: scene bounds, the light-power CDF and the emitter
// table (computeBounds, src/kernels/rebuilder.cu:291-309; cudaCreateLightsCDF,
// src/kernels/light.cu:20-60; cudaCreateEmitters, src/kernels/emitter.cu:20-119).
//
// The emitter table is a stable compaction of the triangles whose shader
// emits, with a normalised prefix sum of their emissions.  It is HBM-bound
// byte work (2 B of shader id per triangle, 36 B of vertices and a 28 B
// record per emitter), so it is built from tiles of 2048 triangles: each of a
// tile's 256 threads owns 8 consecutive triangles (one 16-byte load of their
// shader ids), counts and sums its emitters, and one block-wide scan gives
// every thread its output position and the emission before it.
//   k_emit_tiles      per tile: emitter count and emission sum
//   k_emit_scan       one block: exclusive prefix over the tiles, totals
//   k_emit_write      per tile: records at their final positions, cdf
// Emission sums are kept in double and each cdf is rounded once to float,
// where the reference runs a float Thrust scan (its association
// unspecified) and divides by the last float value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "context.h"

static_assert(sizeof(aur_emitter) == 28, "aur_emitter layout (src/util/emitter.h:10-22)");

namespace aur {
namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kPerThread = 8;
constexpr uint32_t kTile = kScanThreads * kPerThread;

// Emitter count and emission sum of a tile, or a prefix of tiles.
struct TileSum {
  uint32_t count, pad;
  double weight;
};

struct Scan {
  TileSum excl, total;
};

// Block-wide exclusive scan of one (count, weight) per thread: a wave scan
// by shuffles, then the waves' totals in order.
__device__ Scan block_scan(uint32_t c, double w, TileSum* lds) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t ic = c;
  double iw = w;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t oc = __shfl_up(ic, d, 64);
    const double ow = __shfl_up(iw, d, 64);
    if (lane >= (uint32_t)d) {
      ic += oc;
      iw = ow + iw;
    }
  }
  uint32_t ec = __shfl_up(ic, 1, 64);
  double ew = __shfl_up(iw, 1, 64);
  if (lane == 0) {
    ec = 0;
    ew = 0.0;
  }
  if (lane == 63) lds[wave] = TileSum{ic, 0u, iw};
  __syncthreads();
  Scan r{TileSum{0u, 0u, 0.0}, TileSum{0u, 0u, 0.0}};
  for (uint32_t k = 0; k < kScanThreads / 64; ++k) {
    if (k < wave) {
      r.excl.count += lds[k].count;
      r.excl.weight += lds[k].weight;
    }
    r.total.count += lds[k].count;
    r.total.weight += lds[k].weight;
  }
  r.excl.count += ec;
  r.excl.weight += ew;
  return r;
}

__device__ inline uint32_t safe_shader(uint16_t id, uint32_t nsh) {
  const uint32_t s = id > 0 ? id - 1u : 0u;  // getSafeID, src/util/array.h:17-18
  return s < nsh ? s : 0u;
}

// The shader ids of this thread's 8 triangles (0 past the end).
__device__ inline void load_ids(const uint16_t* __restrict__ ids, uint32_t n, uint32_t first, uint16_t out[8]) {
  if (first + kPerThread <= n) {
    const uint4 q = *reinterpret_cast<const uint4*>(ids + first);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      out[2 * k] = (uint16_t)(w[k] & 0xFFFFu);
      out[2 * k + 1] = (uint16_t)(w[k] >> 16);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kPerThread; ++k) out[k] = first + k < n ? ids[first + k] : 0;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_emit_tiles(const uint16_t* __restrict__ ids, uint32_t n,
                                                             const aur_shader* __restrict__ shaders, uint32_t nsh,
                                                             TileSum* __restrict__ tiles) {
  __shared__ TileSum lds[kScanThreads / 64];
  const uint32_t first = blockIdx.x * kTile + threadIdx.x * kPerThread;
  uint16_t id[kPerThread];
  load_ids(ids, n, first, id);
  uint32_t c = 0;
  double w = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kPerThread; ++k) {
    const float e = first + k < n ? shaders[safe_shader(id[k], nsh)].emission : 0.0f;
    if (e > 0.0f) {
      ++c;
      w += (double)e;
    }
  }
  const Scan s = block_scan(c, w, lds);
  if (threadIdx.x == 0) tiles[blockIdx.x] = s.total;
}

// One block: tiles_pre[k] = sum of tiles [0, k); *total = sum of all.
__global__ __launch_bounds__(kScanThreads) void k_emit_scan(const TileSum* __restrict__ tiles, uint32_t ntiles,
                                                            TileSum* __restrict__ tiles_pre,
                                                            TileSum* __restrict__ total) {
  __shared__ TileSum lds[kScanThreads / 64];
  TileSum carry{0u, 0u, 0.0};
  for (uint32_t base = 0; base < ntiles; base += kScanThreads) {
    const uint32_t k = base + threadIdx.x;
    const TileSum t = k < ntiles ? tiles[k] : TileSum{0u, 0u, 0.0};
    const Scan s = block_scan(t.count, t.weight, lds);
    if (k < ntiles) tiles_pre[k] = TileSum{carry.count + s.excl.count, 0u, carry.weight + s.excl.weight};
    carry.count += s.total.count;
    carry.weight += s.total.weight;
    __syncthreads();  // lds is reused by the next chunk's scan
  }
  if (threadIdx.x == 0) *total = carry;
}

// Primitive3::area (src/kernels/lib/primitive.cuh:74-85) of the triangle at
// v (points layout), in the reference's operation order.
__device__ inline float triangle_area(const float* __restrict__ v) {
  const V3 v1 = mk(v[0], v[3], v[6]), v2 = mk(v[1], v[4], v[7]), v3 = mk(v[2], v[5], v[8]);
  const V3 va = v1 - v2, vb = v1 - v3, vc = v2 - v3;
  const float a = dot(va, va), b = dot(vb, vb), c = dot(vc, vc);
  return sqrtf((2 * a * b + 2 * b * c + 2 * c * a - a * a - b * b - c * c) / 16.0f);
}

__global__ __launch_bounds__(kScanThreads) void k_emit_write(const uint16_t* __restrict__ ids, uint32_t n,
                                                             const aur_shader* __restrict__ shaders, uint32_t nsh,
                                                             const float* __restrict__ verts,
                                                             const TileSum* __restrict__ tiles_pre,
                                                             const TileSum* __restrict__ total,
                                                             aur_emitter* __restrict__ out) {
  __shared__ TileSum lds[kScanThreads / 64];
  const uint32_t first = blockIdx.x * kTile + threadIdx.x * kPerThread;
  uint16_t id[kPerThread];
  load_ids(ids, n, first, id);
  float e[kPerThread];
  uint32_t c = 0;
  double w = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kPerThread; ++k) {
    e[k] = first + k < n ? shaders[safe_shader(id[k], nsh)].emission : 0.0f;
    if (e[k] > 0.0f) {
      ++c;
      w += (double)e[k];
    }
  }
  const Scan s = block_scan(c, w, lds);
  if (c == 0) return;
  const TileSum tp = tiles_pre[blockIdx.x];
  const double all = total->weight;
  uint32_t pos = tp.count + s.excl.count;
  double run = tp.weight + s.excl.weight;
  for (uint32_t k = 0; k < kPerThread; ++k) {
    if (!(e[k] > 0.0f)) continue;
    const uint32_t t = first + k;
    const aur_shader& sh = shaders[safe_shader(id[k], nsh)];
    run += (double)e[k];
    aur_emitter r;
    r.area = triangle_area(verts + (size_t)t * 9);
    r.pdf = 1.0f / r.area;
    r.cdf = (float)(run / all);
    r.power[0] = sh.emission_color[0];
    r.power[1] = sh.emission_color[1];
    r.power[2] = sh.emission_color[2];
    r.triangle = t;
    out[pos++] = r;
  }
}

// Scene bounds: per block min and max of each axis over a grid-stride range
// of triangles (points layout), then one block over the partials.
__global__ __launch_bounds__(256) void k_bounds_partial(const float* __restrict__ verts, uint32_t n,
                                                        float* __restrict__ partial) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const float* v = verts + (size_t)t * 9;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lo[a] = fminf(lo[a], v[3 * a + k]);
        hi[a] = fmaxf(hi[a], v[3 * a + k]);
      }
  }
  __shared__ float red[6][256];
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = lo[a];
    red[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (uint32_t d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + d]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + d]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bounds_final(const float* __restrict__ partial, uint32_t nb,
                                                      float* __restrict__ bounds) {
  __shared__ float red[6][256];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t b = threadIdx.x; b < nb; b += 256)
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = c < 3 ? fminf(v[c], partial[b * 6 + c]) : fmaxf(v[c], partial[b * 6 + c]);
#pragma unroll
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  for (uint32_t d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d)
#pragma unroll
      for (int c = 0; c < 6; ++c)
        red[c][threadIdx.x] = c < 3 ? fminf(red[c][threadIdx.x], red[c][threadIdx.x + d])
                                    : fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + d]);
    __syncthreads();
  }
  if (threadIdx.x < 6) bounds[threadIdx.x] = red[threadIdx.x][0];
}

// Light::power (src/kernels/lib/light.cuh:91-106) with luminosity
// (src/util/math.h:248-250) and getBoundingSphereRadius
// (src/kernels/lib/common.cuh:74-77), in the reference's operation order.
__device__ inline float light_power(const aur_light& L, V3 bmin, V3 bmax) {
  const float lum = 0.2126f * L.color[0] + 0.7152f * L.color[1] + 0.0722f * L.color[2];
  switch (L.type) {
    case AUR_LIGHT_AMBIENT:
      return 4.0f * kPi * lum * L.intensity * (1.0f - L.ambient);
    case AUR_LIGHT_POINT:
      return 4.0f * kPi * lum * L.intensity;
    case AUR_LIGHT_DIRECTIONAL: {
      const float R = 0.5f * length(bmin - bmax);
      return lum * L.intensity * kPi * R * R;
    }
    default:
      return lum * L.intensity * L.area * kPi;
  }
}

// One block: thread t owns lights [t*per, (t+1)*per); powers in float, sums
// in double, cdf[i] = prefix / total rounded once (uniform for a zero total).
__global__ __launch_bounds__(kScanThreads) void k_lights_cdf(const aur_light* __restrict__ lights, uint32_t n,
                                                             const float* __restrict__ bounds,
                                                             float* __restrict__ cdf) {
  __shared__ TileSum lds[kScanThreads / 64];
  const V3 bmin = mk(bounds[0], bounds[1], bounds[2]), bmax = mk(bounds[3], bounds[4], bounds[5]);
  const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
  const uint32_t b = threadIdx.x * per, e = min(n, b + per);
  double w = 0.0;
  for (uint32_t i = b; i < e; ++i) w += (double)light_power(lights[i], bmin, bmax);
  const Scan s = block_scan(0u, w, lds);
  const double all = s.total.weight;
  double run = s.excl.weight;
  for (uint32_t i = b; i < e; ++i) {
    run += (double)light_power(lights[i], bmin, bmax);
    cdf[i] = all == 0.0 ? (float)((double)(i + 1) / (double)n) : (float)(run / all);
  }
}

}  // namespace

// Scene bounds of the built scene into bounds[6] (device): min xyz, max xyz.
static aur_status scene_bounds_device(aur_context* ctx, float* bounds) {
  const uint32_t n = ctx->count;
  const uint32_t nb = std::max<uint32_t>(1u, std::min<uint32_t>(1024u, (n + 255) / 256));
  AUR_HIP(ctx, ctx->bounds_partial.reserve((size_t)nb * 6));
  k_bounds_partial<<<nb, 256, 0, ctx->stream>>>(ctx->vertices.p, n, ctx->bounds_partial.p);
  k_bounds_final<<<1, 256, 0, ctx->stream>>>(ctx->bounds_partial.p, nb, bounds);
  AUR_LAUNCH_CHECK(ctx, "scene bounds");
  return AUR_OK;
}

// The emitter table with tile scratch at tile_mem (2 * ntiles + 1 TileSum).
// Every scratch word is written before it is read, so the scratch needs no
// clearing between calls.
static aur_status emitters_run(aur_context* ctx, aur_emitter* out, uint64_t capacity, uint32_t* count,
                               void* tile_mem) {
  const uint32_t n = ctx->count;
  const uint32_t ntiles = (n + kTile - 1) / kTile;
  TileSum* tiles = static_cast<TileSum*>(tile_mem);
  TileSum* pre = tiles + ntiles;
  TileSum* tot = pre + ntiles;
  hipStream_t s = ctx->stream;
  k_emit_tiles<<<ntiles, kScanThreads, 0, s>>>(ctx->shader_ids.p, n, ctx->shaders.p, ctx->num_shaders, tiles);
  k_emit_scan<<<1, kScanThreads, 0, s>>>(tiles, ntiles, pre, tot);
  AUR_LAUNCH_CHECK(ctx, "emitter tiles");
  TileSum h{};
  AUR_HIP(ctx, hipMemcpyAsync(&h, tot, sizeof(h), hipMemcpyDeviceToHost, s));
  AUR_HIP(ctx, hipStreamSynchronize(s));
  *count = h.count;
  if (h.count == 0) return AUR_OK;
  if (h.count > capacity) return ctx->fail(AUR_ERR_INVALID_ARG, "emitters: more emitters than the capacity");
  k_emit_write<<<ntiles, kScanThreads, 0, s>>>(ctx->shader_ids.p, n, ctx->shaders.p, ctx->num_shaders,
                                               ctx->vertices.p, pre, tot, out);
  AUR_LAUNCH_CHECK(ctx, "k_emit_write");
  AUR_HIP(ctx, hipStreamSynchronize(s));
  return AUR_OK;
}

// Tile scratch of aur_create_emitters_noalloc: a fixed device array, one
// copy per GPU (2 MiB), for scenes of at most kNoAllocTiles tiles (2^27
// triangles).  Host code never passes `scratch` itself to a kernel: its
// device address on each GPU comes from hipGetSymbolAddress, taken once per
// device at context creation (scratch_setup) with that device current.
constexpr uint32_t kNoAllocTiles = 1u << 16;
static __device__ __align__(256) TileSum scratch[2 * kNoAllocTiles + 1];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};

hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_done[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  g_scratch_done[device] = true;
  return hipSuccess;
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

static aur_status tables_check(aur_context* ctx) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "light tables: geometry not built");
  return AUR_OK;
}

}  // namespace aur

extern "C" {

aur_status aur_scene_bounds(aur_context* ctx, float bmin[3], float bmax[3]) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!bmin || !bmax) return ctx->fail(AUR_ERR_INVALID_ARG, "bounds: null");
  aur_status st = aur::tables_check(ctx);
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, ctx->bounds.reserve(6));
  st = aur::scene_bounds_device(ctx, ctx->bounds.p);
  if (st != AUR_OK) return st;
  float h[6];
  AUR_HIP(ctx, hipMemcpyAsync(h, ctx->bounds.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < 3; ++a) {
    bmin[a] = h[a];
    bmax[a] = h[3 + a];
  }
  return AUR_OK;
}

aur_status aur_create_lights_cdf(aur_context* ctx, float* cdf, uint32_t* count) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!count) return ctx->fail(AUR_ERR_INVALID_ARG, "lights cdf: null count");
  aur_status st = aur::tables_check(ctx);
  if (st != AUR_OK) return st;
  const uint32_t n = ctx->num_lights;
  *count = n;
  if (n == 0) return AUR_OK;
  if (!cdf) return ctx->fail(AUR_ERR_INVALID_ARG, "lights cdf: null output");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  AUR_HIP(ctx, ctx->bounds.reserve(6));
  st = aur::scene_bounds_device(ctx, ctx->bounds.p);
  if (st != AUR_OK) return st;
  aur::k_lights_cdf<<<1, aur::kScanThreads, 0, ctx->stream>>>(ctx->lights.p, n, ctx->bounds.p, cdf);
  AUR_LAUNCH_CHECK(ctx, "k_lights_cdf");
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

aur_status aur_create_emitters(aur_context* ctx, aur_emitter* emitters, uint64_t capacity, uint32_t* count) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!count) return ctx->fail(AUR_ERR_INVALID_ARG, "emitters: null count");
  aur_status st = aur::tables_check(ctx);
  if (st != AUR_OK) return st;
  if (ctx->num_shaders == 0) return ctx->fail(AUR_ERR_STATE, "emitters: no shaders uploaded");
  if (capacity && !emitters) return ctx->fail(AUR_ERR_INVALID_ARG, "emitters: null output");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t ntiles = (ctx->count + aur::kTile - 1) / aur::kTile;
  AUR_HIP(ctx, ctx->emit_scratch.reserve((2 * (size_t)ntiles + 1) * sizeof(aur::TileSum) / 8));
  return aur::emitters_run(ctx, emitters, capacity, count, ctx->emit_scratch.p);
}

}  // extern "C"

extern "C" aur_status aur_create_emitters_noalloc(aur_context* ctx, aur_emitter* emitters, uint64_t capacity,
                                                  uint32_t* count) {
  if (!ctx) return AUR_ERR_INVALID_ARG;
  if (!count) return ctx->fail(AUR_ERR_INVALID_ARG, "emitters: null count");
  aur_status st = aur::tables_check(ctx);
  if (st != AUR_OK) return st;
  if (ctx->num_shaders == 0) return ctx->fail(AUR_ERR_STATE, "emitters: no shaders uploaded");
  if (capacity && !emitters) return ctx->fail(AUR_ERR_INVALID_ARG, "emitters: null output");
  const uint32_t ntiles = (ctx->count + aur::kTile - 1) / aur::kTile;
  if (ntiles > aur::kNoAllocTiles)
    return ctx->fail(AUR_ERR_INVALID_ARG, "emitters (no allocation): more than 2^27 triangles");
  AUR_HIP(ctx, hipSetDevice(ctx->device));
  void* mem = aur::scratch_ptr(ctx->device);
  if (!mem) return ctx->fail(AUR_ERR_DEVICE, "emitters (no allocation): device scratch address unavailable");
  return aur::emitters_run(ctx, emitters, capacity, count, mem);
}
