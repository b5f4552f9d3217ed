// Standalone traversal launches (src/kernels/lib/intersect.cuh:36-152).
//
// k_trace_pt: persistent waves (trace_rays below): each wave owns a
// contiguous range of rays and refills idle lanes from it, so lanes whose
// rays finish early do not idle until the slowest ray of the wave ends.
// Each lane keeps its deferred ranges in an LDS column of levels - 1 float2
// entries (or the first rows in LDS and the rest in global memory).
//
// k_trace_flat: one lane per ray, no refill (reference schedule; used for
// comparison and by the instrumented counting launch).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.h"
#include "traverse.h"

namespace aur {
namespace {

constexpr int kMaxBlock = 256;

__device__ __forceinline__ void store_result(uint64_t i, const Trav& s, aur_hit* __restrict__ hits,
                                             int32_t* __restrict__ ids) {
  if (ids)
    ids[i] = s.id;
  else
    reinterpret_cast<float4*>(hits)[i] = make_float4(__int_as_float(s.id), s.t, s.u, s.v);
}

// Persistent traversal of rays [0, n) with the refill of persistent_trace;
// the tree is wave-uniform, so its root slab is read once per wave and a
// refill waits on the ray loads only.  (Keeping the next 64 rays in
// registers and handing them out by lane permutes measured slower: the
// permutes' register reuse makes every refill wait for them.)
template <bool kAny, class Stk>
__device__ __forceinline__ void trace_rays(TreeRef tree, const aur_ray* __restrict__ rays, uint64_t n,
                                           uint32_t refill, const Stk& stk,
                                           aur_hit* __restrict__ hits, int32_t* __restrict__ ids) {
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t chunk = (n + waves - 1) / waves;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint64_t next = wave * chunk;
  const uint64_t end = next + chunk < n ? next + chunk : n;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const float2 root = root_slab(tree.nodes, tree.n_nodes, tree.rec_bytes);  // once per wave
  Trav s;
  bool active = false;
  uint64_t item = 0;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (nidle >= refill || nidle == 64) {
      if (next < end) {
        if (!active) {
          const uint64_t i = next + __popcll(idle & lt);
          if (i < end) {
            const float4* r = reinterpret_cast<const float4*>(rays + i);
            const float4 a = r[0], b = r[1];
            item = i;
            active = trav_init_root<false>(s, root, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                           nullptr);
            if (!active) store_result(i, s, hits, ids);
          }
        }
        next += nidle;
      } else if (nidle == 64) {
        break;
      }
    }
    if (active && trav_step<kAny, false>(s, tree.nodes, tree.n_nodes, tree.rec_bytes, stk, nullptr)) {
      store_result(item, s, hits, ids);
      active = false;
    }
  }
}

template <bool kAny, bool kSplit>
__global__ __launch_bounds__(kMaxBlock) void k_trace_pt(const NodeRec* __restrict__ nodes,
                                                        uint32_t n_nodes, uint32_t rec_bytes,
                                                        const aur_ray* __restrict__ rays, uint64_t n,
                                                        aur_hit* __restrict__ hits,
                                                        int32_t* __restrict__ ids, uint32_t refill,
                                                        float2* __restrict__ gstack, uint32_t klds,
                                                        uint32_t gstack_rows) {
  extern __shared__ float2 stk_mem[];
  if constexpr (kSplit) {
    // the first klds stack rows in LDS, deeper ones in global memory (the
    // launch keeps the buffer under 2^32 bytes)
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t bytes = (gstack_rows * lanes) * 8u;
    const SplitStack stk{(lds_vf2*)(stk_mem + threadIdx.x), blockDim.x,
                         (glb_vf2*)(gstack + gtid),
                         __builtin_amdgcn_make_buffer_rsrc(gstack, (short)0, (int)bytes, 0x00020000),
                         gtid * 8u, lanes, klds};
    trace_rays<kAny>(TreeRef{nodes, n_nodes, {}, rec_bytes}, rays, n, refill, stk, hits, ids);
  } else {
    trace_rays<kAny>(TreeRef{nodes, n_nodes, {}, rec_bytes}, rays, n, refill,
                     lds_stack(stk_mem + threadIdx.x, blockDim.x), hits, ids);
  }
}

template <bool kAny>
__global__ __launch_bounds__(kMaxBlock) void k_trace_flat(const NodeRec* __restrict__ nodes,
                                                          uint32_t n_nodes, uint32_t rec_bytes,
                                                          const aur_ray* __restrict__ rays,
                                                          uint64_t n, aur_hit* __restrict__ hits,
                                                          int32_t* __restrict__ ids) {
  extern __shared__ float2 stk_mem[];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* r = reinterpret_cast<const float4*>(rays + i);
  const float4 a = r[0], b = r[1];
  Trav s;
  if (trav_init<false>(s, nodes, n_nodes, rec_bytes, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                       nullptr)) {
    const LdsStack stk = lds_stack(stk_mem + threadIdx.x, blockDim.x);
    while (!trav_step<kAny, false>(s, nodes, n_nodes, rec_bytes, stk, nullptr)) {
    }
  }
  store_result(i, s, hits, ids);
}

__global__ __launch_bounds__(kMaxBlock) void k_trace_count(const NodeRec* __restrict__ nodes,
                                                           uint32_t n_nodes, uint32_t rec_bytes,
                                                           const aur_ray* __restrict__ rays,
                                                           uint64_t n, aur_hit* __restrict__ hits,
                                                           unsigned long long* __restrict__ sums) {
  extern __shared__ float2 stk_mem[];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  TraceCount c;
  if (i < n) {
    const float4* r = reinterpret_cast<const float4*>(rays + i);
    const float4 a = r[0], b = r[1];
    const TraceResult res = trace_ray<false, true>(nodes, n_nodes, rec_bytes, mk(a.x, a.y, a.z),
                                                   mk(b.x, b.y, b.z), a.w, stk_mem + threadIdx.x,
                                                   blockDim.x, &c);
    reinterpret_cast<float4*>(hits)[i] = make_float4(__int_as_float(res.id), res.t, res.u, res.v);
  }
  uint32_t p = c.popped, q = c.passed;
  for (int off = 32; off > 0; off >>= 1) {
    p += __shfl_xor(p, off);
    q += __shfl_xor(q, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(&sums[0], (unsigned long long)p);
    atomicAdd(&sums[1], (unsigned long long)q);
  }
}

__global__ void k_zero2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

// Self-check of the division shortcuts used by the traversal:
//   sums[0] += mismatches of recip_rn(x) against 1.0f / x over every float x
//              with |x| in [2^-60, 2^64] (bit patterns lo..hi, both signs);
//   sums[1] += mismatches of quot_fast(n, d, RN(1/d)) against n / d over
//              `samples` hashed pairs with 2^-60 <= |d| < 16 and n = 0 or
//              2^-96 <= |n| < 2^60, random signs and mantissas.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_division_check(uint32_t lo, uint32_t hi, uint64_t samples,
                                                        uint64_t seed,
                                                        unsigned long long* __restrict__ sums) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bad_r = 0, bad_q = 0;
  for (uint64_t b = lo + gtid; b <= hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    bad_r += __float_as_uint(recip_rn(x)) != __float_as_uint(1.0f / x);
    bad_r += __float_as_uint(recip_rn(-x)) != __float_as_uint(1.0f / -x);
  }
  for (uint64_t i = gtid; i < samples; i += stride) {
    const uint64_t h = mix64(seed ^ mix64(i));
    const uint32_t md = (uint32_t)h & 0x7FFFFFu, mn = (uint32_t)(h >> 23) & 0x7FFFFFu;
    const uint32_t ed = 127u - 60u + (uint32_t)((h >> 46) % 64u);   // 2^-60 .. 2^3
    const uint32_t en = 127u - 96u + (uint32_t)((h >> 52) % 156u);  // 2^-96 .. 2^59
    const uint32_t sd = (uint32_t)(h >> 62) & 1u, sn = (uint32_t)(h >> 63);
    const float d = __uint_as_float((sd << 31) | (ed << 23) | md);
    const float n = (i % 1024u == 7u) ? 0.0f : __uint_as_float((sn << 31) | (en << 23) | mn);
    const float y = 1.0f / d;
    bad_q += __float_as_uint(quot_fast(n, d, y)) != __float_as_uint(n / d);
  }
  for (int off = 32; off > 0; off >>= 1) {
    bad_r += __shfl_xor(bad_r, off);
    bad_q += __shfl_xor(bad_q, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    if (bad_r) atomicAdd(&sums[0], (unsigned long long)bad_r);
    if (bad_q) atomicAdd(&sums[1], (unsigned long long)bad_q);
  }
}

template <bool kAny>
aur_status launch_trace(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                        int32_t* ids) {
  const aur_trace_params& tp = ctx->trace_params;
  const uint32_t block = tp.block;
  const size_t lds = (size_t)ctx->levels * block * sizeof(float2);
  if (!tp.persistent) {
    const uint64_t nb = (n + block - 1) / block;
    if (nb > 0x7FFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: too many rays");
    k_trace_flat<kAny><<<(uint32_t)nb, block, lds, ctx->stream>>>(ctx->nodes.p, ctx->count / 2,
                                                                  ctx->rec_bytes, rays, n, hits, ids);
    AUR_LAUNCH_CHECK(ctx, "k_trace_flat");
    return AUR_OK;
  }
  // stack rows: at most levels - 1 deferred entries; the first `rows` in
  // LDS, the rest in global memory
  const uint32_t L = ctx->levels;
  const uint32_t maxrows = L > 1 ? L - 1 : 1;
  const uint32_t rows = (tp.stack_lds_rows == 0 || tp.stack_lds_rows >= maxrows) ? maxrows : tp.stack_lds_rows;
  const uint32_t klds = rows < maxrows ? rows : 0;
  const size_t lds_pt = (size_t)rows * block * sizeof(float2);
  // resident waves: occupancy-limited blocks per CU times CUs (or the request)
  int per_cu = 0;
  if (tp.refill < 1 || tp.refill > 64)
    return ctx->fail(AUR_ERR_INVALID_ARG, "trace: refill must be in 1..64");
  using Kern = void (*)(const NodeRec*, uint32_t, uint32_t, const aur_ray*, uint64_t, aur_hit*,
                        int32_t*, uint32_t, float2*, uint32_t, uint32_t);
  const Kern fn = klds != 0 ? k_trace_pt<kAny, true> : k_trace_pt<kAny, false>;
  AUR_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, (int)block, lds_pt));
  if (per_cu < 1) per_cu = 1;
  uint64_t blocks_total = (uint64_t)per_cu * ctx->num_cus;
  if (tp.waves_per_cu) blocks_total = std::max<uint64_t>(1, (uint64_t)tp.waves_per_cu * ctx->num_cus * 64 / block);
  const uint64_t max_waves = (n + 63) / 64;  // more waves than rays/64 would idle
  uint64_t nwaves = blocks_total * block / 64;
  if (nwaves > max_waves) nwaves = max_waves;
  const uint32_t grid = (uint32_t)((nwaves * 64 + block - 1) / block);
  float2* gstack = nullptr;
  if (klds) {
    const size_t entries = (size_t)(maxrows - rows) * grid * block;
    if (entries * sizeof(float2) >= (1ull << 31))
      return ctx->fail(AUR_ERR_INVALID_ARG, "trace: global stack rows exceed 2 GiB");
    AUR_HIP(ctx, ctx->gstack.reserve(entries));
    gstack = ctx->gstack.p;
  }
  hipLaunchKernelGGL(fn, dim3(grid), dim3(block), lds_pt, ctx->stream, ctx->nodes.p, ctx->count / 2,
                     ctx->rec_bytes, rays, n, hits, ids, tp.refill, gstack, klds, maxrows - rows);
  AUR_LAUNCH_CHECK(ctx, "k_trace_pt");
  return AUR_OK;
}

}  // namespace

aur_status trace_closest_count(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                               uint64_t* popped, uint64_t* passed) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (!rays || !hits) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  DevBuf<unsigned long long> sums;
  AUR_HIP(ctx, sums.reserve(2));
  k_zero2<<<1, 64, 0, ctx->stream>>>(sums.p);
  if (n) {
    const uint32_t block = 128;
    const uint64_t nb = (n + block - 1) / block;
    if (nb > 0x7FFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: too many rays");
    const size_t lds = (size_t)ctx->levels * block * sizeof(float2);
    k_trace_count<<<(uint32_t)nb, block, lds, ctx->stream>>>(ctx->nodes.p, ctx->count / 2,
                                                              ctx->rec_bytes, rays, n, hits, sums.p);
    AUR_LAUNCH_CHECK(ctx, "k_trace_count");
  }
  unsigned long long h[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (popped) *popped = h[0];
  if (passed) *passed = h[1];
  return AUR_OK;
}

aur_status division_check(aur_context* ctx, uint64_t samples, uint64_t seed, uint64_t* recip_checked,
                          uint64_t* recip_bad, uint64_t* quot_bad) {
  DevBuf<unsigned long long> sums;
  AUR_HIP(ctx, sums.reserve(2));
  k_zero2<<<1, 64, 0, ctx->stream>>>(sums.p);
  const uint32_t lo = 0x21800000u, hi = 0x5F800000u;  // 2^-60 .. 2^64
  k_division_check<<<4096, 256, 0, ctx->stream>>>(lo, hi, samples, seed, sums.p);
  AUR_LAUNCH_CHECK(ctx, "k_division_check");
  unsigned long long h[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (recip_checked) *recip_checked = 2 * ((uint64_t)(hi - lo) + 1);
  if (recip_bad) *recip_bad = h[0];
  if (quot_bad) *quot_bad = h[1];
  return AUR_OK;
}

aur_status trace_closest(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (n == 0) return AUR_OK;
  if (!rays || !hits) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  return launch_trace<false>(ctx, n, rays, hits, nullptr);
}

aur_status trace_any(aur_context* ctx, uint64_t n, const aur_ray* rays, int32_t* ids) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (n == 0) return AUR_OK;
  if (!rays || !ids) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  return launch_trace<true>(ctx, n, rays, nullptr, ids);
}

}  // namespace aur
