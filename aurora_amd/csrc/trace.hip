// Standalone traversal launches: one lane per ray, closest hit or any hit
// (src/kernels/lib/intersect.cuh:36-152).  Each lane keeps its deferred
// ranges in an LDS column of `levels` float2 entries.
#include <hip/hip_runtime.h>

#include "context.h"
#include "traverse.h"

namespace aur {
namespace {

constexpr int kTraceBlock = 128;

template <bool kAny>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const NodeRec* __restrict__ nodes,
                                                       uint32_t n_nodes,
                                                       const aur_ray* __restrict__ rays, uint64_t n,
                                                       aur_hit* __restrict__ hits,
                                                       int32_t* __restrict__ ids) {
  extern __shared__ float2 stk_mem[];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* r = reinterpret_cast<const float4*>(rays + i);
  const float4 a = r[0], b = r[1];
  const TraceResult res = trace_ray<kAny>(nodes, n_nodes, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                          stk_mem + threadIdx.x, blockDim.x);
  if (kAny) {
    ids[i] = res.id;
  } else {
    reinterpret_cast<float4*>(hits)[i] =
        make_float4(__int_as_float(res.id), res.t, res.u, res.v);
  }
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_count(const NodeRec* __restrict__ nodes,
                                                             uint32_t n_nodes,
                                                             const aur_ray* __restrict__ rays,
                                                             uint64_t n, aur_hit* __restrict__ hits,
                                                             unsigned long long* __restrict__ sums) {
  extern __shared__ float2 stk_mem[];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  TraceCount c;
  if (i < n) {
    const float4* r = reinterpret_cast<const float4*>(rays + i);
    const float4 a = r[0], b = r[1];
    const TraceResult res = trace_ray<false, true>(nodes, n_nodes, mk(a.x, a.y, a.z),
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

}  // namespace

aur_status trace_closest_count(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits,
                               uint64_t* popped, uint64_t* passed) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (!rays || !hits) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  DevBuf<unsigned long long> sums;
  AUR_HIP(ctx, sums.reserve(2));
  k_zero2<<<1, 64, 0, ctx->stream>>>(sums.p);
  if (n) {
    const uint64_t nb = (n + kTraceBlock - 1) / kTraceBlock;
    if (nb > 0x7FFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: too many rays");
    const size_t lds = (size_t)ctx->levels * kTraceBlock * sizeof(float2);
    k_trace_count<<<(uint32_t)nb, kTraceBlock, lds, ctx->stream>>>(ctx->nodes.p, ctx->count / 2, rays,
                                                                    n, hits, sums.p);
    AUR_LAUNCH_CHECK(ctx, "k_trace_count");
  }
  unsigned long long h[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (popped) *popped = h[0];
  if (passed) *passed = h[1];
  return AUR_OK;
}

aur_status trace_closest(aur_context* ctx, uint64_t n, const aur_ray* rays, aur_hit* hits) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (n == 0) return AUR_OK;
  if (!rays || !hits) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  const uint64_t nb = (n + kTraceBlock - 1) / kTraceBlock;
  if (nb > 0x7FFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: too many rays");
  const size_t lds = (size_t)ctx->levels * kTraceBlock * sizeof(float2);
  k_trace<false><<<(uint32_t)nb, kTraceBlock, lds, ctx->stream>>>(ctx->nodes.p, ctx->count / 2, rays,
                                                                   n, hits, nullptr);
  AUR_LAUNCH_CHECK(ctx, "k_trace<closest>");
  return AUR_OK;
}

aur_status trace_any(aur_context* ctx, uint64_t n, const aur_ray* rays, int32_t* ids) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "trace: geometry not built");
  if (n == 0) return AUR_OK;
  if (!rays || !ids) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: null buffer");
  const uint64_t nb = (n + kTraceBlock - 1) / kTraceBlock;
  if (nb > 0x7FFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "trace: too many rays");
  const size_t lds = (size_t)ctx->levels * kTraceBlock * sizeof(float2);
  k_trace<true><<<(uint32_t)nb, kTraceBlock, lds, ctx->stream>>>(ctx->nodes.p, ctx->count / 2, rays,
                                                                  n, nullptr, ids);
  AUR_LAUNCH_CHECK(ctx, "k_trace<any>");
  return AUR_OK;
}

}  // namespace aur
