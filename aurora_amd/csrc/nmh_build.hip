// NMH build on gfx950: replaces cudaRebuildNMH (src/kernels/rebuilder.cu:311-450).
//
// Result (bit-identical to the reference's canonical semantics): the tree
// order of the triangles.  Node k of the implicit heap owns tree positions
// 2k and 2k+1; its children are nodes 2k+1 and 2k+2.
//
// The reference sorts the whole array by (node label, min key) at every level
// (sortTriangles, rebuilder.cu:262-289) with two Thrust radix sorts, swaps the
// first maximum of the max key into slot +1 (rebuilder.cu:127-151), relabels
// nodes and emits child splits through a host round trip per level
// (rebuilder.cu:247-260).  Here:
//   * subtree sizes and offsets of the left-complete tree are closed forms
//     (desc_count), so split emission and all host synchronisation vanish;
//   * only the pending region is sorted; finished node pairs are not moved
//     again -- their final in-pair order, which the reference obtains by
//     re-sorting them at every later level, is computed once as a
//     lexicographic comparison of the later levels' keys (newest first; only
//     the last three levels can differ since axes cycle with period 3);
//   * while node groups are larger than one workgroup's LDS tile, a level is a
//     segmented stable LSD radix sort (8-bit digits, tile histograms,
//     per-group digit offsets, wave-ballot stable ranking) followed by a
//     tiled first-maximum reduction and an emit pass that writes finished
//     pairs and compacts the remainders for the next level.  The digits are
//     those of key - (the group's smallest key), so a group whose keys span
//     fewer than 2^24 values needs 3 passes (2^16: 2, ...): the emit pass
//     records each tile's key range, the next level's first group scan
//     reduces them per group and hands (smallest key, passes) to every tile
//     and chunk of the group;
//   * once every group fits (<= kSubT elements) one workgroup takes a whole
//     subtree and runs all remaining levels in LDS with a bitonic sort on
//     (group, key, position) composite keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "context.h"

namespace aur {
namespace {

constexpr uint32_t kTile = 2048;  // radix / emit tile (elements)
constexpr uint32_t kChunk = 32;   // tiles per chunk of the radix offset scan
constexpr uint32_t kSubT = 2048;  // max elements of an LDS-resident subtree
constexpr uint64_t kCompactNodeArrayMax = 128ull << 20;  // bytes
constexpr int kThreads = 256;

struct TileDesc {
  uint32_t start, count, child, group;
};
struct ChildDesc {
  uint32_t off, size, tb, te;
};
struct GroupDesc {
  uint32_t off, size, tb, te, cb, ce;  // tiles [tb, te), offset chunks [cb, ce)
};
struct SubGroup {
  uint32_t off, size, first, count;
};

// Triangle bounds on one axis, (min, max), stored [axis][triangle].
// kx, ky: the min keys of x and y on the axis of the last level (L - 1),
// loaded by the caller; level <= L - 2.
// orig (optional): triangle id of each renumbered index x, y (the subtree
// phase works on positions, renumbered by the last k_emit).
__device__ inline void write_pair_k(uint32_t h, uint32_t x, uint32_t y, float kx, float ky,
                                    int level, int L, const float2* __restrict__ bb, uint32_t N,
                                    uint32_t* __restrict__ P, const uint32_t* __restrict__ orig = nullptr) {
  bool y_first = ky < kx;
  if (!y_first && !(kx < ky)) {
    const int lo = max(level + 1, L - 3);
    for (int lev = L - 2; lev >= lo; --lev) {
      const float2* km = bb + (size_t)(lev % 3) * N;
      const float ax = km[x].x, ay = km[y].x;
      if (ay < ax) {
        y_first = true;
        break;
      }
      if (ax < ay) break;
    }
  }
  const uint32_t a = y_first ? y : x, b = y_first ? x : y;
  P[2 * (size_t)h] = orig ? orig[a] : a;
  P[2 * (size_t)h + 1] = orig ? orig[b] : b;
}

__device__ inline void write_pair(uint32_t h, uint32_t x, uint32_t y, int level, int L,
                                  const float2* __restrict__ bb, uint32_t N,
                                  uint32_t* __restrict__ P) {
  const float2* km = bb + (size_t)((L - 1) % 3) * N;
  write_pair_k(h, x, y, km[x].x, km[y].x, level, L, bb, N, P);
}

// computeTriangleBounds (rebuilder.cu:72-91), (min, max) pairs [axis][triangle]:
// one 8-byte load gives both keys of a triangle on an axis.
__global__ void k_bounds(const float* __restrict__ v, uint32_t N, float2* __restrict__ bb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float* p = v + (size_t)i * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    bb[(size_t)a * N + i] = make_float2(fminf(fminf(p[3 * a], p[3 * a + 1]), p[3 * a + 2]),
                                        fmaxf(fmaxf(p[3 * a], p[3 * a + 1]), p[3 * a + 2]));
}

// Identity order with the axis-0 max keys (the sort's payload: triangle id,
// max key), and (when K is given) the first level's axis-0 sort keys.
__global__ void k_iota(uint2* __restrict__ V, uint32_t N, const float2* __restrict__ bb0,
                       uint32_t* __restrict__ K) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float2 b = bb0[i];
  V[i] = make_uint2(i, fkey(b.y));
  if (K) K[i] = fkey(b.x);
}

// Indices read back from device arrays are clamped to [0, N) before they
// address another array: a defect then shows up as a wrong tree in the parity
// tests, never as an out-of-bounds access.
__device__ __forceinline__ uint32_t clampi(uint32_t i, uint32_t N) { return i < N ? i : N - 1; }

// Radix passes a group needs: its keys minus the smallest, k - r.x, span
// [0, r.y - r.x]; digits above the span's top byte are zero in every key, and
// a stable sort by a digit that is equal for all keys keeps the order.  Pass
// 0 runs for every group (its digit, (k - r.x) & 255, is a rotation of the
// key's low byte, and the group's range is known only after the pass's
// histograms).  Pass p of a level moves the groups with group_passes > p,
// from buffer p & 1 to the other, so a group's sorted data ends in buffer
// group_passes & 1.  Kernels read a tile's (smallest key, passes) as
// TP[tile] and a chunk's passes as CP[chunk] (by block index, so the load
// does not wait for the tile's descriptor).
__device__ __forceinline__ uint32_t group_passes(uint2 r) {
  const uint32_t d = r.y - r.x;
  return d < 0x10000u ? (d < 0x100u ? 1u : 2u) : (d < 0x1000000u ? 3u : 4u);
}

__global__ void k_fill0(uint32_t* __restrict__ p, uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) p[i] = 0;
}

// Per-tile digit histogram H[tile][digit], also added into its chunk's digit
// counts CS[chunk][digit] (integer atomics: the sums are exact whatever the
// order).  CS starts at zero: zeroed once per build and again by
// k_radix_group_scan after it reads a pass's counts.  Pass 0 counts the
// key's low byte (indexed so; k_radix_group_scan orders the digits).
__global__ __launch_bounds__(kThreads) void k_radix_hist(const TileDesc* __restrict__ tiles,
                                                         const GroupDesc* __restrict__ groups,
                                                         const uint2* __restrict__ TP,
                                                         const uint32_t* __restrict__ K,
                                                         uint32_t pass,
                                                         uint32_t* __restrict__ H,
                                                         uint32_t* __restrict__ CS) {
  __shared__ uint32_t h[256];
  const TileDesc t = tiles[blockIdx.x];
  uint32_t kmin = 0;
  if (pass > 0) {
    const uint2 tp = TP[blockIdx.x];
    if (tp.y <= pass) return;  // uniform per block
    kmin = tp.x;
  }
  const uint32_t shift = 8u * pass;
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < t.count; i += kThreads)
    atomicAdd(&h[((K[t.start + i] - kmin) >> shift) & 255u], 1u);
  __syncthreads();
  const uint32_t v = h[threadIdx.x];
  H[(size_t)blockIdx.x * 256 + threadIdx.x] = v;
  const GroupDesc g = groups[t.group];
  const uint32_t chunk = g.cb + (blockIdx.x - g.tb) / kChunk;
  if (v) atomicAdd(&CS[(size_t)chunk * 256 + threadIdx.x], v);
}

// Per group: exclusive digit offsets of the group (digit-major), and the base
// B[chunk][digit] of every chunk of the group (the counts of the group's
// earlier chunks added); then the chunk counts are zeroed for the next pass.
// (A group that skips the pass has no counts: they stay zero.)
// Pass 0 first sets the group's key range: from the previous level's emit
// tiles of the group's parent node (prev_children[group], tile ranges TR),
// or the full range on the first level (prev_children null); GP[group],
// TP[tile] and CP[chunk] of the group take (smallest key, passes).  Its
// digits are the keys' low bytes, scanned in the order of
// (low byte - smallest key) & 255.
__global__ __launch_bounds__(kThreads) void k_radix_group_scan(const GroupDesc* __restrict__ groups,
                                                               uint2* __restrict__ GP,
                                                               uint2* __restrict__ TP,
                                                               uint32_t* __restrict__ CP,
                                                               const ChildDesc* __restrict__ prev_children,
                                                               const uint2* __restrict__ TR,
                                                               uint32_t pass,
                                                               uint32_t* __restrict__ S,
                                                               uint32_t* __restrict__ B) {
  __shared__ uint32_t s[256];
  __shared__ uint32_t rmin[kThreads / 64], rmax[kThreads / 64];
  const uint32_t d = threadIdx.x;
  uint32_t rot = 0;  // scan position of digit d: (d - rot) & 255
  const GroupDesc g = groups[blockIdx.x];
  const bool reduce = pass == 0 && prev_children != nullptr;
  const ChildDesc pc = reduce ? prev_children[blockIdx.x] : ChildDesc{0u, 0u, 0u, 0u};
  if (pass > 0 && GP[blockIdx.x].y <= pass) return;  // uniform per block
  // the group's digit counts first: their loads then overlap the range's
  uint32_t total = 0;
#pragma unroll 8
  for (uint32_t c = g.cb; c < g.ce; ++c) total += S[(size_t)c * 256 + d];
  if (pass == 0) {
    uint2 r = make_uint2(0u, 0xFFFFFFFFu);
    if (reduce) {
      uint32_t mn = 0xFFFFFFFFu, mx = 0u;
      for (uint32_t t = pc.tb + d; t < pc.te; t += kThreads) {
        const uint2 tr = TR[t];
        mn = min(mn, tr.x);
        mx = max(mx, tr.y);
      }
      for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, off));
        mx = max(mx, (uint32_t)__shfl_xor(mx, off));
      }
      if ((d & 63u) == 0) {
        rmin[d >> 6] = mn;
        rmax[d >> 6] = mx;
      }
      __syncthreads();
      r = make_uint2(rmin[0], rmax[0]);
#pragma unroll
      for (int w = 1; w < kThreads / 64; ++w) r = make_uint2(min(r.x, rmin[w]), max(r.y, rmax[w]));
    }
    const uint2 gp = make_uint2(r.x, group_passes(r));
    if (d == 0) GP[blockIdx.x] = gp;
    for (uint32_t t = g.tb + d; t < g.te; t += kThreads) TP[t] = gp;
    for (uint32_t c = g.cb + d; c < g.ce; c += kThreads) CP[c] = gp.y;
    rot = r.x & 255u;
  }
  const uint32_t sp = (d - rot) & 255u;
  s[sp] = total;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t v = d >= off ? s[d - off] : 0u;
    __syncthreads();
    s[d] += v;
    __syncthreads();
  }
  uint32_t run = g.off + s[sp] - total;
#pragma unroll 8
  for (uint32_t c = g.cb; c < g.ce; ++c) {
    B[(size_t)c * 256 + d] = run;
    run += S[(size_t)c * 256 + d];
    S[(size_t)c * 256 + d] = 0u;
  }
}

// Output offsets of every tile, O[tile][digit]: the chunk's base plus the
// counts of the chunk's earlier tiles (one block per chunk; the chunk's
// tiles are found from the tile's group).
__global__ __launch_bounds__(kThreads) void k_radix_chunk_offsets(const uint32_t* __restrict__ chunk_tile,
                                                                  const uint32_t* __restrict__ CP,
                                                                  uint32_t pass,
                                                                  const uint32_t* __restrict__ H,
                                                                  const uint32_t* __restrict__ B,
                                                                  uint32_t* __restrict__ O) {
  const uint32_t tb = chunk_tile[blockIdx.x], te = chunk_tile[blockIdx.x + 1];
  if (pass > 0 && CP[blockIdx.x] <= pass) return;  // uniform per block
  const uint32_t d = threadIdx.x;
  uint32_t run = B[(size_t)blockIdx.x * 256 + d];
#pragma unroll 8
  for (uint32_t t = tb; t < te; ++t) {
    O[(size_t)t * 256 + d] = run;
    run += H[(size_t)t * 256 + d];
  }
}

// Inclusive scan of one value per thread over the 256 threads (4 waves):
// shuffles within each wave, then the wave totals through LDS.
__device__ __forceinline__ uint32_t block_incl_scan256(uint32_t v, uint32_t* __restrict__ wsum) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off);
    if (lane >= (uint32_t)off) v += u;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  for (uint32_t w = 0; w < wave; ++w) v += wsum[w];
  return v;
}

// Stable scatter of one 2048-element tile, staged through LDS.  Ranking:
// wave w owns the tile's elements [512w, 512w + 512) in rounds of 64; lanes
// with the same digit are found with 8 ballots, and the wave keeps its
// running count per digit in its own LDS row (read, then written by the
// digit's first lane, in the wave's program order: no workgroup barrier per
// round).  Each element (key and 8-byte payload) then lands in LDS at its
// digit-sorted tile position, and the tile is written out in that order: a
// digit's run is one contiguous store at the tile's output offset for the
// digit, O[tile][digit].  (Ranking in rounds of 256 elements across the
// workgroup, with four barriers per round, took 20-35% longer; staging only
// a 16-bit tile index and re-reading the payload at write-out measured 6%
// slower despite the smaller LDS footprint.)
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const TileDesc* __restrict__ tiles, const uint2* __restrict__ TP,
    const uint32_t* __restrict__ Kin, const uint2* __restrict__ Vin, uint32_t pass,
    const uint32_t* __restrict__ O, uint32_t m, uint32_t* __restrict__ Kout,
    uint2* __restrict__ Vout) {
  constexpr uint32_t kWaves = kThreads / 64;
  constexpr uint32_t kRounds = kTile / kThreads;  // rounds of 64 per wave
  __shared__ uint32_t sk[kTile];
  __shared__ uint2 sv[kTile];
  __shared__ uint32_t wc[kWaves][256];  // per-wave running digit counts, then tile positions
  __shared__ uint32_t delta[256];       // output position - tile position, per digit
  __shared__ uint32_t wsum[kWaves];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const TileDesc t = tiles[blockIdx.x];
  uint32_t kmin = 0;  // pass 0: the low byte itself (the offsets O follow the group's digit order)
  if (pass > 0) {
    const uint2 tp = TP[blockIdx.x];
    if (tp.y <= pass) return;  // uniform per block
    kmin = tp.x;
  }
  const uint32_t shift = 8u * pass;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) wc[wave][q * 64 + lane] = 0;
  const uint32_t w0 = wave * (kRounds * 64);
  uint32_t k[kRounds], loc[kRounds];
  uint2 v[kRounds];
#pragma unroll
  for (uint32_t r = 0; r < kRounds; ++r) {
    const uint32_t i = w0 + r * 64 + lane;
    const bool valid = i < t.count;
    k[r] = valid ? Kin[t.start + i] : 0u;
    v[r] = valid ? Vin[t.start + i] : make_uint2(0u, 0u);
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (uint32_t r = 0; r < kRounds; ++r) {
    const bool valid = w0 + r * 64 + lane < t.count;
    const uint32_t dg = ((k[r] - kmin) >> shift) & 255u;
    uint64_t mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      mask &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(mask & lt_mask);
    const uint32_t off = wc[wave][dg];
    if (valid && rank == 0) wc[wave][dg] = off + (uint32_t)__popcll(mask);
    loc[r] = off + rank;
  }
  __syncthreads();
  // thread tid = digit d: the tile's count, its first tile position, and
  // the first tile position of each wave's elements with digit d
  uint32_t c[kWaves], cnt = 0;
#pragma unroll
  for (uint32_t w = 0; w < kWaves; ++w) {
    c[w] = wc[w][tid];
    cnt += c[w];
  }
  const uint32_t first = block_incl_scan256(cnt, wsum) - cnt;
  {
    uint32_t run = first;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
      wc[w][tid] = run;
      run += c[w];
    }
  }
  delta[tid] = O[(size_t)blockIdx.x * 256 + tid] - first;
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kRounds; ++r) {
    if (w0 + r * 64 + lane < t.count) {
      const uint32_t pos = wc[wave][((k[r] - kmin) >> shift) & 255u] + loc[r];
      if (pos < kTile) {
        sk[pos] = k[r];
        sv[pos] = v[r];
      }
    }
  }
  __syncthreads();
  const uint32_t n_out = min(t.count, kTile);
  for (uint32_t j = tid; j < n_out; j += kThreads) {
    const uint32_t key = sk[j];
    const uint32_t pos = delta[((key - kmin) >> shift) & 255u] + j;
    if (pos < m) {
      Kout[pos] = key;
      Vout[pos] = sv[j];
    }
  }
}

// First maximum of the max key over child positions [off+1, off+size)
// restricted to this tile: (key, position), key 0 = none.  The group's sorted
// data is in V0 or V1 (group_passes).
__global__ __launch_bounds__(kThreads) void k_argmax_partial(const TileDesc* __restrict__ tiles,
                                                             const ChildDesc* __restrict__ children,
                                                             const uint2* __restrict__ TP,
                                                             const uint2* __restrict__ V0,
                                                             const uint2* __restrict__ V1,
                                                             uint2* __restrict__ partial) {
  __shared__ uint32_t sk[kThreads / 64], sp[kThreads / 64];
  const TileDesc t = tiles[blockIdx.x];
  const ChildDesc c = children[t.child];
  const uint2* __restrict__ V = (TP[blockIdx.x].y & 1u) ? V1 : V0;
  const uint32_t lo = max(t.start, c.off + 1), hi = t.start + t.count;
  uint32_t bk = 0, bp = 0xFFFFFFFFu;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kThreads) {
    const uint32_t k = V[i].y;  // the element's max key, carried by the sort
    if (k > bk) {
      bk = k;
      bp = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t ok = __shfl_xor(bk, off), op = __shfl_xor(bp, off);
    if (ok > bk || (ok == bk && op < bp)) {
      bk = ok;
      bp = op;
    }
  }
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sk[wave] = bk;
    sp[wave] = bp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w)
      if (sk[w] > bk || (sk[w] == bk && sp[w] < bp)) {
        bk = sk[w];
        bp = sp[w];
      }
    partial[blockIdx.x] = make_uint2(bk, bp);
  }
}

// One wave per child: (max key, lowest position) over the child's tiles.
__global__ __launch_bounds__(kThreads) void k_argmax_combine(const ChildDesc* __restrict__ children,
                                                             uint32_t nchildren,
                                                             const uint2* __restrict__ partial,
                                                             uint32_t* __restrict__ argpos) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t c = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (c >= nchildren) return;  // uniform per wave
  const ChildDesc d = children[c];
  uint32_t bk = 0, bp = 0xFFFFFFFFu;
  for (uint32_t t = d.tb + lane; t < d.te; t += 64) {
    const uint2 p = partial[t];
    if (p.x > bk || (p.x == bk && p.y < bp)) {
      bk = p.x;
      bp = p.y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t ok = __shfl_xor(bk, off), op = __shfl_xor(bp, off);
    if (ok > bk || (ok == bk && op < bp)) {
      bk = ok;
      bp = op;
    }
  }
  if (lane == 0) argpos[c] = bp;
}

// Writes finished pairs (tree positions 2h, 2h+1) and compacts remainders
// (read from V0 or V1 as the group's sort left them, written to Vnext); with
// Knext, also the next level's sort keys and max keys of the compacted
// remainders (one 8-byte load of their bounds on the next axis), and the
// tile's (smallest, largest) key in TR[tile] ((~0, 0) if it has none).  A
// tile's remainders all go to the next level's group of the tile's node.
// With orig (the last level before the subtree phase),
// the remainders are renumbered instead: position dst takes the place of
// triangle id val, orig[dst] = val, and its bounds on all three axes go to
// bbR[axis][dst] (mR per axis), so that every subtree group reads one
// contiguous window of bounds.
__global__ __launch_bounds__(kThreads) void k_emit(const TileDesc* __restrict__ tiles,
                                                   const ChildDesc* __restrict__ children,
                                                   const uint32_t* __restrict__ argpos,
                                                   const uint2* __restrict__ TP,
                                                   const uint2* __restrict__ V0,
                                                   const uint2* __restrict__ V1,
                                                   uint2* __restrict__ Vnext, uint32_t first,
                                                   int level, int L, const float2* __restrict__ bb,
                                                   uint32_t N, uint32_t* __restrict__ P,
                                                   uint32_t* __restrict__ Knext,
                                                   uint2* __restrict__ TR,
                                                   float2* __restrict__ bbR, uint32_t* __restrict__ orig,
                                                   uint32_t mR) {
  const TileDesc t = tiles[blockIdx.x];
  const ChildDesc c = children[t.child];
  const uint2* __restrict__ V = (TP[blockIdx.x].y & 1u) ? V1 : V0;
  const float2* bnext = bb + (size_t)((level + 1) % 3) * N;
  uint32_t am = argpos[t.child];
  if (am <= c.off || am >= c.off + c.size) am = c.off + 1;  // cannot happen for sizes >= 2
  uint32_t kmin = 0xFFFFFFFFu, kmax = 0u;
  for (uint32_t r = threadIdx.x; r < t.count; r += kThreads) {
    const uint32_t i = t.start + r;
    if (i == c.off) {
      write_pair(first + t.child, clampi(V[c.off].x, N), clampi(V[am].x, N), level, L, bb, N, P);
    } else if (i >= c.off + 2) {
      const uint32_t val = clampi((i == am) ? V[c.off + 1].x : V[i].x, N);
      const uint32_t dst = i - 2 * (t.child + 1);
      if (Knext) {
        const float2 b = bnext[val];
        const uint32_t key = fkey(b.x);
        Knext[dst] = key;
        Vnext[dst] = make_uint2(val, fkey(b.y));
        kmin = min(kmin, key);
        kmax = max(kmax, key);
      } else if (orig) {
        const float2 b0 = bb[val], b1 = bb[(size_t)N + val], b2 = bb[2 * (size_t)N + val];
        if (dst < mR) {
          orig[dst] = val;
          bbR[dst] = b0;
          bbR[(size_t)mR + dst] = b1;
          bbR[2 * (size_t)mR + dst] = b2;
        }
        Vnext[dst] = make_uint2(dst, 0u);
      } else {
        Vnext[dst] = make_uint2(val, 0u);
      }
    }
  }
  if (Knext) {  // uniform: a kernel argument
    __shared__ uint32_t rmin[kThreads / 64], rmax[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
      kmin = min(kmin, (uint32_t)__shfl_xor(kmin, off));
      kmax = max(kmax, (uint32_t)__shfl_xor(kmax, off));
    }
    if ((threadIdx.x & 63u) == 0) {
      rmin[threadIdx.x >> 6] = kmin;
      rmax[threadIdx.x >> 6] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int w = 1; w < kThreads / 64; ++w) {
        kmin = min(kmin, rmin[w]);
        kmax = max(kmax, rmax[w]);
      }
      TR[blockIdx.x] = make_uint2(kmin, kmax);
    }
  }
}

// One workgroup per group: all remaining levels of the group's subtree(s).
// With orig, the pending region is renumbered (by the last k_emit): element values
// are positions, bb is bbR with stride N = m, and orig maps positions back to
// triangle ids where the tree order is written.
// 37 KB of LDS and at most 128 VGPRs, so four groups reside per CU.
__global__ __launch_bounds__(kThreads, 4) void k_subtree(const SubGroup* __restrict__ groups,
                                                         const uint2* __restrict__ V,
                                                         const float2* __restrict__ bb, uint32_t N,
                                                         uint32_t n, int L, int level0,
                                                         uint32_t* __restrict__ P,
                                                         const uint32_t* __restrict__ orig) {
  __shared__ uint32_t E[kSubT];
  __shared__ uint16_t G[kSubT];
  __shared__ unsigned long long K[kSubT];
  __shared__ uint32_t nodeOff[kSubT / 2 + 1];
  __shared__ uint32_t nodeArg[kSubT / 2];
  __shared__ uint32_t gmax;  // largest sort group of the level
  // once the sorted order is read out, the sort buffer holds the sorted ids
  // (E2) and their max keys (MK)
  uint32_t* E2 = reinterpret_cast<uint32_t*>(K);
  uint32_t* MK = E2 + kSubT;

  const SubGroup g = groups[blockIdx.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t m = g.size;
  for (uint32_t i = tid; i < m; i += kThreads) {
    E[i] = orig ? g.off + i : clampi(V[g.off + i].x, N);
    G[i] = 0;
  }
  uint32_t f = g.first, c = g.count;
  __syncthreads();
  for (int level = level0; level < L && m > 0; ++level) {
    const int a = level % 3;
    const float2* ba = bb + (size_t)a * N;  // (min, max) on this level's axis
    // every thread's key loads are issued together, before the offsets
    constexpr int kPer = kSubT / kThreads;  // elements per thread
    uint32_t kv[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      kv[r] = i < m ? fkey(ba[E[i]].x) : 0u;
    }
    for (uint32_t j = tid; j <= c; j += kThreads) nodeOff[j] = 2u * (uint32_t)desc_count(f, f + j, n);
    if (tid == 0) gmax = 0;
    __syncthreads();
    // Sort groups are the remainders of one node of the previous level: the
    // ranges of this level's node pairs (2q, 2q+1), contiguous and in order,
    // with G = q.  When every group fits a power-of-two slot of S entries and
    // the slots fit the sort buffer, each group is sorted in its own slot
    // (log2(S) bitonic phases instead of log2(M)); otherwise the whole region
    // is sorted on (group, key, position).
    const uint32_t ng = (c + 1) / 2;
    for (uint32_t q = tid; q < ng; q += kThreads)
      atomicMax(&gmax, nodeOff[min(2 * q + 2, c)] - nodeOff[2 * q]);
    uint32_t M = 1;
    while (M < m) M <<= 1;
    __syncthreads();
    uint32_t S = 2;
    while (S < gmax) S <<= 1;
    const bool seg = (uint64_t)ng * S <= kSubT;
    const uint32_t span = seg ? ng * S : M;  // sorted entries
    const uint32_t top = seg ? S : M;        // size of one sorted run
    for (uint32_t q = tid; q < span; q += kThreads) K[q] = ~0ull;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      if (i >= m) break;
      const unsigned long long key = kv[r];
      if (seg) {
        const uint32_t q = G[i], p = i - nodeOff[2 * q];
        K[q * S + p] = (key << 11) | p;
      } else {
        K[i] = ((unsigned long long)G[i] << 43) | (key << 11) | i;
      }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= top; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t t = tid; t < span / 2; t += kThreads) {
          const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));  // j is a power of two
          const uint32_t ixj = i + j;
          const bool up = (i & k & (top - 1)) == 0;  // every run ascending at k == top
          const unsigned long long x = K[i], y = K[ixj];
          if ((x > y) == up) {
            K[i] = y;
            K[ixj] = x;
          }
        }
        // With j <= 64, the pairs of thread t lie in entries [2t - 2t%128,
        // +128): wave w owns entries 128w + 512r .. +127 in this stage and in
        // the next one if that also has j <= 64, so a wave barrier orders it.
        const uint32_t jn = j > 1 ? (j >> 1) : k;  // next stage: (k, j/2) or (2k, k)
        if ((j == 1 && k == top) || j > 64 || jn > 64) {
          __syncthreads();
        } else {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      }
    }
    const bool last = level == L - 1;
    uint32_t ev[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      uint32_t e = 0;
      if (i < m) {
        if (seg) {
          const uint32_t q = G[i], gs = nodeOff[2 * q];
          e = E[gs + (uint32_t)(K[q * S + (i - gs)] & 0x7FFull)];
        } else {
          e = E[(uint32_t)(K[i] & 0x7FFull)];
        }
      }
      ev[r] = e;
      kv[r] = (i < m && !last) ? fkey(ba[e].y) : 0u;  // max keys, issued together
    }
    __syncthreads();  // every sorted slot is read: E2 and MK overwrite them
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      if (i < m) {
        E2[i] = ev[r];
        MK[i] = kv[r];
      }
    }
    __syncthreads();
    if (last) {
      for (uint32_t j = tid; j < c; j += kThreads) {
        const uint32_t o = nodeOff[j];
        P[2 * (size_t)(f + j)] = orig ? orig[E2[o]] : E2[o];
        P[2 * (size_t)(f + j) + 1] = orig ? orig[E2[o + 1]] : E2[o + 1];
      }
      break;
    }
    // first maximum of the max key over each node's [o + 1, e): one wave per
    // node while nodes are large, one thread per node once a sort group (two
    // nodes) holds at most 64 entries
    if (gmax > 64) {
      for (uint32_t j = wave; j < c; j += kThreads / 64) {
        const uint32_t o = nodeOff[j], e = nodeOff[j + 1];
        uint32_t bk = 0, bp = 0xFFFFFFFFu;
        for (uint32_t i = o + 1 + lane; i < e; i += 64) {
          const uint32_t k = MK[i];
          if (k > bk) {
            bk = k;
            bp = i;
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const uint32_t ok = __shfl_xor(bk, off), op = __shfl_xor(bp, off);
          if (ok > bk || (ok == bk && op < bp)) {
            bk = ok;
            bp = op;
          }
        }
        if (lane == 0) nodeArg[j] = (bp > o && bp < e) ? bp : o + 1;  // sizes are >= 2
      }
    } else {
      for (uint32_t j = tid; j < c; j += kThreads) {
        const uint32_t o = nodeOff[j], e = nodeOff[j + 1];
        uint32_t bk = 0, bp = 0xFFFFFFFFu;
        for (uint32_t i = o + 1; i < e; ++i) {
          const uint32_t k = MK[i];
          if (k > bk) {
            bk = k;
            bp = i;
          }
        }
        nodeArg[j] = (bp > o && bp < e) ? bp : o + 1;
      }
    }
    __syncthreads();
    {
      // each thread's pairs (c <= kSubT / 2): last-level keys loaded together
      constexpr int kPairs = kSubT / 2 / kThreads;
      const float2* knew = bb + (size_t)((L - 1) % 3) * N;
      uint32_t px[kPairs], py[kPairs];
      float fx[kPairs], fy[kPairs];
#pragma unroll
      for (int r = 0; r < kPairs; ++r) {
        const uint32_t j = tid + r * kThreads;
        px[r] = j < c ? E2[nodeOff[j]] : 0u;
        py[r] = j < c ? E2[nodeArg[j]] : 0u;
        fx[r] = j < c ? knew[px[r]].x : 0.0f;
        fy[r] = j < c ? knew[py[r]].x : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < kPairs; ++r) {
        const uint32_t j = tid + r * kThreads;
        if (j < c) write_pair_k(f + j, px[r], py[r], fx[r], fy[r], level, L, bb, N, P, orig);
      }
    }
    // compaction: the node of sorted position i is one of its sort group's
    // two nodes (groups are contiguous and keep their ranges through the
    // sort); all reads of G come before the writes
    uint32_t nodeOf[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      uint32_t j = 0;
      if (i < m) {
        j = 2u * G[i];
        if (j + 1 < c && i >= nodeOff[j + 1]) ++j;
      }
      nodeOf[r] = j;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const uint32_t i = tid + r * kThreads;
      if (i >= m) break;
      const uint32_t lo = nodeOf[r], o = nodeOff[lo];
      if (i >= o + 2) {
        const uint32_t val = (i == nodeArg[lo]) ? E2[o + 1] : E2[i];
        const uint32_t dst = i - 2 * (lo + 1);
        E[dst] = val;
        G[dst] = (uint16_t)lo;
      }
    }
    __syncthreads();
    m -= 2 * c;
    const uint64_t nf = 2ull * f + 1;
    if (nf >= n) break;
    const uint64_t nc = std::min<uint64_t>(2ull * c, n - nf);
    f = (uint32_t)nf;
    c = (uint32_t)nc;
  }
}

// applyIndices (rebuilder.cu:217-245): gather every per-triangle array.
__global__ void k_gather_geometry(const uint32_t* __restrict__ P, uint32_t N,
                                  const float* __restrict__ iv, const float* __restrict__ in,
                                  const float* __restrict__ iuv, const uint16_t* __restrict__ ish,
                                  float* __restrict__ ov, float* __restrict__ on,
                                  float* __restrict__ ouv, uint16_t* __restrict__ osh) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t s = clampi(P[i], N);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    ov[(size_t)i * 9 + k] = iv[s * 9 + k];
    on[(size_t)i * 9 + k] = in[s * 9 + k];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) ouv[(size_t)i * 6 + k] = iuv[s * 6 + k];
  osh[i] = ish[s];
}

__device__ inline void slab_of(const float* v, int a, float& lo, float& hi) {
  // intersect.cuh:25-33, 109-111 over the node's two triangles
  const float* p = v;
  const float* q = v + 9;
  lo = fminf(fminf(fminf(p[3 * a], p[3 * a + 1]), p[3 * a + 2]),
             fminf(fminf(q[3 * a], q[3 * a + 1]), q[3 * a + 2]));
  hi = fmaxf(fmaxf(fmaxf(p[3 * a], p[3 * a + 1]), p[3 * a + 2]),
             fmaxf(fmaxf(q[3 * a], q[3 * a + 1]), q[3 * a + 2]));
}

// Record k at k * rec_bytes; bytes after the 96 of a record are zero.
__device__ inline void put_record(NodeRec* __restrict__ nodes, uint32_t k, uint32_t rec_bytes,
                                  const NodeRec& r) {
  NodeRec* dst = node_at(nodes, k, rec_bytes);
  *dst = r;
  float4* pad = reinterpret_cast<float4*>(dst + 1);
  for (uint32_t q = 0; q < (rec_bytes - kNodeBytes) / 16; ++q) pad[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_node_records(const float* __restrict__ v, uint32_t n, uint32_t rec_bytes,
                               NodeRec* __restrict__ nodes) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  NodeRec r;
  if (k == n) {  // the root's own slab, on axis 0
#pragma unroll
    for (int t = 0; t < 18; ++t) (&r.tri[0][0])[t] = 0.0f;
    slab_of(v, 0, r.cslab[0], r.cslab[1]);
    r.cslab[2] = r.cslab[3] = 0.0f;
    r.flags = 0;
    r.spare = 0;
    put_record(nodes, k, rec_bytes, r);
    return;
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const float* p = v + (size_t)(2 * k + t) * 9;  // points layout
    const V3 v1 = mk(p[0], p[3], p[6]), v2 = mk(p[1], p[4], p[7]), v3 = mk(p[2], p[5], p[8]);
    const V3 e1 = v2 - v1, e2 = v3 - v1;
    r.tri[t][0] = v1.x;
    r.tri[t][1] = v1.y;
    r.tri[t][2] = v1.z;
    r.tri[t][3] = e1.x;
    r.tri[t][4] = e1.y;
    r.tri[t][5] = e1.z;
    r.tri[t][6] = e2.x;
    r.tri[t][7] = e2.y;
    r.tri[t][8] = e2.z;
  }
  const uint32_t d = depth_of(k);
  const int ca = (int)((d + 1) % 3);
  r.cslab[0] = r.cslab[2] = 0.0f;
  r.cslab[1] = r.cslab[3] = 0.0f;
  const uint64_t cl = 2ull * k + 1;
  if (cl < n) slab_of(v + cl * 18, ca, r.cslab[0], r.cslab[1]);
  if (cl + 1 < n) slab_of(v + (cl + 1) * 18, ca, r.cslab[2], r.cslab[3]);
  bool normal = true;
#pragma unroll
  for (int c = 0; c < 4; ++c) normal = normal && normal_coord(r.cslab[c]);
  r.flags = normal ? kNodeSlabsNormal : 0u;
  r.spare = 0;
  put_record(nodes, k, rec_bytes, r);
}

// cudaGenerateTB (utility.cu:136-180).
__global__ void k_generate_tb(const float* __restrict__ normals, uint32_t N,
                              float* __restrict__ tangents, float* __restrict__ bitangents) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const V3 up = mk(0.0f, 1.0f, 0.0f), fwd = mk(0.0f, 0.0f, 1.0f);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3 nv = ld3(normals + (size_t)i * 9 + 3 * k);
    V3 T = cross(nv, up);
    if (length(T) < kEpsilon) T = cross(nv, fwd);
    const V3 B = cross(nv, T);
    float* tp = tangents + (size_t)i * 9 + 3 * k;
    float* bp = bitangents + (size_t)i * 9 + 3 * k;
    tp[0] = T.x;
    tp[1] = T.y;
    tp[2] = T.z;
    bp[0] = B.x;
    bp[1] = B.y;
    bp[2] = B.z;
  }
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

struct LevelPlan {
  int level;
  uint32_t m, first, nchildren, ngroups, ntiles, nchunks;
  size_t child_base, group_base, tile_base, chunk_base;
};

}  // namespace

aur_status generate_tb(aur_context* ctx) {
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "generate_tb: geometry not built");
  const uint32_t N = ctx->count;
  AUR_HIP(ctx, ctx->tangents.reserve((size_t)N * 9));
  AUR_HIP(ctx, ctx->bitangents.reserve((size_t)N * 9));
  k_generate_tb<<<blocks(N, 256), 256, 0, ctx->stream>>>(ctx->normals.p, N, ctx->tangents.p,
                                                         ctx->bitangents.p);
  AUR_LAUNCH_CHECK(ctx, "k_generate_tb");
  ctx->has_tb = true;
  return AUR_OK;
}

aur_status nmh_build(aur_context* ctx, aur_build_stats* stats) {
  const uint32_t N = ctx->count;
  if (N < 2 || (N & 1u)) return ctx->fail(AUR_ERR_STATE, "rebuild: no even-sized geometry uploaded");
  const uint32_t n = N / 2;
  int L = 0;
  for (uint32_t x = n; x; x >>= 1) ++L;  // log2i(n) + 1
  hipStream_t s = ctx->stream;

  // ---- host plan (closed forms only; no device data needed) ----
  int level0 = 0;
  if (N > kSubT) {
    level0 = L - 1;
    for (int l = 1; l < L; ++l) {
      const uint64_t fp = (1ull << (l - 1)) - 1;  // first node of the parent level
      const uint64_t maxg = 2 * desc_count(fp, fp + 1, n) - 2;
      if (maxg <= kSubT) {
        level0 = l;
        break;
      }
    }
  }
  std::vector<TileDesc> tiles;
  std::vector<ChildDesc> children;
  std::vector<GroupDesc> groups;
  uint32_t chunk_count = 0;  // chunks of kChunk tiles of one group, per level
  std::vector<uint32_t> chunk_tiles;  // per level: first tile of each chunk, then the level's tile count
  std::vector<LevelPlan> plan;
  size_t max_tiles = 1, max_children = 1, max_chunks = 1;
  for (int l = 0; l < level0; ++l) {
    LevelPlan lp{};
    lp.level = l;
    lp.first = (uint32_t)((1ull << l) - 1);
    const uint64_t last = std::min<uint64_t>((2ull << l) - 1, n);
    lp.m = (uint32_t)(N - 2ull * lp.first);
    lp.child_base = children.size();
    lp.group_base = groups.size();
    lp.tile_base = tiles.size();
    uint64_t off = 0;
    int64_t cur_parent = -1;
    for (uint64_t h = lp.first; h < last; ++h) {
      const uint64_t size = 2 * desc_count(h, h + 1, n);
      const int64_t parent = l == 0 ? 0 : (int64_t)((h - 1) / 2);
      if (parent != cur_parent) {
        if (!groups.empty() && groups.size() > lp.group_base) groups.back().te = (uint32_t)(tiles.size() - lp.tile_base);
        groups.push_back(GroupDesc{(uint32_t)off, 0, (uint32_t)(tiles.size() - lp.tile_base), 0, 0, 0});
        cur_parent = parent;
      }
      GroupDesc& g = groups.back();
      g.size += (uint32_t)size;
      ChildDesc cd{(uint32_t)off, (uint32_t)size, (uint32_t)(tiles.size() - lp.tile_base), 0};
      for (uint64_t st = 0; st < size; st += kTile)
        tiles.push_back(TileDesc{(uint32_t)(off + st), (uint32_t)std::min<uint64_t>(kTile, size - st),
                                 (uint32_t)(h - lp.first),
                                 (uint32_t)(groups.size() - 1 - lp.group_base)});
      cd.te = (uint32_t)(tiles.size() - lp.tile_base);
      children.push_back(cd);
      off += size;
    }
    groups.back().te = (uint32_t)(tiles.size() - lp.tile_base);
    lp.nchildren = (uint32_t)(children.size() - lp.child_base);
    lp.ngroups = (uint32_t)(groups.size() - lp.group_base);
    lp.ntiles = (uint32_t)(tiles.size() - lp.tile_base);
    chunk_count = 0;
    lp.chunk_base = chunk_tiles.size();
    for (size_t gi = lp.group_base; gi < groups.size(); ++gi) {
      GroupDesc& g = groups[gi];
      g.cb = chunk_count;
      for (uint32_t t = g.tb; t < g.te; t += kChunk) chunk_tiles.push_back(t);
      chunk_count += (g.te - g.tb + kChunk - 1) / kChunk;
      g.ce = chunk_count;
    }
    chunk_tiles.push_back(lp.ntiles);
    lp.nchunks = chunk_count;
    max_chunks = std::max<size_t>(max_chunks, lp.nchunks);
    max_tiles = std::max<size_t>(max_tiles, lp.ntiles);
    max_children = std::max<size_t>(max_children, lp.nchildren);
    plan.push_back(lp);
  }
  std::vector<SubGroup> sub;
  if (level0 == 0) {
    sub.push_back(SubGroup{0, N, 0, 1});
  } else {
    const uint64_t fp = (1ull << (level0 - 1)) - 1;
    const uint64_t lastp = std::min<uint64_t>((1ull << level0) - 1, n);
    const uint64_t fc = (1ull << level0) - 1;
    for (uint64_t p = fp; p < lastp; ++p) {
      const uint64_t c0 = 2 * p + 1;
      if (c0 >= n) break;
      const uint64_t c1 = std::min<uint64_t>(c0 + 2, n);
      const uint64_t off = 2 * desc_count(fc, c0, n);
      const uint64_t size = 2 * desc_count(c0, c1, n);
      sub.push_back(SubGroup{(uint32_t)off, (uint32_t)size, (uint32_t)c0, (uint32_t)(c1 - c0)});
    }
  }

  // ---- device buffers ----
  DevBuf<float2> bb;          // (min, max) per axis and triangle
  DevBuf<float2> bbR;         // the same for the subtree phase's pending region, by position
  // sort payload (triangle id, max key on the level's axis): the level's
  // input, the other side of the radix passes, and the emit pass's output
  DevBuf<uint2> Va, Vb, Vc;
  DevBuf<uint2> GP, TP, TR;  // (smallest key, passes) per group and tile; emit tiles' key ranges
  DevBuf<uint32_t> CP;        // passes per offset chunk
  DevBuf<uint32_t> Ka, Kb, H, O, CS, CB, argpos, dchunk_tiles;
  DevBuf<uint2> partial;
  DevBuf<TileDesc> dtiles;
  DevBuf<ChildDesc> dchildren;
  DevBuf<GroupDesc> dgroups;
  DevBuf<SubGroup> dsub;
  AUR_HIP(ctx, bb.reserve((size_t)3 * N));
  AUR_HIP(ctx, Va.reserve(N));
  AUR_HIP(ctx, Vb.reserve(N));
  AUR_HIP(ctx, ctx->perm.reserve(N));
  if (!plan.empty()) {
    AUR_HIP(ctx, Vc.reserve(N));
    AUR_HIP(ctx, GP.reserve(max_children));
    AUR_HIP(ctx, TP.reserve(max_tiles));
    AUR_HIP(ctx, CP.reserve(max_chunks));
    AUR_HIP(ctx, TR.reserve(max_tiles));
    AUR_HIP(ctx, Ka.reserve(N));
    AUR_HIP(ctx, Kb.reserve(N));
    AUR_HIP(ctx, H.reserve(max_tiles * 256));
    AUR_HIP(ctx, O.reserve(max_tiles * 256));
    AUR_HIP(ctx, dchunk_tiles.reserve(chunk_tiles.size()));
    AUR_HIP(ctx, hipMemcpy(dchunk_tiles.p, chunk_tiles.data(), chunk_tiles.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice));
    AUR_HIP(ctx, CS.reserve(max_chunks * 256));
    AUR_HIP(ctx, CB.reserve(max_chunks * 256));
    AUR_HIP(ctx, partial.reserve(max_tiles));
    AUR_HIP(ctx, argpos.reserve(max_children));
    AUR_HIP(ctx, dtiles.reserve(tiles.size()));
    AUR_HIP(ctx, dchildren.reserve(children.size()));
    AUR_HIP(ctx, dgroups.reserve(groups.size()));
    AUR_HIP(ctx, hipMemcpy(dtiles.p, tiles.data(), tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice));
    AUR_HIP(ctx, hipMemcpy(dchildren.p, children.data(), children.size() * sizeof(ChildDesc),
                           hipMemcpyHostToDevice));
    AUR_HIP(ctx, hipMemcpy(dgroups.p, groups.data(), groups.size() * sizeof(GroupDesc),
                           hipMemcpyHostToDevice));
  }
  // pending triangles when the subtree phase starts
  const uint32_t m0 = level0 > 0 ? (uint32_t)(N - 2 * ((1ull << level0) - 1)) : N;
  if (level0 > 0) AUR_HIP(ctx, bbR.reserve((size_t)3 * m0));
  AUR_HIP(ctx, dsub.reserve(sub.size()));
  AUR_HIP(ctx, hipMemcpy(dsub.p, sub.data(), sub.size() * sizeof(SubGroup), hipMemcpyHostToDevice));

  hipEvent_t e0, e1;
  AUR_HIP(ctx, hipEventCreate(&e0));
  AUR_HIP(ctx, hipEventCreate(&e1));
  AUR_HIP(ctx, hipEventRecord(e0, s));

  k_bounds<<<blocks(N, 256), 256, 0, s>>>(ctx->in_vertices.p, N, bb.p);
  // level 0's keys come with the identity order; every later level's keys
  // and max keys are gathered by the previous level's emit pass
  uint2* Vcur = Va.p;  // the level's input (k_iota's, then the previous emit's output)
  uint2* Valt = Vb.p;  // the other side of the radix passes
  uint2* Vout = Vc.p;  // the emit pass's output
  k_iota<<<blocks(N, 256), 256, 0, s>>>(Vcur, N, bb.p, plan.empty() ? nullptr : Ka.p);
  k_fill0<<<blocks(N, 256), 256, 0, s>>>(ctx->perm.p, N);
  if (!plan.empty()) k_fill0<<<blocks(max_chunks * 256, 256), 256, 0, s>>>(CS.p, (uint32_t)(max_chunks * 256));
  for (size_t li = 0; li < plan.size(); ++li) {
    const LevelPlan& lp = plan[li];
    const TileDesc* tl = dtiles.p + lp.tile_base;
    const ChildDesc* ch = dchildren.p + lp.child_base;
    const GroupDesc* gr = dgroups.p + lp.group_base;
    // the previous level's nodes: group g of this level is node g's remainder
    const ChildDesc* prev = li == 0 ? nullptr : dchildren.p + plan[li - 1].child_base;
    uint32_t *kin = Ka.p, *kout = Kb.p;
    uint2 *vin = Vcur, *vout = Valt;
    // every launch of a pass skips the groups that need fewer passes
    for (uint32_t pass = 0; pass < 4; ++pass) {
      k_radix_hist<<<lp.ntiles, kThreads, 0, s>>>(tl, gr, TP.p, kin, pass, H.p, CS.p);
      k_radix_group_scan<<<lp.ngroups, kThreads, 0, s>>>(gr, GP.p, TP.p, CP.p, prev, TR.p, pass, CS.p, CB.p);
      k_radix_chunk_offsets<<<lp.nchunks, kThreads, 0, s>>>(dchunk_tiles.p + lp.chunk_base, CP.p, pass,
                                                            H.p, CB.p, O.p);
      k_radix_scatter<<<lp.ntiles, kThreads, 0, s>>>(tl, TP.p, kin, vin, pass, O.p, lp.m, kout, vout);
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    // each group's sorted data is in Vcur or Valt (group_passes); the keys
    // are not read again, so Ka takes the next level's
    const bool last = li + 1 == plan.size();
    k_argmax_partial<<<lp.ntiles, kThreads, 0, s>>>(tl, ch, TP.p, Vcur, Valt, partial.p);
    k_argmax_combine<<<blocks(lp.nchildren, kThreads / 64), kThreads, 0, s>>>(ch, lp.nchildren,
                                                                              partial.p, argpos.p);
    k_emit<<<lp.ntiles, kThreads, 0, s>>>(tl, ch, argpos.p, TP.p, Vcur, Valt, Vout, lp.first, lp.level, L,
                                          bb.p, N, ctx->perm.p, last ? nullptr : Ka.p, TR.p,
                                          last ? bbR.p : nullptr, last ? Ka.p : nullptr, m0);
    // rotate: the emit output is the next level's input
    uint2* const free_buf = Vcur;
    Vcur = Vout;
    Vout = Valt;
    Valt = free_buf;
  }
  AUR_LAUNCH_CHECK(ctx, "nmh level kernels");
  if (level0 > 0) {
    // the subtree phase on positions (renumbered by the last k_emit)
    k_subtree<<<(uint32_t)sub.size(), kThreads, 0, s>>>(dsub.p, Vcur, bbR.p, m0, n, L, level0,
                                                        ctx->perm.p, Ka.p);
  } else {
    k_subtree<<<(uint32_t)sub.size(), kThreads, 0, s>>>(dsub.p, Vcur, bb.p, N, n, L, level0,
                                                        ctx->perm.p, nullptr);
  }
  AUR_LAUNCH_CHECK(ctx, "k_subtree");

  AUR_HIP(ctx, ctx->vertices.reserve((size_t)N * 9));
  AUR_HIP(ctx, ctx->normals.reserve((size_t)N * 9));
  AUR_HIP(ctx, ctx->texcoords.reserve((size_t)N * 6));
  AUR_HIP(ctx, ctx->shader_ids.reserve(N));
  // records one cache line apart once the compact array would fill more than
  // half of the 256 MB Infinity Cache (node visits then go to HBM)
  const uint32_t rec_bytes =
      (uint64_t)(n + 1) * kNodeBytes > kCompactNodeArrayMax ? kNodeLineBytes : kNodeBytes;
  AUR_HIP(ctx, ctx->nodes.reserve(((size_t)(n + 1) * rec_bytes + sizeof(NodeRec) - 1) /
                                  sizeof(NodeRec)));  // + the root's own slab
  k_gather_geometry<<<blocks(N, 256), 256, 0, s>>>(ctx->perm.p, N, ctx->in_vertices.p,
                                                   ctx->in_normals.p, ctx->in_texcoords.p,
                                                   ctx->in_shader_ids.p, ctx->vertices.p,
                                                   ctx->normals.p, ctx->texcoords.p,
                                                   ctx->shader_ids.p);
  k_node_records<<<blocks((uint64_t)n + 1, 256), 256, 0, s>>>(ctx->vertices.p, n, rec_bytes,
                                                               ctx->nodes.p);
  AUR_LAUNCH_CHECK(ctx, "gather/node records");
  AUR_HIP(ctx, hipEventRecord(e1, s));
  AUR_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0.0f;
  AUR_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the root's own slab (cslab[0..1] of record n) for launches that start
  // every ray at the root without reading it
  float root[2] = {0.0f, 0.0f};
  AUR_HIP(ctx, hipMemcpyAsync(root, &node_at(ctx->nodes.p, n, rec_bytes)->cslab[0], sizeof(root),
                              hipMemcpyDeviceToHost, s));
  AUR_HIP(ctx, hipStreamSynchronize(s));
  ctx->root_slab = make_float2(root[0], root[1]);
  ctx->rec_bytes = rec_bytes;
  ctx->levels = (uint32_t)L;
  ctx->built = true;
  ctx->has_tb = false;
  if (stats) {
    stats->count = N;
    stats->levels = (uint32_t)L;
    stats->build_ms = ms;
  }
  return AUR_OK;
}

}  // namespace aur
