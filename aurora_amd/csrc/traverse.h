// NMH traversal on gfx950: the reference's intersect()/intersectAny()
// (src/kernels/lib/intersect.cuh:36-152) with identical results.
//
// The reference pops a state, tests the node's slab on the node's axis
// (Ray::intersect(slab), src/kernels/lib/ray.cuh:49-61), tests both triangles
// (one-sided Moller-Trumbore, ray.cuh:22-47, strict t < ray.t so the first
// hit wins ties), then pushes R and L (L popped first; L/R swapped when the
// ray direction on the children's axis is negative).  The same tests run here
// in the same order, with changes that do not alter any result:
//   * a node's children slabs are stored in the node record, so both
//     children are tested when their parent is visited; the t-dependent part
//     of the test (range.x <= ray.t) is still applied when a child is taken,
//     exactly when the reference would pop it;
//   * the near child is taken directly; only the far child is deferred;
//   * a deferred child is the sibling of the current path's node at its
//     depth, so the per-lane stack holds only its range (8 bytes, a
//     compact LIFO stack in depth order) plus one trail bit per depth; the
//     child's heap index is recomputed from the last visited node when it is
//     resumed;
//   * the slab quotient (s - p) / d uses the ray's correctly rounded
//     reciprocal y = RN(1/d) and one residual correction,
//         q0 = RN(n y),  e = n - d q0 (exact, FMA),  q = RN(q0 + e y),
//     which equals RN(n / d) (Markstein's theorem) while no step under- or
//     overflows: used only for 2^-60 <= |d| <= 16 and n = 0 or
//     2^-96 <= |n| <= 2^60, true division otherwise.  Inside the loop the
//     numerator range follows from a per-node flag (all four child slab
//     values pass normal_coord) and a per-ray check of the origin;
//   * both triangles of a node are tested without branches (every value is
//     computed, the reference's early outs become one accept predicate with
//     the same NaN behaviour), and 1/det is HIP's correctly rounded
//     reciprocal sequence without its scaling and fix-up steps, which are
//     identities for 2^-24 <= det <= 2^64 (det < kEpsilon is rejected before
//     the value is used; det > 2^64 takes true division); the ray's
//     reciprocals use the same sequence (needed only for 2^-60 <= |d| <= 16).
//
// The traversal is a small state machine (trav_init / trav_step, one node
// visit per step) so that persistent kernels can refill finished lanes.
#pragma once

#include "common.h"

namespace aur {

// Bitwise (not short-circuit) conditions keep these tests free of branches.
__device__ __forceinline__ bool fast_num(float n) {
  const float a = fabsf(n);
  return (n == 0.0f) | ((a >= 0x1p-96f) & (a <= 0x1p60f));
}

// RN(n / d) by the reciprocal quotient (valid when fast_num(n) and fast_d).
__device__ __forceinline__ float quot_fast(float n, float d, float y) {
  const float q = n * y;
  return __builtin_fmaf(__builtin_fmaf(-d, q, n), y, q);
}

// v_min_f32 / v_max_f32 on operands that are never signaling NaNs (every
// operand here is the result of arithmetic, or a copy of one): with IEEE mode
// on, these instructions return the other operand for a quiet NaN, exactly
// fminf / fmaxf.  The compiler cannot prove that through the join of the fast
// and exact quotient paths or the loop-carried range, and would quiet each
// operand first (one more instruction per operand).
__device__ __forceinline__ float min_q(float a, float b) {
  float r;
  __asm__("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float max_q(float a, float b) {
  float r;
  __asm__("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

__device__ __forceinline__ float2 clip(float2 r, float d0, float d1) {
  return make_float2(max_q(min_q(d0, d1), r.x), min_q(max_q(d0, d1), r.y));
}

__device__ __forceinline__ float2 narrow(float2 r, float slo, float shi, float p, float d, float y,
                                         bool fast_d) {
  const float n0 = slo - p, n1 = shi - p;
  float d0 = quot_fast(n0, d, y), d1 = quot_fast(n1, d, y);
  const bool f0 = fast_num(n0), f1 = fast_num(n1);
  if (!(fast_d & f0 & f1)) {
    d0 = n0 / d;
    d1 = n1 / d;
  }
  return clip(r, d0, d1);
}

// Both children's slabs on one axis; `fast`: the divisor and all four
// numerators are in the reciprocal quotient's range (else one shared, rarely
// taken branch to true division).
__device__ __forceinline__ void narrow2(float2 r, float4 slabs, float p, float d, float y,
                                        bool fast, float2& rl, float2& rr) {
  const float n0 = slabs.x - p, n1 = slabs.y - p, n2 = slabs.z - p, n3 = slabs.w - p;
  float d0 = quot_fast(n0, d, y), d1 = quot_fast(n1, d, y);
  float d2 = quot_fast(n2, d, y), d3 = quot_fast(n3, d, y);
  if (!fast) {
    d0 = n0 / d;
    d1 = n1 / d;
    d2 = n2 / d;
    d3 = n3 / d;
  }
  rl = clip(r, d0, d1);
  rr = clip(r, d2, d3);
}

// RN(1/x) for 2^-60 <= |x| <= 2^64: the steps of HIP's IEEE division 1/x
// (reciprocal estimate, one Newton step, quotient and two residual
// corrections) without the operand scaling and result fix-up, which do
// nothing on this range.
constexpr float kRecipMax = 0x1p64f;
__device__ __forceinline__ float recip_rn(float x) {
  const float r0 = __builtin_amdgcn_rcpf(x);
  const float r1 = __builtin_fmaf(__builtin_fmaf(-x, r0, 1.0f), r0, r0);
  const float q1 = __builtin_fmaf(__builtin_fmaf(-x, r1, 1.0f), r1, r1);
  return __builtin_fmaf(__builtin_fmaf(-x, q1, 1.0f), r1, q1);
}

// The reference's one-sided Moller-Trumbore test after its determinant:
// accept iff !(det < eps), u in [0,1], v >= 0 and u + v <= 1 (NaN passes
// these as in the reference and is rejected by the caller's t test).
__device__ __forceinline__ bool tri_finish(V3 o, V3 d, V3 v1, V3 e1, V3 e2, V3 P, float det,
                                           float invdet, float& u, float& v, float& t) {
  const V3 T = o - v1;
  u = dot(T, P) * invdet;
  const V3 Q = cross(T, e1);
  v = dot(d, Q) * invdet;
  t = dot(e2, Q) * invdet;
  return !(det < kEpsilon) & !(u < 0.0f) & !(u > 1.0f) & !(v < 0.0f) & !(u + v > 1.0f);
}

struct TraceResult {
  int32_t id;
  float t, u, v;
};

// Traversal counters of the reference's loop: popped = states popped with an
// in-range index (P), passed = states whose slab test passed (S).  Exact for
// closest-hit traversal, where every pushed state is eventually popped.
struct TraceCount {
  uint32_t popped = 0, passed = 0;
};

struct Trav {
  V3 o, d, y;      // origin, direction, RN(1/direction)
  uint32_t fast;   // bit a: 2^-60 <= |d_a| <= 16 and normal_coord(o_a)
  float t, u, v;   // current closest hit (t = ray.t)
  int32_t id;
  uint32_t k, depth, trail;
  uint32_t axis;   // depth % 3
  bool any;        // any-hit query (persistent_trace's mixed launches)
  float2 r;        // range of node k after its slab test
};

__device__ __forceinline__ float compf(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Per-lane stack of deferred ranges.  A deferred child is always deeper
// than every entry already pending (entries sit on the current path's
// siblings and are popped deepest first), so the entries form a LIFO stack
// sorted by depth and entry i is at index i = popcount(trail) at push time:
// at most levels - 1 entries, indices <= levels - 2.
//
// LdsStack: every row in LDS (lane-strided, conflict-free).
struct LdsStack {
  float2* lds;  // row r of this lane at lds[r * stride]
  uint32_t stride;
  __device__ __forceinline__ void store(uint32_t r, float2 v) const { lds[r * stride] = v; }
  __device__ __forceinline__ float2 load(uint32_t r) const { return lds[r * stride]; }
};

// SplitStack: rows [0, klds) in LDS, deeper rows in a per-lane column of
// global memory, so that a small LDS window leaves room for more resident
// waves.  The LDS side uses an LDS-qualified pointer and the global side
// stores through a global pointer and loads through a buffer resource, so
// the two stay separate instructions (merged, they would become generic
// "flat" accesses, slower for the common LDS case).
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) vf2 lds_vf2;
typedef __attribute__((address_space(1))) vf2 glb_vf2;
struct SplitStack {
  lds_vf2* lds;            // row r < klds at lds[r * lstride]
  uint32_t lstride;
  glb_vf2* glb;            // row r >= klds at glb[(r - klds) * gstride]
  __amdgpu_buffer_rsrc_t rsrc;  // the whole global stack buffer
  uint32_t lane_bytes;     // byte offset of glb from the buffer start
  uint32_t gstride;
  uint32_t klds;
  __device__ __forceinline__ void store(uint32_t r, float2 v) const {
    const vf2 w = {v.x, v.y};
    if (r < klds)
      lds[r * lstride] = w;
    else
      glb[(size_t)(r - klds) * gstride] = w;
  }
  __device__ __forceinline__ float2 load(uint32_t r) const {
    if (r < klds) {
      const vf2 w = lds[r * lstride];
      return make_float2(w.x, w.y);
    }
    const auto w = __builtin_amdgcn_raw_buffer_load_b64(rsrc, lane_bytes + (r - klds) * gstride * 8u,
                                                        0, 0);
    return make_float2(__uint_as_float(w[0]), __uint_as_float(w[1]));
  }
};

__device__ __forceinline__ LdsStack lds_stack(float2* lane_column, uint32_t stride) {
  return LdsStack{lane_column, stride};
}

__device__ __forceinline__ uint32_t fast_axis(float o, float d) {
  const float a = fabsf(d);
  return (uint32_t)((a >= 0x1p-60f) & (a <= 16.0f) & normal_coord(o));
}

// The root's own slab: cslab[0..1] of the record after the last node.
__device__ __forceinline__ float2 root_slab(const NodeRec* __restrict__ nodes, uint32_t n_nodes,
                                            uint32_t rec_bytes) {
  const float4 rs = reinterpret_cast<const float4*>(node_at(nodes, n_nodes, rec_bytes))[4];
  return make_float2(rs.z, rs.w);
}

// Sets up a ray; returns false when the root's slab culls it (finished, miss).
// The reciprocals y serve only axes with 2^-60 <= |d| <= 16, where recip_rn
// equals 1.0f / d.
template <bool kCount>
__device__ __forceinline__ bool trav_init_root(Trav& s, float2 root, V3 o, V3 d, float tmax,
                                               TraceCount* cnt) {
  s.o = o;
  s.d = d;
  s.y = mk(recip_rn(d.x), recip_rn(d.y), recip_rn(d.z));
  // keep the reciprocals in registers: without this barrier the compiler
  // re-derives them at every use inside the loop
  __asm__ volatile("" : "+v"(s.y.x), "+v"(s.y.y), "+v"(s.y.z));
  s.fast = fast_axis(o.x, d.x) | (fast_axis(o.y, d.y) << 1) | (fast_axis(o.z, d.z) << 2);
  s.t = tmax;
  s.u = s.v = 0.0f;
  s.id = -1;
  s.k = s.depth = s.trail = s.axis = 0;
  s.any = false;
  if (kCount) cnt->popped += 1;  // the root
  s.r = narrow(make_float2(kEpsilon, kInfinity), root.x, root.y, o.x, d.x, s.y.x, s.fast & 1u);
  return (s.r.x <= s.r.y) && (s.r.x <= s.t);
}

template <bool kCount>
__device__ __forceinline__ bool trav_init(Trav& s, const NodeRec* __restrict__ nodes,
                                          uint32_t n_nodes, uint32_t rec_bytes, V3 o, V3 d,
                                          float tmax, TraceCount* cnt) {
  return trav_init_root<kCount>(s, root_slab(nodes, n_nodes, rec_bytes), o, d, tmax, cnt);
}

// Visits node s.k; returns true when the traversal has finished.  All paths
// (take a child, resume a deferred child, finish) leave through one exit and
// the next node is chosen by selects: with several early exits the compiler's
// structured control flow ran the node visit separately for lanes arriving
// from different paths, doubling the instructions issued per visit.
// kMixed: the query kind is the lane's s.any instead of kAny.
template <bool kAny, bool kCount, bool kMixed = false, class Stk>
__device__ __forceinline__ bool trav_step(Trav& s, const NodeRec* __restrict__ nodes,
                                          uint32_t n_nodes, uint32_t rec_bytes, const Stk& stk,
                                          TraceCount* cnt) {
  const bool any = kAny || (kMixed && s.any);
  const uint32_t k = s.k;
  if (kCount) {
    cnt->passed += 1;
    cnt->popped += (2 * k + 1 < n_nodes) + (2 * k + 2 < n_nodes);
  }
  // the whole record in one round of loads (the right child's slab too, used
  // only when the node has two children: a load issued inside that branch
  // would wait a second memory latency per visit)
  const float4* rec = reinterpret_cast<const float4*>(node_at(nodes, k, rec_bytes));
  const float4 a0 = rec[0], a1 = rec[1], a2 = rec[2], a3 = rec[3], a4 = rec[4];
  const float4 a5 = rec[5];  // right child's slab, flags
  // both triangles up to their determinants, then both reciprocals
  const V3 v1a = mk(a0.x, a0.y, a0.z), e1a = mk(a0.w, a1.x, a1.y), e2a = mk(a1.z, a1.w, a2.x);
  const V3 v1b = mk(a2.y, a2.z, a2.w), e1b = mk(a3.x, a3.y, a3.z), e2b = mk(a3.w, a4.x, a4.y);
  const V3 Pa = cross(s.d, e2a), Pb = cross(s.d, e2b);
  const float deta = dot(e1a, Pa), detb = dot(e1b, Pb);
  float ia = recip_rn(deta), ib = recip_rn(detb);
  if ((deta > kRecipMax) | (detb > kRecipMax)) {
    ia = deta > kRecipMax ? 1.0f / deta : ia;
    ib = detb > kRecipMax ? 1.0f / detb : ib;
  }
  float ua, va, ta, ub, vb, tb;
  const bool hita = tri_finish(s.o, s.d, v1a, e1a, e2a, Pa, deta, ia, ua, va, ta) &
                    (ta > kEpsilon) & (ta < s.t);
  if (hita) {
    s.t = ta;
    s.u = ua;
    s.v = va;
    s.id = (int32_t)(2 * k);
  }
  const bool hitb = tri_finish(s.o, s.d, v1b, e1b, e2b, Pb, detb, ib, ub, vb, tb) &
                    (tb > kEpsilon) & (tb < s.t) & !(any & hita);
  if (hitb) {
    s.t = tb;
    s.u = ub;
    s.v = vb;
    s.id = (int32_t)(2 * k + 1);
  }
  const bool stop = any & (hita | hitb);  // any hit found
  bool found = false;
  uint32_t nk = 0, nd = 0, na = 0;
  float2 nrg = s.r;
  const uint32_t cl = 2 * k + 1;
  if (!stop && cl < n_nodes) {
    const uint32_t cd = s.depth + 1;
    const int ca = s.axis == 2 ? 0 : (int)s.axis + 1;
    const float oc = compf(s.o, ca), dc = compf(s.d, ca), yc = compf(s.y, ca);
    const bool fast = ((s.fast >> ca) & __float_as_uint(a5.z) & kNodeSlabsNormal) != 0u;
    float2 rl, rr;
    narrow2(s.r, make_float4(a4.z, a4.w, a5.x, a5.y), oc, dc, yc, fast, rl, rr);
    const bool okl = rl.x <= rl.y;
    const bool okr = (cl + 1 < n_nodes) & (rr.x <= rr.y);
    const bool sw = dc < 0.0f;
    const float2 near_r = sw ? rr : rl, far_r = sw ? rl : rr;
    // lane-mask logic rather than selects of booleans
    const bool nok = (sw & okr) | (!sw & okl), fok = (sw & okl) | (!sw & okr);
    const bool take_near = nok & (near_r.x <= s.t);
    const bool take_far = !take_near & fok & (far_r.x <= s.t);
    if (take_near && fok) {
      stk.store((uint32_t)__popc(s.trail), far_r);
      s.trail |= 1u << cd;
    }
    found = take_near || take_far;
    nk = (take_near != sw) ? cl : cl + 1;  // near = cl unless swapped; far is the other
    nrg = take_near ? near_r : far_r;
    nd = cd;
    na = (uint32_t)ca;
  }
  if (!stop && !found) {
    while (s.trail) {
      const uint32_t dd = 31u - (uint32_t)__clz(s.trail);
      const float2 frv = stk.load((uint32_t)__popc(s.trail) - 1u);
      // both halves in one LDS read (else the second is issued after the
      // test and the next step's loads wait for it)
      float fx = frv.x, fy = frv.y;
      __asm__ volatile("" : "+v"(fx), "+v"(fy));
      const float2 fr = make_float2(fx, fy);
      s.trail &= ~(1u << dd);
      if (fr.x <= s.t) {
        const uint32_t anc = ((k + 1) >> (s.depth - dd)) - 1;
        nk = (anc & 1u) ? anc + 1 : anc - 1;
        nd = dd;
        na = dd % 3;
        nrg = fr;
        found = true;
        break;
      }
    }
  }
  s.k = nk;
  s.depth = nd;
  s.axis = na;
  s.r = nrg;
  return !found;
}

// The tree a lane traverses: node records, their count, the root's own
// slab (cslab[0..1] of the record after the last node) and the distance
// between records.
struct TreeRef {
  const NodeRec* nodes;
  uint32_t n_nodes;
  float2 root;
  uint32_t rec_bytes;
};

// Persistent-wave traversal of a work range [0, n) of closest-hit and
// any-hit queries.  Wave w owns [w*chunk, (w+1)*chunk); between node visits,
// once at least `refill` lanes are idle (or all are), idle lanes take the
// next items of the range, ranked by ballot (no atomics).
// fetch(i, o, d, tmax, tree, any) loads item i, may replace the lane's tree
// (initially tree0; batches of scenes), sets any for an any-hit query and
// returns false if the item needs no traversal; store(i, s) receives the
// finished state.  A fetch that never writes `tree` keeps the tree
// wave-uniform.  (Splitting a refill over two iterations, so that the item
// loads overlap the active lanes' node step, measured slower: the new lanes
// lose a step.)
template <class Stk, class Fetch, class Store>
__device__ __forceinline__ void persistent_trace(TreeRef tree0, uint64_t n, uint32_t refill,
                                                 const Stk& stk, Fetch fetch, Store store) {
  TreeRef tree = tree0;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t chunk = (n + waves - 1) / waves;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint64_t next = wave * chunk;
  const uint64_t end = next + chunk < n ? next + chunk : n;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
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
            V3 o, d;
            float tmax;
            bool any = false;
            if (fetch(i, o, d, tmax, tree, any)) {
              item = i;
              active = trav_init_root<false>(s, tree.root, o, d, tmax, nullptr);
              s.any = any;
              if (!active) store(i, s);
            }
          }
        }
        next += nidle;
      } else if (nidle == 64) {
        break;
      }
    }
    if (active && trav_step<false, false, true>(s, tree.nodes, tree.n_nodes, tree.rec_bytes, stk, nullptr)) {
      store(item, s);
      active = false;
    }
  }
}

// Whole traversal of one ray by one lane (used by the render kernels).
// stk: this lane's stack column, stk[depth * stride].
template <bool kAny, bool kCount = false>
__device__ __forceinline__ TraceResult trace_ray(const NodeRec* __restrict__ nodes, uint32_t n_nodes,
                                                 uint32_t rec_bytes, V3 o, V3 d, float tmax,
                                                 float2* stk, uint32_t stride,
                                                 TraceCount* cnt = nullptr) {
  Trav s;
  const LdsStack st = lds_stack(stk, stride);
  if (trav_init<kCount>(s, nodes, n_nodes, rec_bytes, o, d, tmax, cnt)) {
    while (!trav_step<kAny, kCount>(s, nodes, n_nodes, rec_bytes, st, cnt)) {
    }
  }
  return TraceResult{s.id, s.t, s.u, s.v};
}

}  // namespace aur
