// NMH traversal on gfx950: the reference's intersect()/intersectAny()
// (src/kernels/lib/intersect.cuh:36-152) with identical results.
//
// The reference pops a state, tests the node's slab on the node's axis
// (Ray::intersect(slab), src/kernels/lib/ray.cuh:49-61), tests both triangles
// (one-sided Moller-Trumbore, ray.cuh:22-47, strict t < ray.t so the first
// hit wins ties), then pushes R and L (L popped first; L/R swapped when the
// ray direction on the children's axis is negative).  The same tests run here
// in the same order, with three changes that do not alter any result:
//   * a node's children slabs are stored in the node record, so both
//     children are tested when their parent is visited; the t-dependent part
//     of the test (range.x <= ray.t) is still applied when a child is taken,
//     exactly when the reference would pop it;
//   * the near child is taken directly; only the far child is deferred;
//   * a deferred child is the sibling of the current path's node at its
//     depth, so the per-lane stack holds only its range (8 bytes in LDS,
//     indexed by depth) plus one trail bit per depth; the child's heap index
//     is recomputed from the last visited node when it is resumed.
// Division stays the correctly rounded IEEE quotient, as the reference's
// culling and t values depend on it.
#pragma once

#include "common.h"

namespace aur {

__device__ __forceinline__ float2 narrow(float2 r, float slo, float shi, float p, float d) {
  const float d0 = (slo - p) / d;
  const float d1 = (shi - p) / d;
  return make_float2(fmaxf(fminf(d0, d1), r.x), fminf(fmaxf(d0, d1), r.y));
}

__device__ __forceinline__ bool tri_test(V3 o, V3 d, V3 v1, V3 e1, V3 e2, float& u, float& v,
                                         float& t) {
  const V3 P = cross(d, e2);
  const float det = dot(e1, P);
  if (det < kEpsilon) return false;
  const float invdet = 1.0f / det;
  const V3 T = o - v1;
  u = dot(T, P) * invdet;
  if (u < 0.0f || u > 1.0f) return false;
  const V3 Q = cross(T, e1);
  v = dot(d, Q) * invdet;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = dot(e2, Q) * invdet;
  return true;
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

// stk: this lane's stack column, stk[depth * stride].
template <bool kAny, bool kCount = false>
__device__ __forceinline__ TraceResult trace_ray(const NodeRec* __restrict__ nodes, uint32_t n_nodes,
                                                 V3 o, V3 d, float tmax, float2* stk,
                                                 uint32_t stride, TraceCount* cnt = nullptr) {
  TraceResult res{-1, tmax, 0.0f, 0.0f};
  if (kCount) cnt->popped += 1;  // the root
  const float4 rs = reinterpret_cast<const float4*>(nodes)[5];  // node 0: cslab.zw, pad
  float2 r = narrow(make_float2(kEpsilon, kInfinity), rs.z, rs.w, o.x, d.x);
  if (!((r.x <= r.y) && (r.x <= res.t))) return res;
  uint32_t k = 0, depth = 0, trail = 0;
  for (;;) {
    if (kCount) {
      cnt->passed += 1;
      cnt->popped += (2 * k + 1 < n_nodes) + (2 * k + 2 < n_nodes);
    }
    const float4* rec = reinterpret_cast<const float4*>(nodes + k);
    const float4 a0 = rec[0], a1 = rec[1], a2 = rec[2], a3 = rec[3], a4 = rec[4];
    float u, v, t;
    if (tri_test(o, d, mk(a0.x, a0.y, a0.z), mk(a0.w, a1.x, a1.y), mk(a1.z, a1.w, a2.x), u, v, t) &&
        t > kEpsilon && t < res.t) {
      res = TraceResult{(int32_t)(2 * k), t, u, v};
      if (kAny) return res;
    }
    if (tri_test(o, d, mk(a2.y, a2.z, a2.w), mk(a3.x, a3.y, a3.z), mk(a3.w, a4.x, a4.y), u, v, t) &&
        t > kEpsilon && t < res.t) {
      res = TraceResult{(int32_t)(2 * k + 1), t, u, v};
      if (kAny) return res;
    }
    const uint32_t cl = 2 * k + 1;
    if (cl < n_nodes) {
      const uint32_t cd = depth + 1;
      const int ca = (int)(cd % 3);
      const float oc = comp(o, ca), dc = comp(d, ca);
      const float2 rl = narrow(r, a4.z, a4.w, oc, dc);
      const bool okl = rl.x <= rl.y;
      float2 rr = make_float2(0.0f, 0.0f);
      bool okr = false;
      if (cl + 1 < n_nodes) {
        const float2 cs = reinterpret_cast<const float2*>(rec + 5)[0];
        rr = narrow(r, cs.x, cs.y, oc, dc);
        okr = rr.x <= rr.y;
      }
      const bool sw = dc < 0.0f;
      const float2 nr = sw ? rr : rl, fr = sw ? rl : rr;
      const bool nok = sw ? okr : okl, fok = sw ? okl : okr;
      const uint32_t ni = sw ? cl + 1 : cl, fi = sw ? cl : cl + 1;
      if (nok && nr.x <= res.t) {
        if (fok) {
          stk[cd * stride] = fr;
          trail |= 1u << cd;
        }
        k = ni;
        r = nr;
        depth = cd;
        continue;
      }
      if (fok && fr.x <= res.t) {
        k = fi;
        r = fr;
        depth = cd;
        continue;
      }
    }
    bool resumed = false;
    while (trail) {
      const uint32_t dd = 31u - (uint32_t)__clz(trail);
      trail &= ~(1u << dd);
      const float2 fr = stk[dd * stride];
      if (fr.x <= res.t) {
        const uint32_t anc = ((k + 1) >> (depth - dd)) - 1;
        k = (anc & 1u) ? anc + 1 : anc - 1;
        depth = dd;
        r = fr;
        resumed = true;
        break;
      }
    }
    if (!resumed) break;
  }
  return res;
}

}  // namespace aur
