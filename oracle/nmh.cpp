// CPU oracle: literal restatement of the reference NMH build.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.h).  This follows
// src/kernels/rebuilder.cu:311-450 step by step, one host loop per kernel:
//   computeTriangleBounds   rebuilder.cu:72-91
//   sortTriangles           rebuilder.cu:262-289  (two whole-array stable sorts)
//   swapMaxTriangleTPS      rebuilder.cu:127-151  (canonical at every level)
//   updateNodesTPS          rebuilder.cu:171-187
//   emitSplits(+Kernel)     rebuilder.cu:189-215, 247-260
//   calcPartition           rebuilder.cu:40-50
// The reference uses a racy atomicCAS chain (swapMaxTriangleTPT,
// rebuilder.cu:93-125) while 2048 or fewer splits are pending; its outcome
// under ties is not deterministic, so the deterministic rule of the
// thread-per-split kernel (first maximum by position wins) is used at every
// level.  Keys are compared as floats with '<' (vertex data is canonicalised
// so that -0.0 does not occur; NaN is rejected by the caller).
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <parallel/algorithm>
#include <vector>

#include "oracle_math.h"

namespace {

struct Split {
  uint32_t index = 0, size = 0;
  bool written = false;  // tracks which output slots this level wrote
};

void calc_partition(uint32_t N, uint32_t& L, uint32_t& R) {
  const uint32_t n = N / 2;
  const uint32_t H = orc::log2i(n);
  const uint32_t s = orc::exp2i(H - 1) - 1;
  const uint32_t S = orc::exp2i(H) - 1;
  const int Oi = int((n - 1) - s - S);
  const uint32_t O = (uint32_t)std::max(0, Oi);
  R = 2 * (s + O);
  L = 2 * (n - 1) - R;
}

// Stable sort, serial or on OpenMP threads (libstdc++ parallel mode); both
// produce the unique stable order.
template <class Cmp>
void stable_sort_n(std::vector<uint32_t>& v, Cmp cmp, int threads) {
  if (threads > 1 && v.size() > 65536) {
    omp_set_num_threads(threads);
    __gnu_parallel::stable_sort(v.begin(), v.end(), cmp);
  } else {
    std::stable_sort(v.begin(), v.end(), cmp);
  }
}

// sortTriangles: permutation = stable sort by key, then stable sort by node.
void sort_triangles(uint32_t count, const float* keys, std::vector<uint32_t>& indices,
                    std::vector<uint32_t>& nodes, std::vector<uint32_t>& perm,
                    std::vector<uint32_t>& tmp, int threads) {
  std::iota(perm.begin(), perm.end(), 0u);
  stable_sort_n(perm, [&](uint32_t a, uint32_t b) { return keys[indices[a]] < keys[indices[b]]; },
                threads);
  stable_sort_n(perm, [&](uint32_t a, uint32_t b) { return nodes[a] < nodes[b]; }, threads);
  for (uint32_t i = 0; i < count; ++i) tmp[i] = indices[perm[i]];
  indices.swap(tmp);
  for (uint32_t i = 0; i < count; ++i) tmp[i] = nodes[perm[i]];
  nodes.swap(tmp);
}

}  // namespace

extern "C" {

// calcPartition exported for the closed-form tests (rebuilder.cu:40-50).
void orc_calc_partition(uint32_t N, uint32_t* L, uint32_t* R) { calc_partition(N, *L, *R); }

// Number of levels of the build: log2i(count/2)+1 (rebuilder.cu:313).
uint32_t orc_nmh_levels(uint32_t count) { return orc::log2i(count / 2) + 1; }

// cudaRebuildNMH restated.  vertices: count triangles in "points" layout
// (x1 x2 x3 y1 y2 y3 z1 z2 z3).  indices_out[i] = original triangle index at
// tree position i.  threads > 1 runs the two stable sorts of every level on
// that many OpenMP threads (same result).  Returns 0 on success, -1 on
// invalid input, -2 if an internal consistency check fails (a split slot
// read that this level did not write, which the reference would read as
// stale memory).
int orc_nmh_build_mt(uint32_t count, const float* vertices, uint32_t* indices_out, int threads) {
  if (count < 2 || (count & 1u)) return -1;
  const uint32_t numLevels = orc::log2i(count / 2) + 1;
  const size_t maxSplits = size_t(1) << numLevels;

  std::vector<uint32_t> indices(count), nodes(count, 0u), perm(count), tmp(count);
  std::iota(indices.begin(), indices.end(), 0u);
  std::vector<Split> inSplits(2 * maxSplits), outSplits(2 * maxSplits);
  inSplits[0].index = 0;
  inSplits[0].size = count;

  // computeTriangleBounds (rebuilder.cu:72-91)
  std::vector<float> vmin[3], vmax[3];
  for (int a = 0; a < 3; ++a) {
    vmin[a].resize(count);
    vmax[a].resize(count);
  }
  for (uint32_t t = 0; t < count; ++t) {
    const float* v = vertices + size_t(t) * 9;
    for (int a = 0; a < 3; ++a) {
      vmin[a][t] = orc::fmin_nn(orc::fmin_nn(v[3 * a], v[3 * a + 1]), v[3 * a + 2]);
      vmax[a][t] = orc::fmax_nn(orc::fmax_nn(v[3 * a], v[3 * a + 1]), v[3 * a + 2]);
    }
  }

  int axis = 0;
  uint32_t pendingSplits = 1;
  uint32_t doneNodes = 0;
  for (uint32_t i = 0; i < numLevels; ++i) {
    const float* keysMin = vmin[axis].data();
    const float* keysMax = vmax[axis].data();
    sort_triangles(count, keysMin, indices, nodes, perm, tmp, threads);
    if (i == numLevels - 1) break;

    // swapMaxTriangleTPS (first maximum by position from slot +1); splits
    // are disjoint ranges, so they run in parallel with the same result
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) if (threads > 1)
    for (uint32_t j = 0; j < pendingSplits; ++j) {
      const Split sp = inSplits[j];
      float maxValue = keysMax[indices[sp.index + 1]];
      uint32_t maxIndex = sp.index + 1;
      for (uint32_t k = sp.index + 2; k < sp.index + sp.size; ++k) {
        const float value = keysMax[indices[k]];
        if (value > maxValue) {
          maxIndex = k;
          maxValue = value;
        }
      }
      if (maxIndex != sp.index + 1) {
        std::swap(indices[sp.index + 1], indices[maxIndex]);
        std::swap(nodes[sp.index + 1], nodes[maxIndex]);
      }
    }
    // updateNodesTPS
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) if (threads > 1)
    for (uint32_t j = 0; j < pendingSplits; ++j) {
      const Split sp = inSplits[j];
      const uint32_t nodeIndex = doneNodes + j;
      nodes[sp.index] = nodeIndex;
      nodes[sp.index + 1] = nodeIndex;
      for (uint32_t k = sp.index + 2; k < sp.index + sp.size; ++k) nodes[k] = 2 * nodeIndex + 1;
    }
    if (int(i) < std::max(0, int(numLevels) - 2)) {
      // emitSplitsKernel + copy_if over the first `generated` output slots
      for (auto& s : outSplits) s.written = false;
      uint32_t generated = 0;
      for (uint32_t j = 0; j < pendingSplits; ++j) {
        const Split sp = inSplits[j];
        if (sp.size <= 2) continue;
        uint32_t numL, numR;
        calc_partition(sp.size, numL, numR);
        const uint32_t pL = sp.index + 2 * (pendingSplits - j);
        const uint32_t pR = pL + numL;
        outSplits[2 * j].index = pL;
        outSplits[2 * j].size = numL;
        outSplits[2 * j].written = true;
        outSplits[2 * j + 1].index = pR;
        outSplits[2 * j + 1].size = numR;
        outSplits[2 * j + 1].written = true;
        generated += (numL > 0) + (numR > 0);
      }
      uint32_t copied = 0;
      for (uint32_t k = 0; k < generated; ++k) {
        if (!outSplits[k].written) return -2;
        if (outSplits[k].size > 0) inSplits[copied++] = outSplits[k];
      }
      if (copied != generated) return -2;
      doneNodes += pendingSplits;
      pendingSplits = generated;
    }
    axis = (axis + 1) % 3;
  }
  std::memcpy(indices_out, indices.data(), sizeof(uint32_t) * count);
  return 0;
}

int orc_nmh_build(uint32_t count, const float* vertices, uint32_t* indices_out) {
  return orc_nmh_build_mt(count, vertices, indices_out, 1);
}

}  // extern "C"
