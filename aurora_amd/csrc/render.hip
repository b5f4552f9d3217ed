// Frame rendering on gfx950: replaces MonteCarloRaytracer::render
// (src/render/monteCarloRaytracer.cpp:81-120) and the kernels it drives
// (cudaGenerateRays, utility.cu:18-94; cudaRaytraceMonteCarlo,
// monteCarloRaytracer.cu:20-124; cudaDrawPixels, utility.cu:199-229).
//
// AUR_MODE_COMPAT reproduces the reference's Monte Carlo ray tracer: each
// pixel sample walks the reference's ray stack (reflection and refraction
// only, MAX_DEPTH 8) with the reference's direct-light estimators
// (radiance.cuh:8-80, light.cuh:24-60), as wavefront passes: per stack
// level, closest hit -> shade (the estimators' draws and shadow rays, child
// rays) -> shadow (any hit) -> resolve (the estimators' sums in light and
// sample order).  cuRAND is replaced by Philox keyed by (seed, pixel,
// sample, draw index).
//
// AUR_MODE_PBR is a wavefront path loop: per sample pass, camera rays ->
// [extend (closest hit) -> shade (emission, one next-event slot per light
// sample, Russian roulette, cosine BSDF sample, compaction of surviving
// paths into the next queue) -> shadow (any hit) -> resolve (adds the
// unoccluded slots in light order)] x max_depth -> accumulate.  Float
// operations follow the same order as the scalar restatement, so every
// pixel replays exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "context.h"
#include "traverse.h"

namespace aur {
namespace {

constexpr int kBlock = 128;

struct SceneDev {
  const NodeRec* nodes;
  uint32_t n_nodes;
  uint32_t rec_bytes;
  const float* normals;
  const uint16_t* shader_ids;
  const aur_shader* shaders;
  uint32_t num_shaders;
  const aur_light* lights;
  uint32_t num_lights;
  const float* texcoords;
  const TexDesc* textures;
  const uint8_t* texels;
  uint32_t num_textures;
};

// Texture colour of a hit: texture sh.texture (1-based) at the barycentric
// interpolation of the triangle's uvs (getTexCoord, src/kernels/lib/common.cuh:45-51;
// texfetch, src/kernels/lib/texunit.cuh:54-94); white when no such texture
// exists (texfetch's default).
__device__ inline V3 tex_color(const float* texcoords, const TexDesc* textures, const uint8_t* texels,
                               uint32_t num_textures, uint32_t texture, uint32_t tri, float u, float v) {
  const uint32_t id = texture - 1;
  if (id >= num_textures) return mk(1.0f, 1.0f, 1.0f);
  const float* t = texcoords + (size_t)tri * 6;
  const float w = 1.0f - u - v;
  const float tu = (w * t[0] + u * t[2]) + v * t[4];
  const float tv = (w * t[1] + u * t[3]) + v * t[5];
  return tex_sample(textures[id], texels, tu, tv);
}

__device__ inline uint32_t shader_index(const SceneDev& sc, uint32_t tri) {
  uint32_t sid = sc.shader_ids[tri];
  sid = sid > 0 ? sid - 1 : 0;  // getSafeID (src/util/array.h:18)
  return sid < sc.num_shaders ? sid : 0;
}

__device__ inline const aur_shader& shader_of(const SceneDev& sc, uint32_t tri) {
  return sc.shaders[shader_index(sc, tri)];
}

__device__ inline V3 interp9(const float* base, uint32_t tri, float u, float v) {
  const float* p = base + (size_t)tri * 9;
  return bclerp(mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]), mk(p[6], p[7], p[8]), u, v);
}

// ---------------------------------------------------------------- compat --

struct CompatRng {
  uint32_t seed, pixel, sample, k;
  U4 cache;
  __device__ float next() {
    if ((k & 3u) == 0) cache = philox(U4{pixel, sample, 0x40000000u | (k >> 2), 0u}, seed, kKeyCompat);
    const uint32_t w = (k & 3u) == 0 ? cache.x : (k & 3u) == 1 ? cache.y : (k & 3u) == 2 ? cache.z : cache.w;
    ++k;
    return u01(w);
  }
};

__device__ inline V3 camera_dir_compat(const aur_camera& c, uint32_t w, uint32_t h, uint32_t x,
                                       uint32_t y, uint32_t sampleID, uint32_t samples) {
  float px, py;
  if (samples == 1) {
    px = float(x) / float(w) - 0.5f;
    py = float(y) / float(h) - 0.5f;
  } else {
    const short swidth = (short)sqrtf((float)(unsigned short)samples);
    const short sx = (short)(sampleID % swidth), sy = (short)(sampleID / swidth);
    const float dx = 1.0f / float(w), dy = 1.0f / float(h);
    const float sdx = dx / swidth, sdy = dy / swidth;
    px = x * dx - 0.5f + (sx * sdx);
    py = y * dy - 0.5f + (sy * sdy);
  }
  return normalize(px * ld3(c.right) * c.tanfov[0] + py * ld3(c.up) * c.tanfov[1] + ld3(c.forward));
}

struct CRay {
  V3 pos, dir;
  float t, weight, eta;
};

// The reference's ray stack (Stack<Ray, MAX_DEPTH + 2>, popped last-in
// first-out) without a dynamically indexed per-lane array: the ray on top of
// the stack is kept with the pixel's state, and a hit pushes a ray into
// memory only when it spawns two (mirror and refraction): the mirror ray,
// pushed first, waits in `pend` below the refraction ray, which becomes the
// top.  Each such hit raises depth, and a child is pushed only while depth
// <= 8, so at most kCompatPend rays wait.
constexpr int kCompatPend = 8;

// ------------------------------------------------------------------- PBR --

// Sample point of (pixel, sample) for dimension pair (bounce, slot): slot 0
// of bounce 0 = pixel position, slot 1 = BSDF direction (x, y) and Russian
// roulette (z), slots 16.. = light samples.
__device__ inline U4 pbr_draw(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                              uint32_t slot) {
  return sobol_owen(seed, pixel, sample, (bounce << 16) | slot);
}

// sin/cos of (pi/4) q, |q| <= 1, Taylor polynomials in Horner form.
__device__ inline void sincos_quarter(float q, float& s, float& c) {
  const float phi = 0.785398163f * q;
  const float z = phi * phi;
  s = phi * (1.0f + z * (-0.166666667f + z * (0.00833333333f + z * (-0.000198412698f + z * 2.75573192e-6f))));
  c = 1.0f + z * (-0.5f + z * (0.0416666667f + z * (-0.00138888889f + z * (2.48015873e-5f + z * -2.75573192e-7f))));
}

// Concentric map (Shirley & Chiu 1997) with the branch structure of
// src/kernels/lib/sampling.cuh:51-73.
__device__ inline void concentric_disk(float u1, float u2, float& dx, float& dy) {
  const float x = 2.0f * u1 - 1.0f;
  const float y = 2.0f * u2 - 1.0f;
  float r, s, c;
  if (x >= -y) {
    if (x > y) {
      r = x;
      sincos_quarter(y / x, s, c);
      dx = r * c;
      dy = r * s;
    } else if (y != 0.0f) {
      r = y;
      sincos_quarter(-x / y, s, c);
      dx = r * -s;
      dy = r * c;
    } else {
      // centre of the square: the reference's 0/0 would give NaN
      dx = 0.0f;
      dy = 0.0f;
    }
  } else {
    if (x < y) {
      r = -x;
      sincos_quarter(y / x, s, c);
      dx = r * -c;
      dy = r * -s;
    } else {
      r = -y;
      if (y != 0.0f) {
        sincos_quarter(-x / y, s, c);
        dx = r * s;
        dy = r * -c;
      } else {
        dx = r;
        dy = 0.0f;
      }
    }
  }
}

// Orthonormal basis (Duff et al., JCGT 2017).
__device__ inline void onb(V3 n, V3& t1, V3& t2) {
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  t1 = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  t2 = mk(b, sign + n.y * n.y * a, -n.y);
}

// Wavefront state shared by the PBR kernels.  A pass traces the global paths
// [g0, g0 + n), n = min(cap, end - g0), with g0 read from device memory so a
// captured pass replays unchanged; per-path arrays are indexed by q = g - g0.
// kMulti = false: one job, passed by value (wave-uniform parameters);
// kMulti = true: jobs[path_job[q]].
struct PbrArgs {
  JobDev job;
  const JobDev* jobs;
  uint32_t njobs;
  uint32_t* path_job;
  const unsigned long long* g0p;
  unsigned long long end;
  uint32_t cap, nslots, max_depth;
  uint32_t refill;  // persistent traversal: refill once this many lanes are idle
  // traversal stack: klds rows in LDS (all when gstack is null), the other
  // gstack_rows per lane in gstack
  float2* gstack;
  uint32_t klds, gstack_rows;
  // rays of bounce k in queue order at [(k & 1) * cap + queue position];
  // hits of the current bounce in queue order
  float4 *ray_o, *ray_d;
  float4 *beta, *radiance;
  aur_hit* hits;
  float4 *sh_o, *sh_d, *sh_c;
  int32_t* sh_occ;
  uint32_t* sh_queue;  // slot indices e = slot * cap + queue position with a shadow ray
  uint32_t* sh_job;    // batches: the job of each shadow-queue entry
  // queue lengths as three 64-bit words used in turn, word k % 3 for bounce
  // k: counters[2w] = paths entering bounce k, counters[2w + 1] = shadow rays
  // shading appended at bounce k - 1 (shading appends to both with one
  // atomic; three words keep bounce k - 1's counts for its resolve while
  // bounce k + 1's are filled)
  uint32_t* counters;
  unsigned long long* stats;
  // path-vertex records of an adjoint replay (null in a plain render),
  // indexed [bounce][q] and [bounce][slot][q] with stride cap: shader index of
  // the hit (-1 ends the path), Russian-roulette survival probability (1 when
  // not applied) and each next-event slot's weight w, zero when occluded, with
  // contribution = beta * albedo * light radiance * w.
  int32_t* rec_mat;
  float* rec_rr;
  float* rec_w;
};

__device__ __forceinline__ uint32_t pass_paths(const PbrArgs& a, unsigned long long g0) {
  const unsigned long long left = a.end > g0 ? a.end - g0 : 0ull;
  return left < a.cap ? (uint32_t)left : a.cap;
}

template <bool kMulti>
__device__ __forceinline__ const JobDev& job_at(const PbrArgs& a, uint32_t jb) {
  if constexpr (kMulti)
    return a.jobs[jb];
  else
    return a.job;
}

template <bool kMulti>
__device__ __forceinline__ const JobDev& job_of(const PbrArgs& a, uint32_t q) {
  if constexpr (kMulti)
    return a.jobs[a.path_job[q]];
  else
    return a.job;
}

// Frame pixel index of region pixel l.
__device__ __forceinline__ uint32_t frame_pixel(const JobDev& J, uint32_t l) {
  return (J.ry0 + l / J.rw) * J.w + J.rx0 + l % J.rw;
}

__device__ __forceinline__ uint32_t job_shader(const JobDev& J, uint32_t tri) {
  uint32_t sid = J.shader_ids[tri];
  sid = sid > 0 ? sid - 1 : 0;  // getSafeID (src/util/array.h:18)
  return sid < J.num_shaders ? sid : 0;
}

template <bool kMulti>
__global__ void k_pbr_camera(PbrArgs a, uint32_t* __restrict__ queue) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long g0 = *a.g0p;
  const uint32_t n = pass_paths(a, g0);
  if (q == 0) {
    a.counters[0] = n;
    a.counters[1] = 0;
  }
  if (q >= n) return;
  const unsigned long long g = g0 + q;
  uint32_t jb = 0;
  if constexpr (kMulti) {
    // last job whose range starts at or before g
    uint32_t lo = 0, hi = a.njobs;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.jobs[mid].path_start <= g) lo = mid; else hi = mid;
    }
    jb = lo;
    a.path_job[q] = jb;
  }
  const JobDev& J = job_at<kMulti>(a, jb);
  const uint32_t r = (uint32_t)(g - J.path_start);
  const uint32_t l = r % J.npix, smp = r / J.npix;
  const uint32_t x = J.rx0 + l % J.rw, y = J.ry0 + l / J.rw;
  const uint32_t p = y * J.w + x;  // frame pixel: the random-number key
  // the pixel's sample points are stratified for every power-of-two prefix
  const U4 rn = pbr_draw(J.seed, p, smp, 0, 0);
  const float fx = u01(rn.x), fy = u01(rn.y);
  const aur_camera& cam = J.cam;
  const float px = ((float)x + fx) / (float)J.w - 0.5f;
  const float py = ((float)y + fy) / (float)J.h - 0.5f;
  const V3 d = normalize(px * ld3(cam.right) * cam.tanfov[0] + py * ld3(cam.up) * cam.tanfov[1] +
                         ld3(cam.forward));
  // a batch carries each ray's job in origin.w (the trace needs no lookup)
  a.ray_o[q] = make_float4(cam.position[0], cam.position[1], cam.position[2],
                           kMulti ? __uint_as_float(jb) : 0.0f);
  a.ray_d[q] = make_float4(d.x, d.y, d.z, 0.0f);
  a.beta[q] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  a.radiance[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  queue[q] = q;
}

// One persistent launch for bounce k: closest hit for every path entering
// bounce k (rays in queue order from ray_o/ray_d + roff, items [0, n_ext);
// hits in queue order) and any hit for every shadow ray shading appended at
// bounce k - 1 (the dense shadow queue, items [n_ext, n_ext + n_sh)); both
// counts in word w = k % 3.  parts bit 0:
// extension rays, bit 1: shadow rays.  The two sets are independent, so one
// launch overlaps their tails.
template <bool kMulti, bool kSplit>
__global__ __launch_bounds__(kBlock) void k_pbr_trace(PbrArgs a, uint32_t w, uint32_t parts,
                                                      uint32_t roff) {
  extern __shared__ float2 stk_mem[];
  const auto stack = [&]() {
    if constexpr (kSplit) {
      const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
      const uint32_t lanes = gridDim.x * blockDim.x;
      return SplitStack{(lds_vf2*)(stk_mem + threadIdx.x), blockDim.x, (glb_vf2*)(a.gstack + gtid),
                        __builtin_amdgcn_make_buffer_rsrc(a.gstack, (short)0,
                                                          (int)(a.gstack_rows * lanes * 8u), 0x00020000),
                        gtid * 8u, lanes, a.klds};
    } else {
      return lds_stack(stk_mem + threadIdx.x, blockDim.x);
    }
  }();
  const uint32_t n_ext = (parts & 1u) ? a.counters[2 * w] : 0u;
  const uint32_t n_sh = (parts & 2u) ? a.counters[2 * w + 1] : 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the next word was last read by the resolve of bounce k - 2, which has
    // completed; the shade kernel that follows appends to it
    const uint32_t wn = w == 2u ? 0u : w + 1u;
    a.counters[2 * wn] = 0;
    a.counters[2 * wn + 1] = 0;
    atomicAdd(&a.stats[0], (unsigned long long)n_ext);
    atomicAdd(&a.stats[1], (unsigned long long)n_sh);
  }
  persistent_trace(
      TreeRef{a.job.nodes, a.job.n_nodes, a.job.root, a.job.rec_bytes}, (uint64_t)n_ext + n_sh,
      a.refill, stack,
      [&](uint64_t i, V3& o, V3& d, float& tmax, TreeRef& tree, bool& any) {
        float4 o4, d4;
        uint32_t jb = 0;
        if (i < n_ext) {
          o4 = a.ray_o[roff + i];
          d4 = a.ray_d[roff + i];
          jb = __float_as_uint(o4.w);
          o4.w = kInfinity;
          any = false;
        } else {
          const uint32_t e = a.sh_queue[i - n_ext];
          if constexpr (kMulti) jb = a.sh_job[i - n_ext];
          o4 = a.sh_o[e];
          d4 = a.sh_d[e];
          any = true;
        }
        if constexpr (kMulti) {
          const JobDev& J = a.jobs[jb];
          tree = TreeRef{J.nodes, J.n_nodes, J.root, J.rec_bytes};
        }
        o = mk(o4.x, o4.y, o4.z);
        d = mk(d4.x, d4.y, d4.z);
        tmax = o4.w;
        return true;
      },
      [&](uint64_t i, const Trav& st) {
        if (i < n_ext)
          a.hits[i] = aur_hit{st.id, st.t, st.u, st.v};
        else
          a.sh_occ[a.sh_queue[i - n_ext]] = st.id >= 0 ? 1 : 0;
      });
}

// What shading one path vertex leaves to the queues.
struct ShadeOut {
  bool survive;      // the path continues with ray (o, d)
  uint64_t shadows;  // bit e: next-event slot e holds a shadow ray
  float4 o, d;
};

// Shades queue entry j (path p of job J): emission, next-event slots,
// Russian roulette and the next direction.
template <bool kMulti, bool kRecord>
__device__ __forceinline__ ShadeOut shade_path(const PbrArgs& a, const JobDev& J, uint32_t j,
                                               uint32_t p, uint32_t bounce, uint32_t roff) {
  ShadeOut out{false, 0ull, make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
  const uint32_t r0 = (uint32_t)(*a.g0p + p - J.path_start);
  const uint32_t pix = frame_pixel(J, r0 % J.npix), smp = r0 / J.npix;
  const size_t base = j, ss = a.cap;  // slot e of queue entry j at base + e * ss
  const size_t np = a.cap;
  const aur_hit hit = a.hits[j];
  if (hit.id < 0) {
    for (uint32_t s = 0; s < a.nslots; ++s) a.sh_d[base + s * ss].w = 0.0f;
    if (kRecord) a.rec_mat[bounce * np + p] = -1;
    return out;
  }
  const uint32_t sid = job_shader(J, (uint32_t)hit.id);
  if (kRecord) a.rec_mat[bounce * np + p] = (int32_t)sid;
  const float4 o4 = a.ray_o[roff + j], d4 = a.ray_d[roff + j], b4 = a.beta[p];
  float4 L4 = a.radiance[p];
  const V3 O = mk(o4.x, o4.y, o4.z), D = mk(d4.x, d4.y, d4.z);
  V3 beta = mk(b4.x, b4.y, b4.z);
  const V3 P = O + (D * hit.t);
  const V3 wo = -D;
  const aur_shader& sh = J.shaders[sid];
  V3 N = normalize(interp9(J.normals, (uint32_t)hit.id, hit.u, hit.v));
  if (dot(N, wo) < 0.0f) N = -N;
  V3 albedo = ld3(sh.diffuse_color) * sh.diffuse;
  if (sh.texture > 0)
    albedo = tex_color(J.texcoords, J.textures, J.texels, J.num_textures, sh.texture, (uint32_t)hit.id, hit.u,
                       hit.v) * sh.diffuse;
  const V3 f = albedo * kInvPi;
  V3 L = mk(L4.x, L4.y, L4.z) + beta * ld3(sh.emission_color);
  a.radiance[p] = make_float4(L.x, L.y, L.z, bounce == 0 ? 1.0f : L4.w);
  const V3 bf = beta * f;
  uint32_t slot = 16, e = 0;
  for (uint32_t li = 0; li < J.num_lights; ++li) {
    const aur_light& Lt = J.lights[li];
    const V3 CI = ld3(Lt.color) * Lt.intensity;
    if (Lt.type == AUR_LIGHT_AREA) {
      const int ns = Lt.samples > 0 ? Lt.samples : 1;
      const float inv = 1.0f / (float)ns;
      for (int k = 0; k < ns; ++k, ++e) {
        const U4 r = pbr_draw(J.seed, pix, smp, bounce, slot++);
        const V3 Q = ld3(Lt.position) + ((ld3(Lt.e1) * u01(r.x)) + (ld3(Lt.e2) * u01(r.y)));
        V3 wi = Q - P;
        const float dist = length(wi);
        wi = wi / dist;
        const float cosL = dot(ld3(Lt.direction), -wi);
        const float cosS = dot(N, wi);
        float flag = 0.0f;
        if (cosL > 0.0f && cosS > 0.0f) {
          const float wgt = (cosS * cosL * Lt.area) / (dist * dist) * inv;
          const V3 c = bf * (CI * wgt);
          const V3 so = P + (kEpsilon * wi);
          a.sh_o[base + e * ss] = make_float4(so.x, so.y, so.z, dist + kEpsilon);
          a.sh_c[base + e * ss] = make_float4(c.x, c.y, c.z, wgt * kInvPi);
          flag = 1.0f;
          out.shadows |= 1ull << e;
        }
        a.sh_d[base + e * ss] = make_float4(wi.x, wi.y, wi.z, flag);
      }
    } else if (Lt.type == AUR_LIGHT_POINT) {
      V3 wi = ld3(Lt.position) - P;
      const float dist = length(wi);
      wi = wi / dist;
      const float cosS = dot(N, wi);
      float flag = 0.0f;
      if (cosS > 0.0f) {
        const V3 c = bf * ((CI / (dist * dist)) * cosS);
        a.sh_o[base + e * ss] = make_float4(P.x, P.y, P.z, dist);
        a.sh_c[base + e * ss] = make_float4(c.x, c.y, c.z, cosS / (dist * dist) * kInvPi);
        flag = 1.0f;
        out.shadows |= 1ull << e;
      }
      a.sh_d[base + e * ss] = make_float4(wi.x, wi.y, wi.z, flag);
      ++e;
    } else if (Lt.type == AUR_LIGHT_DIRECTIONAL) {
      const V3 wi = -ld3(Lt.direction);
      const float cosS = dot(N, wi);
      float flag = 0.0f;
      if (cosS > 0.0f) {
        const V3 c = bf * (CI * cosS);
        a.sh_o[base + e * ss] = make_float4(P.x, P.y, P.z, kInfinity);
        a.sh_c[base + e * ss] = make_float4(c.x, c.y, c.z, cosS * kInvPi);
        flag = 1.0f;
        out.shadows |= 1ull << e;
      }
      a.sh_d[base + e * ss] = make_float4(wi.x, wi.y, wi.z, flag);
      ++e;
    } else {
      // ambient: the fraction `ambient` lights unshadowed; the rest comes from
      // the light's position with no falloff and a shadow ray (compat mode's
      // estimateDirectRadianceDelta, radiance.cuh:13-28)
      const V3 c = bf * (CI * Lt.ambient);
      a.sh_c[base + e * ss] = make_float4(c.x, c.y, c.z, Lt.ambient * kInvPi);
      a.sh_d[base + e * ss] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);  // no shadow ray
      ++e;
      V3 wi = ld3(Lt.position) - P;
      const float dist = length(wi);
      wi = wi / dist;
      const float cosS = dot(N, wi);
      const float rest = 1.0f - Lt.ambient;
      float flag = 0.0f;
      if (cosS > 0.0f && rest != 0.0f) {
        const V3 c2 = bf * ((CI * rest) * cosS);
        a.sh_o[base + e * ss] = make_float4(P.x, P.y, P.z, dist);
        a.sh_c[base + e * ss] = make_float4(c2.x, c2.y, c2.z, rest * cosS * kInvPi);
        flag = 1.0f;
        out.shadows |= 1ull << e;
      }
      a.sh_d[base + e * ss] = make_float4(wi.x, wi.y, wi.z, flag);
      ++e;
    }
  }
  if constexpr (kMulti) {
    for (; e < a.nslots; ++e) a.sh_d[base + e * ss].w = 0.0f;  // slots of jobs with more lights
  }
  if (bounce + 1 >= J.max_depth) return out;
  const U4 r = pbr_draw(J.seed, pix, smp, bounce, 1);
  if (kRecord) a.rec_rr[bounce * np + p] = 1.0f;
  if (bounce >= J.rr_start) {
    const float q = fminf(0.95f, fmaxf(beta.x, fmaxf(beta.y, beta.z)));
    if (u01(r.z) >= q) {
      if (kRecord) a.rec_mat[(bounce + 1) * np + p] = -1;
      return out;
    }
    beta = beta / q;
    if (kRecord) a.rec_rr[bounce * np + p] = q;
  }
  float dx, dy;
  concentric_disk(u01(r.x), u01(r.y), dx, dy);
  const float dz = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
  V3 t1, t2;
  onb(N, t1, t2);
  const V3 wi = t1 * dx + t2 * dy + N * dz;
  beta = beta * albedo;
  out.o = make_float4(P.x, P.y, P.z, 0.0f);
  out.d = make_float4(wi.x, wi.y, wi.z, 0.0f);
  a.beta[p] = make_float4(beta.x, beta.y, beta.z, 0.0f);
  out.survive = true;
  return out;
}

// Shading of every queued path, then the appends with one atomic per wave,
// made where every lane of the wave is active: the next bounce's path queue
// (with its rays, in queue order) and the shadow queue (slot-major within
// the wave, so neighbouring items trace to the same light).
template <bool kMulti, bool kRecord>
__global__ __launch_bounds__(128) void k_pbr_shade(PbrArgs a, uint32_t w, uint32_t bounce,
                                                   const uint32_t* __restrict__ qin,
                                                   uint32_t* __restrict__ qout,
                                                   const JobDev* __restrict__ jobs) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n_in = a.counters[2 * w];
  const uint32_t roff = (bounce & 1u) * a.cap, roff_next = a.cap - roff;
  ShadeOut out{false, 0ull, make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
  uint32_t p = 0, jb = 0;
  if constexpr (kMulti) {
    // one round per distinct job among the wave's paths (almost always one:
    // paths are numbered job by job), each reading its job record at a
    // wave-uniform address; a ray carries its job in origin.w
    const bool act = j < n_in;
    if (act) {
      p = qin[j];
      jb = __float_as_uint(a.ray_o[roff + j].w);
    }
    uint64_t todo = __ballot(act);
    while (todo) {
      const uint32_t jbu = (uint32_t)__builtin_amdgcn_readlane((int)jb, (int)(__ffsll((unsigned long long)todo) - 1));
      const bool mine = act && jb == jbu;
      // read where the whole wave is active, through the read-only kernel
      // argument: a wave-uniform record comes through scalar loads
      const JobDev J = jobs[jbu];
      if (mine) out = shade_path<kMulti, kRecord>(a, J, j, p, bounce, roff);
      todo &= ~__ballot(mine);
    }
  } else {
    if (j < n_in) {
      p = qin[j];  // path index in the pass
      out = shade_path<kMulti, kRecord>(a, a.job, j, p, bounce, roff);
    }
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t nshadow = 0;
  for (uint32_t sl = 0; sl < a.nslots; ++sl) nshadow += (uint32_t)__popcll(__ballot((out.shadows >> sl) & 1ull));
  const uint64_t ms = __ballot(out.survive);
  if (nshadow == 0 && ms == 0) return;
  unsigned long long old = 0;
  if (lane == 0)
    old = atomicAdd(reinterpret_cast<unsigned long long*>(a.counters + 2 * (w == 2u ? 0u : w + 1u)),
                    (unsigned long long)__popcll(ms) | ((unsigned long long)nshadow << 32));
  old = __shfl(old, 0);
  if (out.survive) {
    const uint32_t pos = (uint32_t)old + (uint32_t)__popcll(ms & below);
    qout[pos] = p;
    if constexpr (kMulti) out.o.w = __uint_as_float(jb);  // the ray's job, for the trace
    a.ray_o[roff_next + pos] = out.o;
    a.ray_d[roff_next + pos] = out.d;
  }
  uint32_t qbase = (uint32_t)(old >> 32);
  for (uint32_t sl = 0; sl < a.nslots; ++sl) {
    const bool mine = (out.shadows >> sl) & 1ull;
    const uint64_t m = __ballot(mine);
    if (mine) {
      const uint32_t qpos = qbase + (uint32_t)__popcll(m & below);
      a.sh_queue[qpos] = sl * a.cap + j;
      if constexpr (kMulti) a.sh_job[qpos] = jb;
    }
    qbase += (uint32_t)__popcll(m);
  }
}

// Adds the unoccluded next-event contributions of bounce `bounce` (queue qin,
// count in word w).  Paths without a hit at this bounce have every slot
// flag 0 (hits[] may already hold the next bounce's result).
template <bool kRecord>
__global__ void k_pbr_resolve(PbrArgs a, uint32_t w, uint32_t bounce, const uint32_t* __restrict__ qin) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.counters[2 * w]) return;
  const uint32_t p = qin[j];
  const size_t np = a.cap;
  float4 L4 = a.radiance[p];
  V3 L = mk(L4.x, L4.y, L4.z);
  for (uint32_t s = 0; s < a.nslots; ++s) {
    const size_t e = (size_t)s * a.cap + j;
    const float flag = a.sh_d[e].w;
    const bool lit = flag == 2.0f || (flag == 1.0f && a.sh_occ[e] == 0);
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (lit) {
      c = a.sh_c[e];
      L = L + mk(c.x, c.y, c.z);
    }
    if (kRecord) a.rec_w[((size_t)bounce * a.nslots + s) * np + p] = c.w;
  }
  a.radiance[p] = make_float4(L.x, L.y, L.z, L4.w);
}

// Adds the pass's samples to each pixel in sample order, so the sum does not
// depend on how paths were grouped into passes.  The thread of a pixel's
// first path in the pass walks that pixel's later samples (stride npix).
template <bool kMulti>
__global__ void k_pbr_accumulate(PbrArgs a) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long g0 = *a.g0p;
  const uint32_t n = pass_paths(a, g0);
  if (q >= n) return;
  const JobDev& J = job_of<kMulti>(a, q);
  const uint32_t r = (uint32_t)(g0 + q - J.path_start);
  if (q >= J.npix && r >= J.npix) return;  // the pixel's previous sample is in this pass
  const uint32_t pix = frame_pixel(J, r % J.npix);
  const float weight = 1.0f / (float)J.spp;
  float4 px = J.pixels[pix];
  for (uint32_t qq = q, rr = r; qq < n && rr < J.paths; qq += J.npix, rr += J.npix) {
    const float4 L = a.radiance[qq];
    if (L.w != 0.0f) {
      px.x += L.x * weight;
      px.y += L.y * weight;
      px.z += L.z * weight;
      px.w += weight;
    }
  }
  J.pixels[pix] = px;
}

// Next pass: g0 += cap (a kernel, so that a captured pass replays).
__global__ void k_pbr_advance(unsigned long long* g0p, uint32_t cap) {
  if (threadIdx.x == 0) *g0p += cap;
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0) *p = v;
}

// Clears one accumulation buffer of n pixels.
__global__ void k_clear_pixels(float4* __restrict__ px, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) px[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Clears each job's accumulation buffer (block b: jobs b, b + grid, ...).
__global__ void k_clear_jobs(const JobDev* __restrict__ jobs, uint32_t njobs) {
  for (uint32_t jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
    float4* px = jobs[jb].pixels;
    const uint32_t n = jobs[jb].npix;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) px[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ---------------------------------------------------------------- adjoint --
//
// Gradient of a scalar image loss with respect to each shader's albedo
// (diffuse_color * diffuse) and each light's radiance (color * intensity),
// given dLoss/dI per pixel.  For one path with vertices b = 0..n-1 (shader
// m_b, albedo rho_b, emission e_b, Russian-roulette probability q_b, and
// next-event weight W_{b,l} per light l), the estimate is
//   sum_b beta_b (e_b + rho_b D_b),   D_b = sum_l CI_l W_{b,l},
//   beta_0 = 1,  beta_{b+1} = beta_b rho_b / q_b
// (cosine-weighted BSDF sampling makes the continuation weight rho_b).  With
// the suffix G_b = e_b + rho_b R_b, R_b = D_b + G_{b+1} / q_b, the
// derivatives are d/d rho_b = beta_b R_b and d/d CI_l = sum_b beta_b rho_b
// W_{b,l}; the RR probabilities are held fixed.  The sweep visits vertices
// forward for beta and the light terms, then backward for R.

// Per-block LDS accumulation is exact integer arithmetic on contributions
// quantised to 2^-28 of the largest per-sample pixel adjoint, so the sum does
// not depend on the order in which lanes add.
constexpr float kAdjQuant = 268435456.0f;  // 2^28

__global__ void k_adj_absmax(const float4* __restrict__ dimg, uint32_t npix, uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const float4 g = dimg[p];
    const float a = fmaxf(fabsf(g.x), fmaxf(fabsf(g.y), fabsf(g.z)));
    m = max(m, __float_as_uint(a));  // non-negative floats order as integers
  }
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
  if ((threadIdx.x & 63u) == 0 && m) atomicMax(out, m);
}

__device__ inline void adj_add(unsigned long long* acc, uint32_t i, float v, float scale) {
  if (v == 0.0f) return;
  const float x = fminf(fmaxf(v * scale, -4.0e18f), 4.0e18f);
  atomicAdd(acc + i, (unsigned long long)__float2ll_rn(x));
}

// One job (the frame); block partials[block][param] as doubles.
template <int kMax>
__global__ __launch_bounds__(256) void k_adj_sweep(PbrArgs a, const float4* __restrict__ dimg,
                                                   const uint32_t* __restrict__ absmax,
                                                   double* __restrict__ partials, uint32_t nparams) {
  extern __shared__ unsigned long long acc[];
  for (uint32_t i = threadIdx.x; i < nparams; i += blockDim.x) acc[i] = 0ull;
  __syncthreads();
  const JobDev& J = a.job;
  const float dmax = __uint_as_float(*absmax);
  const float scale = dmax > 0.0f ? kAdjQuant * (float)J.spp / dmax : 0.0f;
  const float inv_spp = 1.0f / (float)J.spp;
  const unsigned long long g0 = *a.g0p;
  const uint32_t np = pass_paths(a, g0);
  const size_t stride = a.cap;
  const uint32_t lbase = 3u * J.num_shaders;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; scale > 0.0f && q < np;
       q += gridDim.x * blockDim.x) {
    const uint32_t pix = frame_pixel(J, (uint32_t)(g0 + q - J.path_start) % J.npix);
    const float4 g = dimg[pix];
    const V3 dl = mk(g.x * inv_spp, g.y * inv_spp, g.z * inv_spp);
    if (dl.x == 0.0f && dl.y == 0.0f && dl.z == 0.0f) continue;
    V3 beta = mk(1.0f, 1.0f, 1.0f);
    V3 betas[kMax], D[kMax];
    int32_t mats[kMax];
    float rr[kMax];
    uint32_t n = 0;
#pragma unroll
    for (int b = 0; b < kMax; ++b) {
      if ((uint32_t)b < J.max_depth && n == (uint32_t)b) {
        const int32_t m = a.rec_mat[(size_t)b * stride + q];
        if (m >= 0) {
          const aur_shader& sh = J.shaders[m];
          const V3 rho = ld3(sh.diffuse_color) * sh.diffuse;
          V3 d = mk(0.0f, 0.0f, 0.0f);
          uint32_t s = 0;
          for (uint32_t li = 0; li < J.num_lights; ++li) {
            const aur_light& Lt = J.lights[li];
            const uint32_t ns = light_slots(Lt);
            float W = 0.0f;
            for (uint32_t k = 0; k < ns; ++k, ++s) W += a.rec_w[((size_t)b * a.nslots + s) * stride + q];
            if (W != 0.0f) {
              const V3 CI = ld3(Lt.color) * Lt.intensity;
              d = d + CI * W;
              const V3 gl = dl * beta * rho * W;
              adj_add(acc, lbase + 3 * li + 0, gl.x, scale);
              adj_add(acc, lbase + 3 * li + 1, gl.y, scale);
              adj_add(acc, lbase + 3 * li + 2, gl.z, scale);
            }
          }
          betas[b] = beta;
          D[b] = d;
          mats[b] = m;
          const bool more = (uint32_t)b + 1 < J.max_depth;
          rr[b] = more ? a.rec_rr[(size_t)b * stride + q] : 1.0f;
          beta = (beta / rr[b]) * rho;
          n = b + 1;
        }
      }
    }
    V3 G = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int b = kMax - 1; b >= 0; --b) {
      if ((uint32_t)b < n) {
        const aur_shader& sh = J.shaders[mats[b]];
        const V3 rho = ld3(sh.diffuse_color) * sh.diffuse;
        const V3 R = D[b] + G / rr[b];
        const V3 ga = dl * betas[b] * R;
        adj_add(acc, 3 * mats[b] + 0, ga.x, scale);
        adj_add(acc, 3 * mats[b] + 1, ga.y, scale);
        adj_add(acc, 3 * mats[b] + 2, ga.z, scale);
        G = ld3(sh.emission_color) + rho * R;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nparams; i += blockDim.x)
    partials[(size_t)blockIdx.x * nparams + i] = scale > 0.0f ? (double)(long long)acc[i] / (double)scale : 0.0;
}

// total[i] += sum over blocks of partials[block][i], blocks in order.
__global__ void k_adj_reduce(const double* __restrict__ partials, uint32_t nblocks, uint32_t nparams,
                             double* __restrict__ total) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nparams) return;
  double s = 0.0;
  for (uint32_t b = 0; b < nblocks; ++b) s += partials[(size_t)b * nparams + i];
  total[i] += s;
}

__global__ void k_adj_finish(const double* __restrict__ total, uint32_t nshader3, uint32_t nlight3,
                             float* __restrict__ g_albedo, float* __restrict__ g_light) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nshader3) g_albedo[i] = (float)total[i];
  else if (i < nshader3 + nlight3 && g_light) g_light[i - nshader3] = (float)total[i];
}

__global__ void k_zero_f64(double* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_adj_zero_u32(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0u;
}

inline uint32_t nblocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Resident grid of a persistent kernel: occupancy-limited blocks on every CU,
// but no more waves than max_items/64.
uint32_t resident_grid(aur_context* ctx, const void* fn, size_t lds, uint64_t max_items) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  uint64_t blocks = (uint64_t)per_cu * ctx->num_cus;
  const uint64_t need = (max_items + kBlock - 1) / kBlock;
  if (blocks > need) blocks = need;
  return (uint32_t)(blocks ? blocks : 1);
}

// ------------------------------------------------------- compat wavefront --
//
// Compat mode as wavefront passes over rounds of (pixel, sample) entries:
// k_cw_init sets up each entry's stack with its camera ray; then per stack
// level a closest-hit launch over the entries still active, k_cw_shade (one
// ray of each entry's stack: the direct-light estimators with their random
// draws, shadow rays appended to a queue, child rays pushed), an any-hit
// launch over the shadow queue and k_cw_resolve (the estimators' sums in
// light and sample order from the shadow results, colour += Li * weight,
// then the entry continues or keeps its final colour); k_cw_accumulate then
// adds each pixel's samples in sample order, as the reference's launch per
// sample does.  Per entry, the rays, draws and sums are those of the
// reference's loop over its stack (monteCarloRaytracer.cu:54-85,
// radiance.cuh:13-80) in the same order; the persistent traversal launches
// refill lanes between node visits.
struct CwArgs {
  SceneDev sc;
  aur_camera cam;
  uint32_t w, h, spp, seed;
  uint4 region;     // (left, bottom, region width, region pixels)
  uint32_t l0;      // first region pixel of the chunk
  uint32_t nchunk;  // pixels in the chunk
  uint32_t s0;      // first sample of the round: entry l is chunk pixel l % nchunk, sample s0 + l / nchunk
  uint32_t n;       // entries of the round
  uint32_t stride;  // entry capacity: stride of the per-slot arrays
  uint32_t nslots;  // shadow slots per hit
  float4 *st_o, *st_d, *st_c;
  uint4* st_i;
  float4* pend;  // rays waiting below the top: [row][chunk pixel], two float4 each
  float4 *emw, *c1, *contrib;
  float4* fin;  // per entry: final colour, 1 when a ray of the stack hit (else 0)
  float4* pixels;
};

constexpr uint32_t kCwHit = 1u, kCwMore = 2u;  // c1.w flags: the ray hit; the pixel continues

// Exclusive prefix of n over the wave's lanes plus one atomic per wave on
// *counter: the position of this lane's first of n appended entries.  Every
// lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, uint32_t n) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t incl = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += t;
  }
  const uint32_t total = __shfl(incl, 63);
  uint32_t base = 0;
  if (lane == 0 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 0);
  return base + incl - n;
}

__device__ __forceinline__ CompatRng compat_rng(uint32_t seed, uint32_t pix, uint32_t sample, uint32_t k) {
  CompatRng r{seed, pix, sample, k, U4{0, 0, 0, 0}};
  if (k & 3u) r.cache = philox(U4{pix, sample, 0x40000000u | (k >> 2), 0u}, seed, kKeyCompat);
  return r;
}

__global__ void k_cw_init(CwArgs c, PbrArgs a, uint32_t* __restrict__ qin) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l == 0) {
    a.counters[0] = c.n;
    a.counters[1] = 0;
  }
  if (l >= c.n) return;
  const uint32_t rl = c.l0 + l % c.nchunk;
  const uint32_t x = c.region.x + rl % c.region.z, y = c.region.y + rl / c.region.z;
  const V3 o = ld3(c.cam.position);
  const V3 d = camera_dir_compat(c.cam, c.w, c.h, x, y, c.s0 + l / c.nchunk, c.spp);
  c.st_o[l] = make_float4(o.x, o.y, o.z, 1.0f);
  c.st_d[l] = make_float4(d.x, d.y, d.z, 1.0f);
  c.st_c[l] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
  c.st_i[l] = make_uint4(0u, 0u, 0u, y * c.w + x);
  a.ray_o[l] = make_float4(o.x, o.y, o.z, 0.0f);
  a.ray_d[l] = make_float4(d.x, d.y, d.z, 0.0f);
  qin[l] = l;
}

// One ray of each active entry's stack, up to the shadow rays; the entries
// whose stacks go on enter the next level's queue with their next ray.
// Queue counts as in the PBR passes: word w holds this level's entries;
// shadow rays and the next level's entries are appended to word wn.
__global__ __launch_bounds__(kBlock) void k_cw_shade(CwArgs c, PbrArgs a, uint32_t w, uint32_t wn,
                                                     const uint32_t* __restrict__ qin,
                                                     uint32_t* __restrict__ qout, uint32_t roff_next) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = a.counters[2 * w];
  const bool act = i < n;
  const int kMaxDepth = 8;
  uint32_t ncast = 0;
  bool more = false;
  uint32_t l = 0;
  CRay next{};
  if (act) {
    l = qin[i];
    const float4 so = c.st_o[l], sd = c.st_d[l], sc4 = c.st_c[l];
    uint4 si = c.st_i[l];
    int depth = (int)si.x, np = (int)si.y;
    const uint32_t pix = si.w;
    int32_t hpId = __float_as_int(sc4.w);
    float4* mine = c.pend + 2 * (size_t)l;
    const size_t row = 2 * (size_t)c.stride;
    CRay top{mk(so.x, so.y, so.z), mk(sd.x, sd.y, sd.z), kInfinity, so.w, sd.w};
    auto pop_pending = [&]() {
      if (np == 0) return false;
      --np;
      const float4 pa = mine[np * row], pb = mine[np * row + 1];
      top = CRay{mk(pa.x, pa.y, pa.z), mk(pb.x, pb.y, pb.z), kInfinity, pa.w, pb.w};
      return true;
    };
    CRay ray = top;
    bool have = false;
    uint32_t flags = 0;
    float4 c1v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const aur_hit hr = a.hits[i];
    if (hr.id < 0) {
      have = pop_pending();
    } else {
      flags = kCwHit;
      ray.t = hr.t;
      hpId = hr.id;
      const V3 position = ray.pos + (ray.dir * ray.t);
      const V3 wo = -ray.dir;
      const aur_shader& sh = shader_of(c.sc, (uint32_t)hpId);
      const V3 N = normalize(interp9(c.sc.normals, hpId, hr.u, hr.v));
      V3 color1 = ld3(sh.diffuse_color) * sh.diffuse;
      if (sh.texture > 0)
        color1 = tex_color(c.sc.texcoords, c.sc.textures, c.sc.texels, c.sc.num_textures, sh.texture,
                           (uint32_t)hpId, hr.u, hr.v) * sh.diffuse;
      const V3 f = color1 * kInvPi;
      CompatRng rng = compat_rng(c.seed, pix, c.s0 + l / c.nchunk, si.z);
      uint32_t slot = 0;
      // a shadow ray of slot `slot` (not queued yet) and its contribution
      // when unoccluded; a slot without a ray contributes zero
      const auto put = [&](bool cast, V3 o, V3 d, float tmax, V3 val) {
        const size_t e = (size_t)slot * c.stride + i;
        c.contrib[e] = make_float4(val.x, val.y, val.z, cast ? 1.0f : 0.0f);
        if (cast) {
          a.sh_o[e] = make_float4(o.x, o.y, o.z, tmax);
          a.sh_d[e] = make_float4(d.x, d.y, d.z, 0.0f);
          ++ncast;
        }
        ++slot;
      };
      for (uint32_t li = 0; li < c.sc.num_lights; ++li) {
        const aur_light& L = c.sc.lights[li];
        const V3 CI = ld3(L.color) * L.intensity;
        if (L.type != AUR_LIGHT_AREA) {
          // estimateDirectRadianceDelta (radiance.cuh:13-28)
          V3 wdir;
          float wt;
          V3 Lv;
          if (L.type == AUR_LIGHT_AMBIENT) {
            wdir = ld3(L.position) - position;
            wt = length(wdir);
            wdir = wdir / wt;
            Lv = CI * (1.0f - L.ambient);
          } else if (L.type == AUR_LIGHT_POINT) {
            wdir = ld3(L.position) - position;
            wt = length(wdir);
            wdir = wdir / wt;
            Lv = CI / (wt * wt);
          } else {
            wdir = -ld3(L.direction);
            wt = kInfinity;
            Lv = CI;
          }
          const bool cast = !is_zero(Lv) && !is_zero(f);
          put(cast, position, wdir, wt, cast ? f * Lv * fmaxf(0.0f, dot(wdir, N)) : mk(0.0f));
        } else {
          // estimateDirectRadiance (radiance.cuh:30-80)
          for (int s = 0; s < L.samples; ++s) {
            const float u1 = rng.next();
            const float u2 = rng.next();
            const V3 Pl = ld3(L.position) + ((ld3(L.e1) * u1) + (ld3(L.e2) * u2));
            V3 wdir = Pl - position;
            float wt = length(wdir);
            wdir = wdir / wt;
            const float lightPdf = 1.0f / fabsf(dot(ld3(L.direction), -wdir));
            const V3 Lv = CI * fmaxf(0.0f, dot(ld3(L.direction), -wdir));
            V3 wpos = position;
            wpos.x += kEpsilon * wdir.x;
            wpos.y += kEpsilon * wdir.y;
            wpos.z += kEpsilon * wdir.z;
            wt += kEpsilon;
            const bool cast = lightPdf > 0.0f && !is_zero(Lv) && !is_zero(f);
            put(cast, wpos, wdir, wt, cast ? f * Lv * fmaxf(0.0f, dot(wdir, N)) : mk(0.0f));
          }
        }
      }
      si.z = rng.k;
      if (sh.reflectivity > 0.0f || sh.translucence > 0.0f) depth++;
      if (depth <= kMaxDepth && sh.reflectivity > 0.0f) {
        top = CRay{position, reflect(ray.dir, N), kInfinity, sh.reflectivity, ray.eta};
        ray.weight *= (1.0f - sh.reflectivity);
        have = true;
      }
      if (depth <= kMaxDepth && sh.translucence > 0.0f) {
        V3 rN = N;
        float etai = ray.eta, etat = sh.refractive_index;
        if (dot(rN, wo) < 0.0f) {
          rN = -rN;
          etai = sh.refractive_index;
          etat = 1.0f;
        }
        V3 tdir;
        if (refract(tdir, ray.dir, rN, etai, etat)) {
          if (have) {  // the mirror ray waits below the refraction ray
            mine[np * row] = make_float4(top.pos.x, top.pos.y, top.pos.z, top.weight);
            mine[np * row + 1] = make_float4(top.dir.x, top.dir.y, top.dir.z, top.eta);
            ++np;
          }
          top = CRay{position, tdir, kInfinity, sh.translucence, etat};
          ray.weight *= (1.0f - sh.translucence);
          have = true;
        }
      }
      if (!have) have = pop_pending();
      const V3 em = ld3(sh.emission_color);
      c.emw[i] = make_float4(em.x, em.y, em.z, ray.weight);
      c1v = make_float4(color1.x, color1.y, color1.z, 0.0f);
    }
    more = depth <= kMaxDepth && have;
    if (more) flags |= kCwMore;
    next = top;
    c1v.w = __uint_as_float(flags);
    c.c1[i] = c1v;
    c.st_o[l] = make_float4(top.pos.x, top.pos.y, top.pos.z, top.weight);
    c.st_d[l] = make_float4(top.dir.x, top.dir.y, top.dir.z, top.eta);
    c.st_c[l].w = __int_as_float(hpId);
    si.x = (uint32_t)depth;
    si.y = (uint32_t)np;
    c.st_i[l] = si;
  }
  // the cast slots enter the shadow queue slot by slot, each slot's rays in
  // lane order (neighbouring pixels' rays toward the same light sample stay
  // together for the traversal); one atomic per wave
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t total = 0;
  for (uint32_t slot = 0; slot < c.nslots; ++slot)
    total += (uint32_t)__popcll(__ballot(act && ncast && c.contrib[(size_t)slot * c.stride + i].w != 0.0f));
  uint32_t run = 0;
  if (lane == 0 && total) run = atomicAdd(&a.counters[2 * wn + 1], total);
  run = __shfl(run, 0);
  for (uint32_t slot = 0; slot < c.nslots && total; ++slot) {
    const size_t e = (size_t)slot * c.stride + i;
    const bool cast = act && ncast && c.contrib[e].w != 0.0f;
    const uint64_t m = __ballot(cast);
    if (cast) a.sh_queue[run + (uint32_t)__popcll(m & lt)] = (uint32_t)e;
    run += (uint32_t)__popcll(m);
  }
  const uint32_t p = wave_append(&a.counters[2 * wn], more ? 1u : 0u);
  if (more) {
    qout[p] = l;
    a.ray_o[roff_next + p] = make_float4(next.pos.x, next.pos.y, next.pos.z, 0.0f);
    a.ray_d[roff_next + p] = make_float4(next.dir.x, next.dir.y, next.dir.z, 0.0f);
  }
}

// The estimators' sums from the shadow results, then the next stack ray or
// the pixel's sample.
// Each pixel of the chunk: its samples of the round in sample order
// (cudaDrawPixels, utility.cu:199-229: colour / spp and coverage alpha when a
// ray of the sample hit).
__global__ void k_cw_accumulate(CwArgs c, uint32_t ns) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= c.nchunk) return;
  const uint32_t rl = c.l0 + p;
  const uint32_t pix = (c.region.y + rl / c.region.z) * c.w + c.region.x + rl % c.region.z;
  const float weight = 1.0f / (float)c.spp;
  float4 px = c.pixels[pix];
  for (uint32_t k = 0; k < ns; ++k) {
    const float4 f = c.fin[(size_t)k * c.nchunk + p];
    if (f.w != 0.0f) {
      px.x += f.x * weight;
      px.y += f.y * weight;
      px.z += f.z * weight;
      px.w += weight;
    }
  }
  c.pixels[pix] = px;
}

__global__ __launch_bounds__(kBlock) void k_cw_resolve(CwArgs c, PbrArgs a, uint32_t w,
                                                       const uint32_t* __restrict__ qin) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = a.counters[2 * w];
  if (i < n) {
    const uint32_t l = qin[i];
    const float4 c1 = c.c1[i];
    const uint32_t flags = __float_as_uint(c1.w);
    float4 sc4 = c.st_c[l];
    V3 color = mk(sc4.x, sc4.y, sc4.z);
    if (flags & kCwHit) {
      const float4 emw = c.emw[i];
      const V3 color1 = mk(c1.x, c1.y, c1.z);
      V3 Li = mk(emw.x, emw.y, emw.z);
      uint32_t slot = 0;
      const auto lit = [&](V3& Ls) {  // Ls + the slot's contribution when unoccluded
        const size_t e = (size_t)slot * c.stride + i;
        const float4 cv = c.contrib[e];
        if (cv.w != 0.0f && a.sh_occ[e] == 0) Ls = Ls + mk(cv.x, cv.y, cv.z);
        ++slot;
      };
      for (uint32_t li = 0; li < c.sc.num_lights; ++li) {
        const aur_light& L = c.sc.lights[li];
        if (L.type == AUR_LIGHT_AMBIENT) Li = Li + (color1 * kInvPi) * (ld3(L.color) * L.intensity * L.ambient);
        if (L.type != AUR_LIGHT_AREA) {
          const size_t e = (size_t)slot * c.stride + i;
          const float4 cv = c.contrib[e];
          V3 add = mk(0.0f);
          if (cv.w != 0.0f && a.sh_occ[e] == 0) add = mk(cv.x, cv.y, cv.z);
          ++slot;
          Li = Li + add;
        } else {
          V3 acc = mk(0.0f);
          for (int s = 0; s < L.samples; ++s) {
            V3 Ls = mk(0.0f);
            lit(Ls);
            acc = acc + Ls;
          }
          const float nsf = (float)L.samples;
          Li = Li + mk(acc.x / nsf, acc.y / nsf, acc.z / nsf);
        }
      }
      color = color + (Li * emw.w);
    }
    if (flags & kCwMore)
      c.st_c[l] = make_float4(color.x, color.y, color.z, sc4.w);
    else
      c.fin[l] = make_float4(color.x, color.y, color.z, __float_as_int(sc4.w) != -1 ? 1.0f : 0.0f);
  }
}

constexpr uint32_t kAdjMaxDepth = 16;
constexpr uint32_t kAdjMaxParams = 6144;  // 48 KiB of LDS accumulators

struct AdjointState {
  const float4* dimg = nullptr;  // dLoss/dI per pixel (RGB used)
  uint32_t* absmax = nullptr;    // max |dLoss/dI| bits
  double* total = nullptr;       // running gradient sums
  uint32_t nparams = 0;
};

// Buffers of a run: per-path state for `cap` paths and `nslots` next-event
// slots, vertex records for the adjoint.
aur_status pbr_reserve(aur_context* ctx, uint32_t cap, uint32_t nslots, uint32_t max_depth, bool records) {
  FrameState& fr = ctx->frame;
  if ((uint64_t)cap * std::max<uint32_t>(nslots, 1u) > 0xFFFFFFFFull)
    return ctx->fail(AUR_ERR_INVALID_ARG, "render: too many paths per pass");
  const size_t nsh = std::max<size_t>(1, (size_t)cap * nslots);
  AUR_HIP(ctx, fr.counters.reserve(kCounterWords));
  AUR_HIP(ctx, fr.pass_g0.reserve(1));
  AUR_HIP(ctx, fr.ray_o.reserve((size_t)cap * 2));
  AUR_HIP(ctx, fr.ray_d.reserve((size_t)cap * 2));
  AUR_HIP(ctx, fr.beta.reserve(cap));
  AUR_HIP(ctx, fr.radiance.reserve(cap));
  AUR_HIP(ctx, fr.hits.reserve(cap));
  AUR_HIP(ctx, fr.queue[0].reserve(cap));
  AUR_HIP(ctx, fr.queue[1].reserve(cap));
  AUR_HIP(ctx, fr.sh_o.reserve(nsh));
  AUR_HIP(ctx, fr.sh_d.reserve(nsh));
  AUR_HIP(ctx, fr.sh_c.reserve(nsh));
  AUR_HIP(ctx, fr.sh_occ.reserve(nsh));
  AUR_HIP(ctx, fr.sh_queue.reserve(nsh));
  if (records) {
    AUR_HIP(ctx, fr.rec_mat.reserve((size_t)cap * max_depth));
    AUR_HIP(ctx, fr.rec_rr.reserve((size_t)cap * max_depth));
    AUR_HIP(ctx, fr.rec_w.reserve(std::max<size_t>(1, (size_t)cap * max_depth * nslots)));
  }
  return AUR_OK;
}

// Launch shape of k_pbr_trace for up to max_items rays: the traversal stack
// takes at most levels - 1 rows per lane, the first `rows` in LDS, the rest
// (with a window) in global memory; the grid is the resident one.
template <bool kMulti>
aur_status trace_shape(aur_context* ctx, uint32_t levels, uint64_t max_items, PbrArgs& a, size_t& lds,
                       uint32_t& grid) {
  FrameState& fr = ctx->frame;
  const uint32_t maxrows = levels > 1 ? levels - 1 : 1;
  const uint32_t want = ctx->render_params.stack_lds_rows;
  const uint32_t rows = want >= maxrows ? maxrows : want;
  lds = (size_t)rows * kBlock * sizeof(float2);
  grid = resident_grid(ctx, rows < maxrows ? (const void*)k_pbr_trace<kMulti, true> : (const void*)k_pbr_trace<kMulti, false>,
                       lds, max_items);
  a.gstack = nullptr;
  a.klds = rows;
  a.gstack_rows = 0;
  if (rows < maxrows) {
    const size_t entries = (size_t)(maxrows - rows) * grid * kBlock;
    if (entries * sizeof(float2) >= (1ull << 31))
      return ctx->fail(AUR_ERR_INVALID_ARG, "render: global stack rows exceed 2 GiB");
    AUR_HIP(ctx, fr.gstack.reserve(entries));
    a.gstack = fr.gstack.p;
    a.gstack_rows = maxrows - rows;
  }
  return AUR_OK;
}

// Kernels of one pass (camera, bounce loop, accumulation or adjoint sweep,
// advance) on the context stream.
template <bool kMulti>
void pbr_pass(aur_context* ctx, PbrArgs& a, size_t lds, uint32_t grid_trace,
              AdjointState* adj, uint32_t sweep_grid) {
  FrameState& fr = ctx->frame;
  hipStream_t s = ctx->stream;
  const uint32_t cap = a.cap;
  k_pbr_camera<kMulti><<<nblocks(cap, 256), 256, 0, s>>>(a, fr.queue[0].p);
  // bounce b: shade, then one launch tracing bounce b + 1's extension rays
  // with bounce b's shadow rays, then bounce b's next-event resolve
  const uint32_t sh = a.nslots ? 2u : 0u;
  const auto trace = [&](uint32_t w, uint32_t parts, uint32_t roff) {
    if (a.gstack)
      k_pbr_trace<kMulti, true><<<grid_trace, kBlock, lds, s>>>(a, w, parts, roff);
    else
      k_pbr_trace<kMulti, false><<<grid_trace, kBlock, lds, s>>>(a, w, parts, roff);
  };
  trace(0u, 1u, 0u);
  for (uint32_t b = 0; b < a.max_depth; ++b) {
    const uint32_t w = b % 3u, wn = (b + 1u) % 3u;
    uint32_t* qin = fr.queue[b & 1u].p;
    uint32_t* qout = fr.queue[(b + 1u) & 1u].p;
    if (adj)
      k_pbr_shade<kMulti, true><<<nblocks(cap, 128), 128, 0, s>>>(a, w, b, qin, qout, a.jobs);
    else
      k_pbr_shade<kMulti, false><<<nblocks(cap, 128), 128, 0, s>>>(a, w, b, qin, qout, a.jobs);
    const uint32_t parts = (b + 1u < a.max_depth ? 1u : 0u) | sh;
    if (parts)
      trace(wn, parts, ((b + 1u) & 1u) * cap);
    if (a.nslots) {
      if (adj)
        k_pbr_resolve<true><<<nblocks(cap, 256), 256, 0, s>>>(a, w, b, qin);
      else
        k_pbr_resolve<false><<<nblocks(cap, 256), 256, 0, s>>>(a, w, b, qin);
    }
  }
  if (!adj) {
    k_pbr_accumulate<kMulti><<<nblocks(cap, 256), 256, 0, s>>>(a);
  } else if constexpr (!kMulti) {
    const size_t slds = (size_t)adj->nparams * sizeof(unsigned long long);
    if (a.max_depth <= 8)
      k_adj_sweep<8><<<sweep_grid, 256, slds, s>>>(a, adj->dimg, adj->absmax, fr.adj_partials.p, adj->nparams);
    else
      k_adj_sweep<16><<<sweep_grid, 256, slds, s>>>(a, adj->dimg, adj->absmax, fr.adj_partials.p, adj->nparams);
    k_adj_reduce<<<nblocks(adj->nparams, 64), 64, 0, s>>>(fr.adj_partials.p, sweep_grid, adj->nparams, adj->total);
  }
  k_pbr_advance<<<1, 64, 0, s>>>(fr.pass_g0.p, cap);
}

// Every pass of a run, launched directly or as replays of one captured pass.
template <bool kMulti>
aur_status pbr_passes(aur_context* ctx, const PbrRun& run, unsigned long long* stats, AdjointState* adj) {
  FrameState& fr = ctx->frame;
  hipStream_t s = ctx->stream;
  const uint32_t cap = run.cap;
  if (run.nslots > kMaxSlots)
    return ctx->fail(AUR_ERR_UNSUPPORTED, "render: more than 64 next-event slots (light samples) per vertex");
  aur_status st = pbr_reserve(ctx, cap, run.nslots, run.max_depth, adj != nullptr);
  if (st != AUR_OK) return st;
  PbrArgs a{};
  a.job = run.jobs[0];
  a.njobs = run.njobs;
  if (kMulti) {
    AUR_HIP(ctx, fr.jobs.reserve(run.njobs));
    AUR_HIP(ctx, fr.path_job.reserve(cap));
    AUR_HIP(ctx, fr.sh_job.reserve(std::max<size_t>(1, (size_t)cap * run.nslots)));
    AUR_HIP(ctx, hipMemcpyAsync(fr.jobs.p, run.jobs, sizeof(JobDev) * run.njobs, hipMemcpyHostToDevice, s));
    AUR_HIP(ctx, hipStreamSynchronize(s));  // run.jobs is pageable host memory
    a.jobs = fr.jobs.p;
    a.path_job = fr.path_job.p;
  }
  // a batch renders its jobs whole: clear their accumulation buffers first
  // (one job or many; frames accumulate onto the existing buffer instead)
  if (run.clear_images) {
    if (kMulti)
      k_clear_jobs<<<std::min<uint32_t>(run.njobs, 1024), 256, 0, s>>>(fr.jobs.p, run.njobs);
    else
      k_clear_pixels<<<nblocks(a.job.npix, 256), 256, 0, s>>>(a.job.pixels, a.job.npix);
  }
  a.g0p = fr.pass_g0.p;
  a.end = run.end;
  a.cap = cap;
  a.nslots = run.nslots;
  a.max_depth = run.max_depth;
  a.refill = ctx->render_params.refill;
  a.ray_o = fr.ray_o.p;
  a.ray_d = fr.ray_d.p;
  a.beta = fr.beta.p;
  a.radiance = fr.radiance.p;
  a.hits = fr.hits.p;
  a.sh_o = fr.sh_o.p;
  a.sh_d = fr.sh_d.p;
  a.sh_c = fr.sh_c.p;
  a.sh_occ = fr.sh_occ.p;
  a.sh_queue = fr.sh_queue.p;
  a.sh_job = kMulti ? fr.sh_job.p : nullptr;
  a.counters = fr.counters.p;
  a.stats = stats;
  if (adj) {
    a.rec_mat = fr.rec_mat.p;
    a.rec_rr = fr.rec_rr.p;
    a.rec_w = fr.rec_w.p;
  }
  size_t lds = 0;
  uint32_t grid_trace = 0;
  st = trace_shape<kMulti>(ctx, run.levels, (uint64_t)cap * (1u + run.nslots), a, lds, grid_trace);
  if (st != AUR_OK) return st;
  uint32_t sweep_grid = 0;
  if (adj) {
    // a few resident blocks per CU; each keeps LDS accumulators for all parameters
    sweep_grid = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * 4, (cap + 255) / 256);
    AUR_HIP(ctx, fr.adj_partials.reserve((size_t)sweep_grid * adj->nparams));
  }
  const uint64_t npass = (run.end - run.begin + cap - 1) / cap;
  k_set_u64<<<1, 64, 0, s>>>(fr.pass_g0.p, run.begin);
  AUR_LAUNCH_CHECK(ctx, "pbr setup");
  std::vector<hipEvent_t> ev;
  if (run.pass_end_ms) {
    ev.resize(npass + 1);
    for (auto& e : ev) AUR_HIP(ctx, hipEventCreate(&e));
    AUR_HIP(ctx, hipEventRecord(ev[0], s));
  }
  aur_status result = AUR_OK;
  const bool use_graph = ctx->render_params.graph && npass > 1;
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      pbr_pass<kMulti>(ctx, a, lds, grid_trace, adj, sweep_grid);
      e = hipStreamEndCapture(s, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) result = ctx->hip_fail(e, "pbr graph capture");
  }
  PassThrottle throttle(s);
  for (uint64_t i = 0; result == AUR_OK && i < npass; ++i) {
    hipError_t te = throttle.wait(i);
    if (te != hipSuccess) {
      result = ctx->hip_fail(te, "pass throttle");
      break;
    }
    if (ctx->take_cancel()) {
      (void)hipStreamSynchronize(s);
      result = ctx->fail(AUR_ERR_CANCELLED, "render cancelled");
      break;
    }
    if (exec) {
      const hipError_t e = hipGraphLaunch(exec, s);
      if (e != hipSuccess) result = ctx->hip_fail(e, "hipGraphLaunch");
    } else {
      pbr_pass<kMulti>(ctx, a, lds, grid_trace, adj, sweep_grid);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) result = ctx->hip_fail(e, "pbr kernels");
    }
    if (result == AUR_OK && run.pass_end_ms) {
      const hipError_t e = hipEventRecord(ev[i + 1], s);
      if (e != hipSuccess) result = ctx->hip_fail(e, "hipEventRecord");
    }
    te = throttle.mark(i);
    if (result == AUR_OK && te != hipSuccess) result = ctx->hip_fail(te, "pass throttle");
  }
  if (exec) {
    (void)hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
  }
  if (run.pass_end_ms) {
    (void)hipStreamSynchronize(s);
    if (result == AUR_OK) {
      run.pass_end_ms->assign(npass, 0.0f);
      for (uint64_t i = 0; i < npass; ++i)
        (void)hipEventElapsedTime(&(*run.pass_end_ms)[i], ev[0], ev[i + 1]);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  return result;
}

// The frame as a single job over the context's current scene.
JobDev frame_job(aur_context* ctx, const aur_camera* cam, const aur_render_opts* o) {
  FrameState& fr = ctx->frame;
  JobDev J{};
  J.nodes = ctx->nodes.p;
  J.normals = ctx->normals.p;
  J.shader_ids = ctx->shader_ids.p;
  J.shaders = ctx->shaders.p;
  J.lights = ctx->lights.p;
  J.pixels = fr.pixels.p;
  J.texcoords = ctx->texcoords.p;
  J.textures = ctx->textures.p;
  J.texels = ctx->texels.p;
  J.num_textures = ctx->num_textures;
  J.path_start = 0;
  J.n_nodes = ctx->count / 2;
  J.root = ctx->root_slab;
  J.rec_bytes = ctx->rec_bytes;
  J.num_shaders = ctx->num_shaders;
  J.num_lights = ctx->num_lights;
  J.nslots = ctx->num_slots;
  J.w = fr.width;
  J.h = fr.height;
  J.rx0 = fr.left;
  J.ry0 = fr.bottom;
  J.rw = fr.right - fr.left + 1;
  J.npix = J.rw * (fr.top - fr.bottom + 1);
  J.spp = o->spp;
  J.max_depth = o->max_depth;
  J.rr_start = o->rr_start;
  J.seed = o->seed;
  J.paths = J.npix * o->spp;
  J.cam = *cam;
  return J;
}

// Compat-mode samples [s0, s1) of the frame region as wavefront passes
// (k_cw_* above): rounds of up to max_paths (pixel, sample) entries, every
// pixel of the region with as many whole samples as fit, or chunks of
// pixels with one sample when the region alone is larger.
aur_status compat_passes(aur_context* ctx, const aur_camera* cam, const aur_render_opts* o, uint32_t s0,
                         uint32_t s1, unsigned long long* stats) {
  FrameState& fr = ctx->frame;
  hipStream_t s = ctx->stream;
  const uint32_t rw = fr.right - fr.left + 1, npix = rw * (fr.top - fr.bottom + 1);
  const uint32_t nslots = ctx->compat_slots;
  const uint64_t per_pixel = 1ull + nslots;
  // entries per round: max_paths, and at most 2^26 shadow slots (72 B each)
  const uint64_t budget = std::min<uint64_t>(std::max<uint32_t>(ctx->render_params.max_paths, 1u),
                                             std::max<uint64_t>(1u, (1ull << 26) / std::max<uint32_t>(nslots, 1u)));
  const uint32_t pix_cap = (uint32_t)std::max<uint64_t>(1u, std::min<uint64_t>(npix, budget));
  const uint32_t ns_max = pix_cap == npix ? (uint32_t)std::max<uint64_t>(1u, budget / npix) : 1u;
  const uint32_t cap = pix_cap * ns_max;  // entries per round
  aur_status st = pbr_reserve(ctx, cap, nslots, 1, false);
  if (st != AUR_OK) return st;
  AUR_HIP(ctx, fr.cw_o.reserve(cap));
  AUR_HIP(ctx, fr.cw_d.reserve(cap));
  AUR_HIP(ctx, fr.cw_c.reserve(cap));
  AUR_HIP(ctx, fr.cw_i.reserve(cap));
  AUR_HIP(ctx, fr.cw_emw.reserve(cap));
  AUR_HIP(ctx, fr.cw_c1.reserve(cap));
  AUR_HIP(ctx, fr.cw_contrib.reserve(std::max<size_t>(1, (size_t)cap * nslots)));
  AUR_HIP(ctx, fr.cw_fin.reserve(cap));
  // waiting mirror rays exist only under shaders that are both reflective
  // and translucent
  float4* pend = nullptr;
  if (ctx->branching_shaders) {
    AUR_HIP(ctx, fr.compat_pend.reserve((size_t)kCompatPend * 2 * cap));
    pend = fr.compat_pend.p;
  }
  PbrArgs a{};
  a.job = frame_job(ctx, cam, o);
  a.cap = cap;
  a.nslots = nslots;
  a.max_depth = 1;
  a.refill = ctx->render_params.refill;
  a.ray_o = fr.ray_o.p;
  a.ray_d = fr.ray_d.p;
  a.hits = fr.hits.p;
  a.sh_o = fr.sh_o.p;
  a.sh_d = fr.sh_d.p;
  a.sh_occ = fr.sh_occ.p;
  a.sh_queue = fr.sh_queue.p;
  a.counters = fr.counters.p;
  a.stats = stats;
  size_t lds = 0;
  uint32_t grid_trace = 0;
  st = trace_shape<false>(ctx, ctx->levels, (uint64_t)cap * per_pixel, a, lds, grid_trace);
  if (st != AUR_OK) return st;
  CwArgs c{};
  c.sc = SceneDev{ctx->nodes.p,      ctx->count / 2,   ctx->rec_bytes,   ctx->normals.p,
                  ctx->shader_ids.p, ctx->shaders.p,   ctx->num_shaders, ctx->lights.p,
                  ctx->num_lights,   ctx->texcoords.p, ctx->textures.p,  ctx->texels.p,
                  ctx->num_textures};
  c.cam = *cam;
  c.w = fr.width;
  c.h = fr.height;
  c.spp = o->spp;
  c.seed = o->seed;
  c.region = make_uint4(fr.left, fr.bottom, rw, npix);
  c.stride = cap;
  c.nslots = nslots;
  c.st_o = fr.cw_o.p;
  c.st_d = fr.cw_d.p;
  c.st_c = fr.cw_c.p;
  c.st_i = fr.cw_i.p;
  c.pend = pend;
  c.emw = fr.cw_emw.p;
  c.c1 = fr.cw_c1.p;
  c.contrib = fr.cw_contrib.p;
  c.fin = fr.cw_fin.p;
  c.pixels = fr.pixels.p;
  const auto trace = [&](uint32_t w, uint32_t parts, uint32_t roff) {
    if (a.gstack)
      k_pbr_trace<false, true><<<grid_trace, kBlock, lds, s>>>(a, w, parts, roff);
    else
      k_pbr_trace<false, false><<<grid_trace, kBlock, lds, s>>>(a, w, parts, roff);
  };
  // every stack ray is a hit of depth <= 8 or below one: at most 1 + 2 * 9
  // rays per pixel sample, so the stacks empty within that many levels
  constexpr uint32_t kMaxLevels = 19;
  PassThrottle throttle(s);
  uint32_t round = 0;
  for (uint32_t sa = s0; sa < s1; sa += ns_max) {
    const uint32_t ns = std::min(ns_max, s1 - sa);
    AUR_HIP(ctx, throttle.wait(round));
    if (ctx->take_cancel()) {
      AUR_HIP(ctx, hipStreamSynchronize(s));
      return ctx->fail(AUR_ERR_CANCELLED, "render cancelled");
    }
    c.s0 = sa;
    for (uint32_t l0 = 0; l0 < npix; l0 += pix_cap) {
      c.l0 = l0;
      c.nchunk = std::min(pix_cap, npix - l0);
      c.n = c.nchunk * ns;
      k_cw_init<<<nblocks(c.n, 256), 256, 0, s>>>(c, a, fr.queue[0].p);
      // level it: one launch traces its rays with level it - 1's shadow
      // rays, then level it - 1's sums, then its shading; after the last
      // level, its shadow rays and sums
      for (uint32_t it = 0;; ++it) {
        const uint32_t w = it % 3u, wn = (it + 1u) % 3u, wp = (it + 2u) % 3u;
        trace(w, 1u | (it > 0 && nslots ? 2u : 0u), (it & 1u) * cap);
        if (it > 0)
          k_cw_resolve<<<nblocks(c.n, kBlock), kBlock, 0, s>>>(c, a, wp, fr.queue[(it + 1u) & 1u].p);
        k_cw_shade<<<nblocks(c.n, kBlock), kBlock, 0, s>>>(c, a, w, wn, fr.queue[it & 1u].p,
                                                          fr.queue[(it + 1u) & 1u].p, ((it + 1u) & 1u) * cap);
        AUR_LAUNCH_CHECK(ctx, "compat wavefront");
        uint32_t more = 0;
        if (ctx->child_shaders) {  // else no stack holds a second ray
          AUR_HIP(ctx, hipMemcpyAsync(&more, fr.counters.p + 2 * wn, sizeof(more), hipMemcpyDeviceToHost, s));
          AUR_HIP(ctx, hipStreamSynchronize(s));
        }
        if (more == 0) {
          if (nslots) trace(wn, 2u, 0u);
          k_cw_resolve<<<nblocks(c.n, kBlock), kBlock, 0, s>>>(c, a, w, fr.queue[it & 1u].p);
          AUR_LAUNCH_CHECK(ctx, "compat wavefront");
          break;
        }
        if (it + 1 >= kMaxLevels) return ctx->fail(AUR_ERR_DEVICE, "render: compat ray stack did not empty");
      }
      k_cw_accumulate<<<nblocks(c.nchunk, 256), 256, 0, s>>>(c, ns);
      AUR_LAUNCH_CHECK(ctx, "compat accumulate");
    }
    AUR_HIP(ctx, throttle.mark(round));
    ++round;
  }
  return AUR_OK;
}

// 3x3 tent filter of the region (cudaFilterPixelsKernel,
// src/kernels/utility.cu:231-258): weights 1/16, 1/8, 1/4, taps clamped to
// the region, sums in the reference's order.  The reference clamps y + 1 to
// region.bottom, a bug; here it is clamped to region.top.
__global__ void k_filter_pixels(uint32_t w, int left, int right, int bottom, int top,
                                const float4* __restrict__ in, float4* __restrict__ out) {
  const int x = left + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = bottom + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (x > right || y > top) return;
  const int xm = max(x - 1, left), xp = min(x + 1, right);
  const int ym = max(y - 1, bottom), yp = min(y + 1, top);
  auto at = [&](int yy, int xx) { return in[(size_t)yy * w + xx]; };
  const float4 t[9] = {at(ym, xm), at(ym, x), at(ym, xp), at(y, xm), at(y, x),
                       at(y, xp),  at(yp, xm), at(yp, x), at(yp, xp)};
  const float wt[9] = {0.0625f, 0.125f, 0.0625f, 0.125f, 0.25f, 0.125f, 0.0625f, 0.125f, 0.0625f};
  float4 v = make_float4(t[0].x * wt[0], t[0].y * wt[0], t[0].z * wt[0], t[0].w * wt[0]);
  for (int i = 1; i < 9; ++i) {
    v.x = v.x + t[i].x * wt[i];
    v.y = v.y + t[i].y * wt[i];
    v.z = v.z + t[i].z * wt[i];
    v.w = v.w + t[i].w * wt[i];
  }
  out[(size_t)y * w + x] = v;
}

}  // namespace

aur_status frame_filter(aur_context* ctx) {
  FrameState& fr = ctx->frame;
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "filter: no frame");
  const size_t n = (size_t)fr.width * fr.height;
  DevBuf<float4> out;
  AUR_HIP(ctx, out.reserve(n));
  // pixels outside the region keep their values
  AUR_HIP(ctx, hipMemcpyAsync(out.p, fr.pixels.p, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
  const dim3 block(32, 8);
  const dim3 grid(nblocks(fr.right - fr.left + 1, 32), nblocks(fr.top - fr.bottom + 1, 8));
  k_filter_pixels<<<grid, block, 0, ctx->stream>>>(fr.width, (int)fr.left, (int)fr.right, (int)fr.bottom,
                                                   (int)fr.top, fr.pixels.p, out.p);
  AUR_LAUNCH_CHECK(ctx, "k_filter_pixels");
  // back into the accumulation buffer, so that its device pointer stays valid
  AUR_HIP(ctx, hipMemcpyAsync(fr.pixels.p, out.p, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
  AUR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AUR_OK;
}

aur_status pbr_run(aur_context* ctx, const PbrRun& run, unsigned long long* stats) {
  if (run.njobs == 0 || run.end <= run.begin) return AUR_OK;
  return run.njobs == 1 ? pbr_passes<false>(ctx, run, stats, nullptr)
                        : pbr_passes<true>(ctx, run, stats, nullptr);
}

aur_status frame_render_check(aur_context* ctx, const aur_render_opts* o, uint32_t s0, uint32_t s1) {
  const FrameState& fr = ctx->frame;
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "render: geometry not built");
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "render: no frame");
  if (ctx->num_shaders == 0) return ctx->fail(AUR_ERR_STATE, "render: no shaders uploaded");
  if (o->spp == 0 || s1 > o->spp || s0 > s1) return ctx->fail(AUR_ERR_INVALID_ARG, "render: bad sample range");
  if (o->mode == AUR_MODE_PBR && o->max_depth == 0) return ctx->fail(AUR_ERR_INVALID_ARG, "render: max_depth 0");
  if (o->mode != AUR_MODE_COMPAT && o->mode != AUR_MODE_PBR) return ctx->fail(AUR_ERR_INVALID_ARG, "render: mode");
  if (o->mode == AUR_MODE_PBR && ctx->num_slots > kMaxSlots)
    return ctx->fail(AUR_ERR_UNSUPPORTED, "render: more than 64 next-event slots (light samples) per vertex");
  if (o->mode == AUR_MODE_COMPAT && ctx->compat_slots > kMaxCompatSlots)
    return ctx->fail(AUR_ERR_UNSUPPORTED, "render: more than 4096 compat shadow rays (light samples) per hit");
  const uint32_t npix = (fr.right - fr.left + 1) * (fr.top - fr.bottom + 1);
  if ((uint64_t)npix * o->spp > 0xFFFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "render: pixels * spp exceeds 2^32");
  return AUR_OK;
}

aur_status frame_render(aur_context* ctx, const aur_camera* cam, const aur_render_opts* o,
                        uint32_t s0, uint32_t s1) {
  const aur_status chk = frame_render_check(ctx, o, s0, s1);
  if (chk != AUR_OK) return chk;
  FrameState& fr = ctx->frame;
  hipStream_t s = ctx->stream;
  const uint32_t rw = fr.right - fr.left + 1, rh = fr.top - fr.bottom + 1, npix = rw * rh;
  AUR_HIP(ctx, fr.counters.reserve(kCounterWords));
  unsigned long long* stats = reinterpret_cast<unsigned long long*>(fr.counters.p + kStatsOffset);
  k_zero_u64<<<1, 64, 0, s>>>(stats, 2);
  if (o->mode == AUR_MODE_COMPAT) {
    if (s1 > s0) {
      const aur_status st = compat_passes(ctx, cam, o, s0, s1, stats);
      if (st != AUR_OK) return st;
    }
  } else if (s1 > s0) {
    const JobDev J = frame_job(ctx, cam, o);
    PbrRun run;
    run.jobs = &J;
    run.njobs = 1;
    run.begin = (unsigned long long)s0 * npix;
    run.end = (unsigned long long)s1 * npix;
    // whole samples per pass: enough paths in flight to fill the device
    const uint32_t batch_cap = std::max<uint32_t>(1u, std::max<uint32_t>(ctx->render_params.max_paths, 1u) / npix);
    run.cap = npix * std::min<uint32_t>(batch_cap, s1 - s0);
    run.levels = ctx->levels;
    run.nslots = ctx->num_slots;
    run.max_depth = o->max_depth;
    const aur_status st = pbr_passes<false>(ctx, run, stats, nullptr);
    if (st != AUR_OK) return st;
  }
  unsigned long long hs[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, s));
  AUR_HIP(ctx, hipStreamSynchronize(s));
  fr.rays_traced = hs[0] + hs[1];
  fr.rays_extension = hs[0];
  fr.rays_shadow = hs[1];
  return AUR_OK;
}

aur_status frame_adjoint(aur_context* ctx, const aur_camera* cam, const aur_render_opts* o,
                         const float* d_image, float* g_albedo, float* g_light) {
  FrameState& fr = ctx->frame;
  if (!ctx->built) return ctx->fail(AUR_ERR_STATE, "adjoint: geometry not built");
  if (fr.width == 0) return ctx->fail(AUR_ERR_STATE, "adjoint: no frame");
  if (ctx->num_shaders == 0) return ctx->fail(AUR_ERR_STATE, "adjoint: no shaders uploaded");
  if (o->mode != AUR_MODE_PBR) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: PBR mode only");
  if (ctx->num_slots > kMaxSlots)
    return ctx->fail(AUR_ERR_UNSUPPORTED, "adjoint: more than 64 next-event slots (light samples) per vertex");
  if (o->spp == 0 || o->max_depth == 0) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: spp and max_depth must be > 0");
  if (o->max_depth > kAdjMaxDepth) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: max_depth > 16");
  if (!d_image || !g_albedo || (ctx->num_lights && !g_light))
    return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: null buffer");
  const uint32_t nparams = 3u * (ctx->num_shaders + ctx->num_lights);
  if (nparams > kAdjMaxParams) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: more than 2048 shaders + lights");
  if (ctx->textured_shaders)
    return ctx->fail(AUR_ERR_UNSUPPORTED, "adjoint: texture-mapped shaders (albedo is not a parameter there)");
  hipStream_t s = ctx->stream;
  const uint32_t npix = (fr.right - fr.left + 1) * (fr.top - fr.bottom + 1);  // region pixels
  const uint32_t nframe = fr.width * fr.height;
  if ((uint64_t)npix * o->spp > 0xFFFFFFFFull) return ctx->fail(AUR_ERR_INVALID_ARG, "adjoint: pixels * spp exceeds 2^32");
  AUR_HIP(ctx, fr.counters.reserve(kCounterWords));
  unsigned long long* stats = reinterpret_cast<unsigned long long*>(fr.counters.p + kStatsOffset);
  k_zero_u64<<<1, 64, 0, s>>>(stats, 2);
  AdjointState adj;
  adj.dimg = reinterpret_cast<const float4*>(d_image);
  adj.nparams = nparams;
  AUR_HIP(ctx, fr.adj_total.reserve(nparams));
  AUR_HIP(ctx, fr.adj_absmax.reserve(1));
  adj.total = fr.adj_total.p;
  adj.absmax = fr.adj_absmax.p;
  k_zero_f64<<<nblocks(nparams, 256), 256, 0, s>>>(adj.total, nparams);
  k_adj_zero_u32<<<1, 64, 0, s>>>(adj.absmax);
  k_adj_absmax<<<std::min<uint32_t>(nblocks(nframe, 256), 1024), 256, 0, s>>>(adj.dimg, nframe, adj.absmax);
  AUR_LAUNCH_CHECK(ctx, "adjoint setup");
  const JobDev J = frame_job(ctx, cam, o);
  PbrRun run;
  run.jobs = &J;
  run.njobs = 1;
  run.begin = 0;
  run.end = (unsigned long long)npix * o->spp;
  const uint32_t batch_cap = std::max<uint32_t>(1u, std::max<uint32_t>(ctx->render_params.max_paths, 1u) / npix);
  run.cap = npix * std::min<uint32_t>(batch_cap, o->spp);
  run.levels = ctx->levels;
  run.nslots = ctx->num_slots;
  run.max_depth = o->max_depth;
  const aur_status st = pbr_passes<false>(ctx, run, stats, &adj);
  if (st != AUR_OK) return st;
  k_adj_finish<<<nblocks(nparams, 256), 256, 0, s>>>(adj.total, 3u * ctx->num_shaders,
                                                     3u * ctx->num_lights, g_albedo, g_light);
  AUR_LAUNCH_CHECK(ctx, "k_adj_finish");
  unsigned long long hs[2] = {0, 0};
  AUR_HIP(ctx, hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, s));
  AUR_HIP(ctx, hipStreamSynchronize(s));
  fr.rays_traced = hs[0] + hs[1];
  fr.rays_extension = hs[0];
  fr.rays_shadow = hs[1];
  return AUR_OK;
}

}  // namespace aur
