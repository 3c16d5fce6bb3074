// ajg_sppm.hip — stochastic progressive photon mapping over the GPU scene
// (SPPMRenderer, ajisai/core/renderer/sppm.cpp:35-303, with
// VisiblePointSearcher, renderer/photon_mapping.h:29-153).
//
// Per render:
//   k_sppm_vp         visible points, once: the reference reseeds every pixel
//                     with x + y*W in every iteration (sppm.cpp:222), so each
//                     iteration traces the same points and adds the same
//                     direct-light terms; they are traced once and the terms
//                     kept, in order, per pixel
// Per iteration (all on the stream, no host round trip):
//   k_sppm_direct     add the pixel's direct-light terms (the reference's
//                     sequence of float additions)
//   k_sppm_grid_*     hash grid of visible points: count per entry, scan,
//                     fill (cell side 1.05 x the largest radius, cells
//                     [min, max) of each radius box)
//   k_sppm_photon     one thread per photon: light choice, SampleEmit, bounce
//                     loop, deposits (atomic adds into the visible points)
//   k_sppm_maxr       largest radius among visible points (next cell side)
//   k_sppm_update     UpdatePixel
// Then k_sppm_final: ComputePixelRadiance.
// Traversal is the exact-order closest/any-hit of the path tracer, called
// per thread.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ajg_scene.h"

namespace ajg {

static constexpr uint32_t kSppmEntries = 40960;  // VisiblePointSearcher entry count (sppm.cpp:209)
static constexpr uint32_t kSppmRegsPerVp = 8;    // cells per visible point (2 per axis at most)

struct SppmParams {
  DevCamera cam;
  uint32_t W, H, n_pix;
  int max_depth;          // forward_max_depth
  uint32_t term_cap;      // direct-light terms per pixel: max_depth + 1
  float init_radius;
  float bmin[3];          // hash-grid bound minimum (world min - init radius)
  float world_c[3];       // world box centre and |max - centre| (Light::Process)
  float world_r;
  float sel_cdf[kMaxLights + 1];  // light selector (FP32 prefix sums of equal weights)
  float sel_pdf;
  uint32_t n_lights;
  uint32_t photons;
  int photon_min_depth, photon_max_depth;
  float cont_prob, alpha;
  uint32_t seed_base;
  uint32_t list_cap;
};

// Per-pixel state, SoA.
struct SppmState {
  uint32_t *valid;
  float *pos, *wr, *beta, *ng, *ns;  // 3 per pixel
  uint32_t *mat;
  uint32_t *n_terms;
  float *terms;                      // term_cap x 3 per pixel
  float *direct, *tau, *phi;         // 3 per pixel
  float *radius, *N;
  int *M;
  uint32_t *grid_count;              // kSppmEntries + 1
  uint32_t *grid_cursor;             // kSppmEntries
  uint32_t *list;                    // pixel of each registration
  float4 *list_pr;                   // its visible point's position and radius
  uint32_t *maxr_bits;               // float bits of the largest radius (radii >= 0)
  unsigned long long *ctr;           // deposits, registrations, list overflow
};

__device__ __forceinline__ f3 ld3v(const float *p, uint32_t i) {
  return mk3(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]);
}
__device__ __forceinline__ void st3v(float *p, uint32_t i, f3 v) {
  p[3 * (size_t)i] = v.x;
  p[3 * (size_t)i + 1] = v.y;
  p[3 * (size_t)i + 2] = v.z;
}
__device__ __forceinline__ bool finite3(f3 v) {
  return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}

#define AJG_SPPM_LDS                                 \
  __shared__ int s_node[kLdsStack * kTravThreads];   \
  __shared__ float s_entry[kLdsStack * kTravThreads]; \
  int *const ln = s_node + threadIdx.x;              \
  float *const le = s_entry + threadIdx.x

template <int SPILL>
__device__ __forceinline__ THit scene_closest(const DevScene &sc, const DevRay &r, int *ln,
                                              float *le) {
  uint32_t nn = 0, nt = 0;
  return trace_closest<SPILL, false>(sc.bvh, make_tray(r.o, r.d, r.tmin, r.tmax), ln, le, &nn, &nt);
}
template <int SPILL>
__device__ __forceinline__ bool scene_occluded(const DevScene &sc, const DevRay &r, int *ln,
                                               float *le) {
  return trace_any<SPILL>(sc.bvh, make_tray(r.o, r.d, r.tmin, r.tmax), ln, le);
}

// MISSampleLight / MISSampleAreaLight / MISSampleEnvLight (helper.h:44-133)
template <int SPILL>
__device__ f3 mis_sample_light(const DevScene &sc, const DevLight &l, const Inct &it,
                               const Bsdf &bs, Pcg32 &rng, int *ln, float *le) {
  const f3 zero = mk3(0.f, 0.f, 0.f);
  LightSample ls;
  if (l.type == AJG_LIGHT_ENV) {
    const float first = pcg_next_float(rng);
    const float second = pcg_next_float(rng);
    ls = env_sample(l.env, it.pos, second, first, sc.radius);
  } else if (l.type == kLightArea) {
    f3 sam;
    sam.x = pcg_next_float(rng);
    sam.y = pcg_next_float(rng);
    sam.z = pcg_next_float(rng);
    ls = area_sample(l, it.pos, sam);
  } else {
    const float u0 = pcg_next_float(rng);
    const float u1 = pcg_next_float(rng);
    ls = dir_sample(l, it.pos, u0, u1, sc.radius);
  }
  if (is_black(ls.radiance) || ls.pdf <= 0.f) return zero;
  const float len = length(sub(ls.pos, it.pos)) - AJG_EPS;
  if (len < AJG_EPS) return zero;
  const f3 to_light = normalize(sub(ls.pos, it.pos));
  if (scene_occluded<SPILL>(sc, make_ray(it.pos, to_light, AJG_EPS, 0.99f * len), ln, le))
    return zero;
  const f3 bf = bsdf_eval_all(bs, to_light, it.wr);
  if (is_black(bf)) return zero;
  const f3 nrm = l.type == kLightArea ? it.ns : normalize(it.ns);
  const f3 f = scale(mul(ls.radiance, bf), fabsf(dot(to_light, nrm)));
  const float bpdf = bsdf_pdf_all(bs, to_light, it.wr);
  return scale(divs(f, ls.pdf), power_heuristic(ls.pdf, bpdf));
}

// MISSampleBsdf (helper.h:208-262)
template <int SPILL>
__device__ f3 mis_sample_bsdf(const DevScene &sc, const Inct &it, const Bsdf &bs, Pcg32 &rng,
                              int *ln, float *le) {
  const f3 zero = mk3(0.f, 0.f, 0.f);
  f3 sam;
  sam.x = pcg_next_float(rng);
  sam.y = pcg_next_float(rng);
  sam.z = pcg_next_float(rng);
  const BsdfSample s = bsdf_sample_all(bs, it.wr, sam);
  if (!s.valid || is_black(s.f) || s.pdf == 0.f) return zero;
  const f3 dn = normalize(s.dir);
  const DevRay r = make_ray(eps_offset(it, dn), dn, AJG_EPS, __int_as_float(0x7f800000));
  const THit h = scene_closest<SPILL>(sc, r, ln, le);
  if (h.slot < 0) {
    if (sc.env_index < 0) return zero;
    const DevEnv &e = sc.lights[sc.env_index].env;
    const f3 lr = env_radiance(e, r.d);
    if (is_black(lr)) return zero;
    const f3 f = scale(mul(lr, s.f), fabsf(dot(it.ns, r.d)));
    return scale(divs(f, s.pdf), power_heuristic(s.pdf, env_pdf(e, r.d)));
  }
  const Inct hi = post_intersect(sc, r, h.t, h.slot, h.u, h.v);
  if (!hi.light || !(dot(hi.ng, hi.wr) > 0)) return zero;
  const DevLight &l = sc.lights[hi.light - 1];
  if (is_black(l.radiance)) return zero;
  const f3 f = scale(mul(l.radiance, s.f), fabsf(dot(it.ns, r.d)));
  return scale(divs(f, s.pdf), power_heuristic(s.pdf, area_pdf(l, r.o, hi.pos, hi.ng)));
}

// HasDiffuseComponent: Lambert always, Disney when metallic < 1 (disney.cpp:79)
__device__ __forceinline__ bool has_diffuse(const DevMaterial &m) {
  return m.type == AJG_MAT_DIFFUSE || m.metallic < 1.f;
}

// TracerVP (sppm.cpp:35-106) for every pixel, once. The direct-light terms
// the reference adds to the pixel's running sum are kept in order.
template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_sppm_vp(DevScene sc, SppmParams P, SppmState S) {
  AJG_SPPM_LDS;
  const uint32_t p = blockIdx.x * kTravThreads + threadIdx.x;
  if (p >= P.n_pix) return;
  const uint32_t x = p % P.W, y = p / P.W;
  Pcg32 rng;
  pcg_seed(rng, (uint64_t)(int64_t)(int)(x + y * P.W));
  const float u = (float)x + pcg_next_float(rng);
  const float v = (float)y + pcg_next_float(rng);
  const float l0 = pcg_next_float(rng);
  const float l1 = pcg_next_float(rng);
  DevRay ray = camera_ray(P.cam, u, v, l0, l1);
  f3 beta = mk3(1.f, 1.f, 1.f);
  uint32_t nt = 0;
  float *terms = S.terms + (size_t)p * P.term_cap * 3;
  auto term = [&](f3 t) {
    terms[3 * nt] = t.x;
    terms[3 * nt + 1] = t.y;
    terms[3 * nt + 2] = t.z;
    ++nt;
  };
  uint32_t valid = 0;
  for (int depth = 0; depth < P.max_depth; ++depth) {
    const THit h = scene_closest<SPILL>(sc, ray, ln, le);
    if (h.slot < 0) {
      if (depth == 0 && sc.env_index >= 0)
        term(mul(beta, env_radiance(sc.lights[sc.env_index].env, ray.d)));
      break;
    }
    const Inct it = post_intersect(sc, ray, h.t, h.slot, h.u, h.v);
    const DevMaterial &m = sc.mats[it.material];
    const Bsdf bs = make_bsdf(m, it.ng, it.ns);
    if (depth == 0 && it.light && dot(it.ng, it.wr) > 0)
      term(mul(beta, sc.lights[it.light - 1].radiance));
    f3 sum_di = mk3(0.f, 0.f, 0.f);
    for (uint32_t k = 0; k < sc.n_lights; ++k)
      sum_di = add(sum_di, mul(beta, mis_sample_light<SPILL>(sc, sc.lights[k], it, bs, rng, ln, le)));
    sum_di = add(sum_di, mul(beta, mis_sample_bsdf<SPILL>(sc, it, bs, rng, ln, le)));
    term(scale(sum_di, 1.f / 1));
    if (has_diffuse(m) || depth == P.max_depth - 1) {
      if (finite3(beta)) {
        valid = 1;
        st3v(S.pos, p, it.pos);
        st3v(S.wr, p, it.wr);
        st3v(S.beta, p, beta);
        st3v(S.ng, p, it.ng);
        st3v(S.ns, p, it.ns);
        S.mat[p] = it.material;
      }
      break;
    }
    f3 sam;
    sam.x = pcg_next_float(rng);
    sam.y = pcg_next_float(rng);
    sam.z = pcg_next_float(rng);
    const BsdfSample s = bsdf_sample_all(bs, it.wr, sam);
    if (!s.valid || is_black(s.f) || s.pdf < 0.f) break;
    beta = mul(beta, divs(scale(s.f, fabsf(dot(it.ng, s.dir))), s.pdf));
    ray = make_ray(eps_offset(it, s.dir), s.dir, AJG_EPS, __int_as_float(0x7f800000));
  }
  S.valid[p] = valid;
  S.n_terms[p] = nt;
}

__global__ void __launch_bounds__(256)
k_sppm_init(SppmParams P, SppmState S) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p == 0) {
    *S.maxr_bits = __float_as_uint(P.init_radius);
    for (int k = 0; k < 3; ++k) S.ctr[k] = 0ull;
  }
  if (p >= P.n_pix) return;
  st3v(S.direct, p, mk3(0.f, 0.f, 0.f));
  st3v(S.tau, p, mk3(0.f, 0.f, 0.f));
  st3v(S.phi, p, mk3(0.f, 0.f, 0.f));
  S.radius[p] = P.init_radius;
  S.N[p] = 0.f;
  S.M[p] = 0;
}

// the iteration's additions to the running direct light (sppm.cpp:48,58-60,74)
__global__ void __launch_bounds__(256)
k_sppm_direct(SppmParams P, SppmState S) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P.n_pix) return;
  f3 d = ld3v(S.direct, p);
  const uint32_t nt = S.n_terms[p];
  const float *t = S.terms + (size_t)p * P.term_cap * 3;
  for (uint32_t k = 0; k < nt; ++k) d = add(d, mk3(t[3 * k], t[3 * k + 1], t[3 * k + 2]));
  st3v(S.direct, p, d);
}

// VisiblePointSearcher::ToGrid / ToEntry (photon_mapping.h:124-138); the
// hash is utility/hash.h's HashCombine over std::hash<int> (the value itself)
__device__ __forceinline__ void sppm_to_grid(const SppmParams &P, float side, f3 p, int *c) {
  c[0] = (int)(std_max(p.x - P.bmin[0], 0.f) / side);
  c[1] = (int)(std_max(p.y - P.bmin[1], 0.f) / side);
  c[2] = (int)(std_max(p.z - P.bmin[2], 0.f) / side);
}
__device__ __forceinline__ uint64_t hash_combine(uint64_t a, uint64_t b) {
  return a ^ (b + 0x9e3779b97f4a7c15ull + (a << 6) + (a >> 2));
}
__device__ __forceinline__ uint32_t sppm_entry(int x, int y, int z) {
  const uint64_t low3 = (uint64_t)((x & 1) | ((y & 1) << 1) | (z & 2));
  const uint64_t h = hash_combine(hash_combine((uint64_t)(int64_t)x, (uint64_t)(int64_t)y),
                                  (uint64_t)(int64_t)z);
  return (uint32_t)(((h << 3) % kSppmEntries) | low3);
}
__device__ __forceinline__ float cell_side(const SppmState &S) {
  return 1.05f * __uint_as_float(*S.maxr_bits);
}

// AddVp (photon_mapping.h:84-101): count, then fill, the cells [min, max)
template <bool FILL>
__device__ __forceinline__ uint32_t sppm_register(const SppmParams &P, const SppmState &S,
                                                  uint32_t p) {
  const float side = cell_side(S);
  const f3 c = ld3v(S.pos, p);
  const float r = S.radius[p];
  int mn[3], mx[3];
  sppm_to_grid(P, side, sub(c, mk3(r, r, r)), mn);
  sppm_to_grid(P, side, add(c, mk3(r, r, r)), mx);
  // at most 2 cells per axis (side >= 1.05 x every radius); anything else
  // (a non-finite point) is counted as an overflow instead of looped over
  for (int k = 0; k < 3; ++k)
    if (mn[k] < 0 || mx[k] - mn[k] > 2) {
      if (!FILL) atomicAdd(&S.ctr[2], 1ull);
      return 0u;
    }
  uint32_t regs = 0;
  for (int z = mn[2]; z < mx[2]; ++z)
    for (int y = mn[1]; y < mx[1]; ++y)
      for (int x = mn[0]; x < mx[0]; ++x) {
        const uint32_t e = sppm_entry(x, y, z);
        if (!FILL) {
          atomicAdd(&S.grid_count[e + 1], 1u);
          ++regs;
        } else {
          const uint32_t at = S.grid_count[e] + atomicAdd(&S.grid_cursor[e], 1u);
          if (at < P.list_cap) {
            S.list[at] = p;
            S.list_pr[at] = make_float4(c.x, c.y, c.z, r);
          } else
            atomicAdd(&S.ctr[2], 1ull);
        }
      }
  return regs;
}
template <bool FILL>
__global__ void __launch_bounds__(256)
k_sppm_grid(SppmParams P, SppmState S) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  uint32_t regs = 0;
  if (p < P.n_pix && S.valid[p]) regs = sppm_register<FILL>(P, S, p);
  if (!FILL) {  // registrations: one atomic per wave
    for (int off = 32; off > 0; off >>= 1) regs += (uint32_t)__shfl_down((int)regs, off, 64);
    if ((threadIdx.x & 63) == 0 && regs) atomicAdd(&S.ctr[1], (unsigned long long)regs);
  }
}

// inclusive scan of grid_count[1..kSppmEntries] in one block (40 per thread)
__global__ void __launch_bounds__(1024)
k_sppm_grid_scan(SppmState S) {
  __shared__ uint32_t part[1024];
  constexpr uint32_t per = kSppmEntries / 1024;  // 40
  const uint32_t t = threadIdx.x;
  uint32_t *c = S.grid_count + 1 + t * per;
  uint32_t s = 0;
  for (uint32_t k = 0; k < per; ++k) s += c[k];
  part[t] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (uint32_t k = 0; k < per; ++k) {
    run += c[k];
    c[k] = run;
  }
}

struct LightEmit {
  f3 pos, dir, normal, radiance;
  float pdf_pos, pdf_dir;
};

// Light::SampleEmit: AreaLight (area_light.cpp:60-73), DirectionalLight
// (directional_light.cpp:60-79), EnvLight (env_light.cpp:105-147).
__device__ LightEmit sample_emit(const SppmParams &P, const DevLight &l, Pcg32 &rng) {
  LightEmit r;
  r.pos = r.dir = r.normal = r.radiance = mk3(0.f, 0.f, 0.f);
  r.pdf_pos = r.pdf_dir = 0.f;
  const f3 center = mk3(P.world_c[0], P.world_c[1], P.world_c[2]);
  const float radius = P.world_r;
  if (l.type == kLightArea) {
    f3 sam;
    sam.x = pcg_next_float(rng);
    sam.y = pcg_next_float(rng);
    sam.z = pcg_next_float(rng);
    f3 ng;
    area_geom_sample(l, sam, &r.pos, &ng, &r.pdf_pos);
    const float s0 = pcg_next_float(rng);
    const float s1 = pcg_next_float(rng);
    const f3 lwi = cosine_hemisphere(s0, s1, &r.pdf_dir);
    r.dir = to_world(make_frame(ng), lwi);
    r.normal = ng;
    r.radiance = l.radiance;
  } else if (l.type == AJG_LIGHT_DIRECTIONAL) {
    const float c0 = pcg_next_float(rng);
    const float c1 = pcg_next_float(rng);
    const f3 ref2light = to_world(l.fr, uniform_cone(c0, c1, l.angle_cos));
    r.pdf_dir = uniform_cone_pdf(l.angle_cos);
    const float d0 = pcg_next_float(rng);
    const float d1 = pcg_next_float(rng);
    const f2 disk = disk_concentric(d0, d1);
    const Frame dc = make_frame(ref2light);
    r.pos = add(center,
                scale(add(add(scale(dc.dx, disk.x), scale(dc.dy, disk.y)), ref2light), radius));
    r.dir = neg(ref2light);
    r.normal = neg(normalize(ref2light));
    r.radiance = l.radiance;
    r.pdf_pos = 1.f / (AJG_PI * radius * radius);
  } else {
    const DevEnv &e = l.env;
    // SampleContinuous(Next1D(), Next1D()): arguments drawn right to left
    const float second = pcg_next_float(rng);
    const float first = pcg_next_float(rng);
    float pdf_v, pdf_u;
    int iv;
    const float v = distrib_sample(e.marg_cdf, e.marg_func, e.marg_int, e.h, second, &pdf_v, &iv);
    float u = distrib_sample(e.cond_cdf + (size_t)iv * (e.w + 1), e.cond_func + (size_t)iv * e.w,
                             e.cond_int[iv], e.w, first, &pdf_u, nullptr);
    float pdf = pdf_u * pdf_v;
    if (pdf == 0.f) return r;
    u -= e.rot / 180 * AJG_PI * AJG_INV_2PI;
    if (u > 1.f) u -= 1.f;
    if (u < 0.f) u += 1.f;
    const float phi = u * AJG_TWO_PI;
    const float theta = v * AJG_PI;
    const float sin_t = sinf(theta);
    pdf = sin_t != 0.f ? pdf / (2.f * AJG_PI * AJG_PI * sin_t) : 0.f;
    f3 dir = mk3(-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi));
    dir = neg(dir);
    const float d0 = pcg_next_float(rng);
    const float d1 = pcg_next_float(rng);
    const f2 disk = disk_concentric(d0, d1);
    const Frame dc = make_frame(neg(dir));
    r.pos = add(center, scale(sub(add(scale(dc.dx, disk.x), scale(dc.dy, disk.y)), dir), radius));
    r.dir = dir;
    r.normal = normalize(dir);
    r.radiance = tex_sample(e, u, v);
    r.pdf_pos = 1.f / (AJG_PI * radius * radius);
    r.pdf_dir = pdf;
  }
  return r;
}

// VisiblePointSearcher::AddPhoton (photon_mapping.h:103-121)
__device__ uint32_t add_photon(const DevScene &sc, const SppmParams &P, const SppmState &S,
                               float side, f3 pos, f3 phi, f3 wr) {
  int c[3];
  sppm_to_grid(P, side, pos, c);
  const uint32_t e = sppm_entry(c[0], c[1], c[2]);
  const uint32_t b = S.grid_count[e], end = min(S.grid_count[e + 1], P.list_cap);
  uint32_t n = 0;
  for (uint32_t k = b; k < end; ++k) {
    const float4 q = S.list_pr[k];
    const f3 d = sub(mk3(q.x, q.y, q.z), pos);
    if (dot(d, d) > q.w * q.w) continue;
    const uint32_t p = S.list[k];
    const Bsdf bs = make_bsdf(sc.mats[S.mat[p]], ld3v(S.ng, p), ld3v(S.ns, p));
    const f3 dphi = mul(phi, bsdf_eval_all(bs, wr, ld3v(S.wr, p)));
    if (!finite3(dphi)) continue;
    atomicAdd(&S.phi[3 * (size_t)p], dphi.x);
    atomicAdd(&S.phi[3 * (size_t)p + 1], dphi.y);
    atomicAdd(&S.phi[3 * (size_t)p + 2], dphi.z);
    atomicAdd(&S.M[p], 1);
    ++n;
  }
  return n;
}

// TracerPhoton (sppm.cpp:108-143); DefaultScene::SampleLight picks the light
// (default_scene.cpp:76-85).
template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_sppm_photon(DevScene sc, SppmParams P, SppmState S, uint32_t iter) {
  AJG_SPPM_LDS;
  const uint32_t i = blockIdx.x * kTravThreads + threadIdx.x;
  uint32_t deposits = 0;
  if (i < P.photons && P.n_lights > 0) {
    Pcg32 rng;
    pcg_seed(rng, sample_seed(P.seed_base, i, iter));
    const float side = cell_side(S);
    const float us = pcg_next_float(rng);
    uint32_t lo = 0, len = P.n_lights + 1;
    while (len > 0) {
      const uint32_t half = len / 2;
      if (!(us <= P.sel_cdf[lo + half])) {
        lo += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    const int idx = clampi((int)lo - 1, 0, (int)P.n_lights - 1);
    const LightEmit em = sample_emit(P, sc.lights[idx], rng);
    if (!(is_black(em.radiance) || em.pdf_dir == 0.f || em.pdf_pos == 0.f)) {
      f3 coef = divs(scale(em.radiance, fabsf(dot(em.normal, em.dir))),
                     P.sel_pdf * em.pdf_pos * em.pdf_dir);
      DevRay ray = make_ray(em.pos, em.dir, AJG_EPS, __int_as_float(0x7f800000));
      for (int depth = 1; depth <= P.photon_max_depth; ++depth) {
        if (depth > P.photon_min_depth) {
          if (pcg_next_float(rng) > P.cont_prob) break;
          coef = divs(coef, P.cont_prob);
        }
        const THit h = scene_closest<SPILL>(sc, ray, ln, le);
        if (h.slot < 0) break;
        const Inct it = post_intersect(sc, ray, h.t, h.slot, h.u, h.v);
        if (depth > 1) deposits += add_photon(sc, P, S, side, it.pos, coef, it.wr);
        const Bsdf bs = make_bsdf(sc.mats[it.material], it.ng, it.ns);
        f3 sam;
        sam.x = pcg_next_float(rng);
        sam.y = pcg_next_float(rng);
        sam.z = pcg_next_float(rng);
        const BsdfSample s = bsdf_sample_all(bs, it.wr, sam);
        if (!s.valid || is_black(s.f) || s.pdf < 0.f) break;
        coef = mul(coef, scale(divs(s.f, s.pdf), fabsf(dot(it.ng, s.dir))));
        ray = make_ray(eps_offset(it, s.dir), s.dir, AJG_EPS, __int_as_float(0x7f800000));
      }
    }
  }
  // one atomic per wave for the deposit count
  unsigned long long w = deposits;
  for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off, 64);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(&S.ctr[0], w);
}

// largest radius among visible points before the update (sppm.cpp:263-271);
// radii are >= 0, so their float bits order like the values
__global__ void __launch_bounds__(256)
k_sppm_maxr_clear(SppmState S) {
  if (threadIdx.x == 0) *S.maxr_bits = 0u;
}
__global__ void __launch_bounds__(256)
k_sppm_maxr(SppmParams P, SppmState S) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  __shared__ uint32_t wmax[4];
  uint32_t b = 0;
  if (p < P.n_pix && S.valid[p]) b = __float_as_uint(S.radius[p]);
  for (int off = 32; off > 0; off >>= 1) b = max(b, (uint32_t)__shfl_down((int)b, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {  // one atomic per block
    const uint32_t m = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    if (m) atomicMax(S.maxr_bits, m);
  }
}

// UpdatePixel (sppm.cpp:145-167); a zero largest radius falls back to the
// initial one for the next cell side
__global__ void __launch_bounds__(256)
k_sppm_update(SppmParams P, SppmState S) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p == 0 && *S.maxr_bits == 0u) *S.maxr_bits = __float_as_uint(P.init_radius);
  if (p >= P.n_pix || !S.valid[p]) return;
  const int m = S.M[p];
  if (m > 0) {
    const float n = S.N[p], r = S.radius[p];
    const float new_n = n + P.alpha * (float)m;
    const float new_r = r * sqrtf(new_n / (n + (float)m));
    const f3 phi = ld3v(S.phi, p);
    const f3 tau = divs(scale(add(ld3v(S.tau, p), mul(ld3v(S.beta, p), phi)), new_r * new_r),
                        r * r);
    st3v(S.tau, p, tau);
    S.N[p] = new_n;
    S.radius[p] = new_r;
    S.M[p] = 0;
    st3v(S.phi, p, mk3(0.f, 0.f, 0.f));
  }
}

// ComputePixelRadiance (sppm.cpp:169-178)
__global__ void __launch_bounds__(256)
k_sppm_final(SppmParams P, SppmState S, uint32_t iteration_count, unsigned long long photon_n,
             float *radiance, ajg_sppm_pixel *state) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P.n_pix) return;
  const f3 direct = ld3v(S.direct, p);
  const float r = S.radius[p];
  const float dem = (float)photon_n * AJG_PI * r * r;
  const f3 L = add(divs(direct, (float)iteration_count), divs(ld3v(S.tau, p), dem));
  st3v(radiance, p, L);
  if (state) {
    ajg_sppm_pixel o;
    const bool v = S.valid[p] != 0;
    const f3 pos = v ? ld3v(S.pos, p) : mk3(0.f, 0.f, 0.f);
    const f3 beta = v ? ld3v(S.beta, p) : mk3(0.f, 0.f, 0.f);
    const f3 tau = ld3v(S.tau, p);
    o.pos[0] = pos.x, o.pos[1] = pos.y, o.pos[2] = pos.z;
    o.beta[0] = beta.x, o.beta[1] = beta.y, o.beta[2] = beta.z;
    o.direct[0] = direct.x, o.direct[1] = direct.y, o.direct[2] = direct.z;
    o.tau[0] = tau.x, o.tau[1] = tau.y, o.tau[2] = tau.z;
    o.radius = r;
    o.n = S.N[p];
    o.valid = v ? 1u : 0u;
    o.pad = 0u;
    state[p] = o;
  }
}

// Device buffers of an SPPM render, kept on the scene and reused.
struct SppmBuffers {
  uint32_t n_pix = 0, term_cap = 0;
  std::vector<void *> bufs;
  SppmState st;
  void release() {
    for (void *p : bufs) (void)hipFree(p);
    bufs.clear();
    n_pix = term_cap = 0;
  }
  ~SppmBuffers() { release(); }
};

static int ensure_sppm(ajg_ctx *ctx, SppmBuffers *b, uint32_t n, uint32_t cap) {
  if (b->n_pix >= n && b->term_cap >= cap && !b->bufs.empty()) return AJG_OK;
  b->release();
  auto al = [&](void **p, size_t bytes, const char *what) {
    int r = dev_alloc(ctx, p, bytes, what);
    if (r == AJG_OK) b->bufs.push_back(*p);
    return r;
  };
  SppmState &s = b->st;
  const size_t N = n;
  int st = 0;
  if ((st = al((void **)&s.valid, 4 * N, "sppm state")) || (st = al((void **)&s.pos, 12 * N, "sppm state")) ||
      (st = al((void **)&s.wr, 12 * N, "sppm state")) || (st = al((void **)&s.beta, 12 * N, "sppm state")) ||
      (st = al((void **)&s.ng, 12 * N, "sppm state")) || (st = al((void **)&s.ns, 12 * N, "sppm state")) ||
      (st = al((void **)&s.mat, 4 * N, "sppm state")) || (st = al((void **)&s.n_terms, 4 * N, "sppm state")) ||
      (st = al((void **)&s.terms, 12 * N * cap, "sppm direct-light terms")) ||
      (st = al((void **)&s.direct, 12 * N, "sppm state")) || (st = al((void **)&s.tau, 12 * N, "sppm state")) ||
      (st = al((void **)&s.phi, 12 * N, "sppm state")) || (st = al((void **)&s.radius, 4 * N, "sppm state")) ||
      (st = al((void **)&s.N, 4 * N, "sppm state")) || (st = al((void **)&s.M, 4 * N, "sppm state")) ||
      (st = al((void **)&s.grid_count, 4 * (kSppmEntries + 1), "sppm grid")) ||
      (st = al((void **)&s.grid_cursor, 4 * kSppmEntries, "sppm grid")) ||
      (st = al((void **)&s.list, 4 * kSppmRegsPerVp * N, "sppm grid lists")) ||
      (st = al((void **)&s.list_pr, 16 * kSppmRegsPerVp * N, "sppm grid lists")) ||
      (st = al((void **)&s.maxr_bits, 256, "sppm counters")) ||
      (st = al((void **)&s.ctr, 256, "sppm counters"))) {
    b->release();
    return st;
  }
  b->n_pix = n;
  b->term_cap = cap;
  return AJG_OK;
}

}  // namespace ajg

using namespace ajg;

void sppm_free(void *cache) { delete (SppmBuffers *)cache; }

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

extern "C" int ajg_render_sppm(ajg_scene *sc, const ajg_camera *camera, const ajg_sppm_args *a,
                               float *radiance, ajg_sppm_pixel *state, ajg_sppm_stats *stats,
                               void *stream) {
  if (!sc || !camera || !a || !radiance) return AJG_ERR_INVALID;
  ajg_ctx *ctx = sc->ctx;
  if (int st = injected(ctx)) return st;
  if (sc->insts) return fail(ctx, AJG_ERR_INVALID, "sppm: instanced scenes are not supported");
  const int W = (int)camera->resolution[0], H = (int)camera->resolution[1];
  if (W <= 0 || H <= 0 || W % 64 || H % 64)
    return fail(ctx, AJG_ERR_INVALID,
                "sppm: film width and height must be positive multiples of 64 (the reference's "
                "64-pixel tiles are not clipped to the film)");
  if (a->forward_max_depth < 1 || a->forward_max_depth > 64 || a->iteration_count < 1 ||
      a->photons_per_iteration < 0 || a->photon_max_depth > 1024 || !(a->init_radius != 0.f) ||
      !std::isfinite(a->init_radius))
    return fail(ctx, AJG_ERR_INVALID, "sppm: bad arguments");
  if ((uint64_t)W * (uint64_t)H > (1ull << 26))
    return fail(ctx, AJG_ERR_INVALID, "sppm: film above 2^26 pixels");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, AJG_ERR_HIP, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  const ajg_accel *acc = sc->accel;
  const int spill = spill_for_depth(acc->depth);
  if (!spill) return fail(ctx, AJG_ERR_INTERNAL, "sppm: tree too deep for the traversal stack");

  SppmParams P;
  memset(&P, 0, sizeof(P));
  memcpy(P.cam.r2c, camera->raster2camera, sizeof(P.cam.r2c));
  memcpy(P.cam.c2w, camera->camera2world, sizeof(P.cam.c2w));
  P.cam.lens_radius = camera->lens_radius;
  P.cam.focal_distance = camera->focal_distance;
  P.W = (uint32_t)W;
  P.H = (uint32_t)H;
  P.n_pix = P.W * P.H;
  P.max_depth = a->forward_max_depth;
  P.term_cap = (uint32_t)a->forward_max_depth + 1;
  // SPPMRenderer::Render (sppm.cpp:186-204): world bound and initial radius
  const f3 wmn = ld3(sc->world), wmx = ld3(sc->world + 3);
  float init_radius = a->init_radius;
  if (init_radius < 0) init_radius = length(sub(wmx, wmn)) / 1000.f;
  P.init_radius = init_radius;
  const f3 bmin = sub(wmn, mk3(init_radius, init_radius, init_radius));
  P.bmin[0] = bmin.x, P.bmin[1] = bmin.y, P.bmin[2] = bmin.z;
  box3 wb;
  wb.mn = wmn;
  wb.mx = wmx;
  const f3 wc = box_center(wb);
  P.world_c[0] = wc.x, P.world_c[1] = wc.y, P.world_c[2] = wc.z;
  P.world_r = sc->radius;
  // light selector: equal weights normalised by their sum (default_scene.cpp:50-55),
  // Distribution::SetFunction in FP32 (distribution.h:40-58)
  P.n_lights = sc->n_lights;
  if (P.n_lights) {
    float tab[kMaxLights], sum = 0.f;
    for (uint32_t i = 0; i < P.n_lights; ++i) tab[i] = 1.f;
    for (uint32_t i = 0; i < P.n_lights; ++i) sum += tab[i];
    const float ratio = 1.f / sum;
    for (uint32_t i = 0; i < P.n_lights; ++i) tab[i] *= ratio;
    P.sel_pdf = tab[0];
    const float inv = 1.f / (float)P.n_lights;
    P.sel_cdf[0] = 0.f;
    for (uint32_t i = 1; i <= P.n_lights; ++i) P.sel_cdf[i] = P.sel_cdf[i - 1] + tab[i - 1] * inv;
    const float fi = P.sel_cdf[P.n_lights];
    for (uint32_t i = 0; i <= P.n_lights; ++i) P.sel_cdf[i] /= fi;
  }
  P.photons = (uint32_t)a->photons_per_iteration;
  P.photon_min_depth = a->photon_min_depth;
  P.photon_max_depth = a->photon_max_depth;
  P.cont_prob = a->photon_cont_prob;
  P.alpha = a->update_alpha;
  P.seed_base = a->photon_seed_base;
  P.list_cap = kSppmRegsPerVp * P.n_pix;

  SppmBuffers *b = (SppmBuffers *)sc->sppm_cache;
  if (!b) {
    b = new SppmBuffers();
    sc->sppm_cache = b;
  }
  if (int st = ensure_sppm(ctx, b, P.n_pix, P.term_cap)) return st;
  const SppmState S = b->st;
  const DevScene ds = dev_scene(sc);
  const bool timing = (a->flags & AJG_RENDER_TIMING) != 0;
  hipEvent_t ev[8];
  for (auto &e : ev) e = nullptr;
  auto done = [&](int code) {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    return code;
  };
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return done(fail(ctx, AJG_ERR_HIP, "sppm events"));
  float ms_grid = 0.f, ms_photons = 0.f;
  const dim3 g256(cdiv(P.n_pix, 256)), gtrav(cdiv(P.n_pix, kTravThreads));
  (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_sppm_init, g256, dim3(256), 0, s, P, S);
  if (spill == kSpillS)
    hipLaunchKernelGGL(k_sppm_vp<kSpillS>, gtrav, dim3(kTravThreads), 0, s, ds, P, S);
  else
    hipLaunchKernelGGL(k_sppm_vp<kSpillL>, gtrav, dim3(kTravThreads), 0, s, ds, P, S);
  (void)hipEventRecord(ev[1], s);
  const dim3 gph(cdiv(std::max(P.photons, 1u), kTravThreads));
  for (int it = 0; it < a->iteration_count; ++it) {
    if (cancelled(ctx)) {  // polled as each iteration is enqueued
      (void)hipStreamSynchronize(s);
      return done(fail(ctx, AJG_ERR_CANCELLED, "sppm cancelled"));
    }
    hipLaunchKernelGGL(k_sppm_direct, g256, dim3(256), 0, s, P, S);
    if (timing) (void)hipEventRecord(ev[2], s);
    if (hipMemsetAsync(S.grid_count, 0, 4 * (kSppmEntries + 1), s) != hipSuccess ||
        hipMemsetAsync(S.grid_cursor, 0, 4 * kSppmEntries, s) != hipSuccess)
      return done(fail(ctx, AJG_ERR_HIP, "sppm grid clear"));
    hipLaunchKernelGGL(k_sppm_grid<false>, g256, dim3(256), 0, s, P, S);
    hipLaunchKernelGGL(k_sppm_grid_scan, dim3(1), dim3(1024), 0, s, S);
    hipLaunchKernelGGL(k_sppm_grid<true>, g256, dim3(256), 0, s, P, S);
    if (timing) (void)hipEventRecord(ev[3], s);
    if (P.photons && P.n_lights) {
      if (spill == kSpillS)
        hipLaunchKernelGGL(k_sppm_photon<kSpillS>, gph, dim3(kTravThreads), 0, s, ds, P, S, (uint32_t)it);
      else
        hipLaunchKernelGGL(k_sppm_photon<kSpillL>, gph, dim3(kTravThreads), 0, s, ds, P, S, (uint32_t)it);
    }
    if (timing) (void)hipEventRecord(ev[4], s);
    hipLaunchKernelGGL(k_sppm_maxr_clear, dim3(1), dim3(64), 0, s, S);
    hipLaunchKernelGGL(k_sppm_maxr, g256, dim3(256), 0, s, P, S);
    hipLaunchKernelGGL(k_sppm_update, g256, dim3(256), 0, s, P, S);
    if (timing) {
      float t0 = 0.f, t1 = 0.f;
      if (hipEventSynchronize(ev[4]) != hipSuccess)
        return done(fail(ctx, AJG_ERR_HIP, "sppm timing"));
      (void)hipEventElapsedTime(&t0, ev[2], ev[3]);
      (void)hipEventElapsedTime(&t1, ev[3], ev[4]);
      ms_grid += t0;
      ms_photons += t1;
    }
  }
  const unsigned long long photon_n =
      (unsigned long long)a->iteration_count * (unsigned long long)a->photons_per_iteration;
  hipLaunchKernelGGL(k_sppm_final, g256, dim3(256), 0, s, P, S, (uint32_t)a->iteration_count,
                     photon_n, radiance, state);
  (void)hipEventRecord(ev[5], s);
  unsigned long long ctr[3] = {0, 0, 0};
  if (hipMemcpyAsync(ctr, S.ctr, sizeof(ctr), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(ctx, AJG_ERR_HIP, "sppm render"));
  if (hipGetLastError() != hipSuccess) return done(fail(ctx, AJG_ERR_HIP, "sppm kernel launch"));
  if (ctr[2]) return done(fail(ctx, AJG_ERR_INTERNAL, "sppm: grid list overflow"));
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->photons = photon_n;
    stats->deposits = ctr[0];
    stats->grid_entries = ctr[1];
    float ms = 0.f, mv = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[5]);
    (void)hipEventElapsedTime(&mv, ev[0], ev[1]);
    stats->ms_total = ms;
    stats->ms_visible_points = mv;
    stats->ms_photons = ms_photons;
    stats->ms_grid = ms_grid;
    uint32_t nv = 0;
    std::vector<uint32_t> valid(P.n_pix);
    if (hipMemcpy(valid.data(), S.valid, 4 * (size_t)P.n_pix, hipMemcpyDeviceToHost) == hipSuccess)
      for (uint32_t v : valid) nv += v ? 1u : 0u;
    stats->visible_points = nv;
  }
  return done(AJG_OK);
}
