// ajg_vcm.hip — vertex connection and merging over the GPU scene
// (VCMRenderer, ajisai/core/renderer/vcm.cpp:130-751, with HashGrid,
// renderer/hashgrid.h:30-175).
//
// Per iteration (stream-ordered, no host round trip):
//   k_vcm_light    one thread per light sub-path (W*H of them): emission,
//                  bounces, vertices kept in the path's own slots, light
//                  tracing splats (ConnectToCamera) added atomically
//   k_vcm_bbox     box of the light vertices (exact min/max on ordered keys)
//   k_vcm_grid_*   HashGrid: per-cell counts (W*H cells), exclusive scan, fill
//   k_vcm_camera   one thread per pixel: camera sub-path with emission hits,
//                  direct illumination, connections to the pixel's own light
//                  path, merging with the light vertices in 8 grid cells
// Reference behaviour kept: the reverse pdf of a scattering event is
// PdfAll(-new_dir, new_dir) (vcm.cpp:416-423); the camera path's first ray is
// offset twice; a scattered path restarts at the hit point without
// EpsOffset; each vertex draws a BSDF sample whose only use is is_delta;
// splats are summed over iterations while colors are averaged. Streams: light
// path i of iteration k draws from PCG32 seeded with mix(seed_base, i, 2k), the
// camera path of pixel p from mix(seed_base, p, 2k+1); the reference draws
// all paths from one sampler in sequence.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "ajg_scene.h"

namespace ajg {

struct VcmParams {
  DevCamera cam;
  float w2c[16], c2r[16], w2r[16];
  float cam_pos[3], cam_dir[3];
  float res_x, res_y;
  uint32_t W, H, n;
  int use_vm, use_vc, min_len, max_len;
  float cont_prob;
  float light_count, mis_vm, mis_vc, vm_norm, radius, radius_sqr;
  float sel_cdf[kMaxLights + 1];
  float sel_pdf;
  uint32_t n_lights;
  float world_c[3], world_r;
  uint32_t seed_base, iter;
  uint32_t vcap;   // vertex slots per light path
};

// Light vertex, 5 x float4: (pos, path length), (in_dir, material),
// (throughput, dVCM), (shading normal, dVC), (geometric normal, dVM)
static constexpr int kVtx = 5;

struct VcmState {
  float4 *verts;
  uint32_t *n_verts;        // per light path
  float *color;             // W*H*3, summed over iterations
  float *splat;             // W*H*3
  uint32_t *bbox;           // 6 ordered keys: min xyz, max xyz
  uint32_t *cell_count;     // n + 1
  uint32_t *cell_start;     // n + 1 (exclusive scan)
  uint32_t *cell_cursor;    // n
  uint32_t *indices;        // vertex slots by cell
  uint32_t *scan_tmp;
  unsigned long long *ctr;  // light vertices
};

__device__ __forceinline__ f3 f3a(const float *p) { return mk3(p[0], p[1], p[2]); }

template <int SPILL>
__device__ __forceinline__ THit vcm_closest(const DevScene &sc, const DevRay &r, int *ln,
                                            float *le) {
  uint32_t nn = 0, nt = 0;
  return trace_closest<SPILL, false>(sc.bvh, make_tray(r.o, r.d, r.tmin, r.tmax), ln, le, &nn, &nt);
}
template <int SPILL>
__device__ __forceinline__ bool vcm_occluded(const DevScene &sc, const DevRay &r, int *ln,
                                             float *le) {
  return trace_any<SPILL>(sc.bvh, make_tray(r.o, r.d, r.tmin, r.tmax), ln, le);
}

#define AJG_VCM_LDS                                  \
  __shared__ int s_node[kLdsStack * kTravThreads];   \
  __shared__ float s_entry[kLdsStack * kTravThreads]; \
  int *const ln = s_node + threadIdx.x;              \
  float *const le = s_entry + threadIdx.x

struct VcmSub {
  f3 origin, direction, throughput;
  uint32_t path_length, is_finite_light, specular_path;
  float dVCM, dVC, dVM;
};

__device__ __forceinline__ int vcm_pick(const VcmParams &P, Pcg32 &rng) {
  const float u = pcg_next_float(rng);
  uint32_t lo = 0, len = P.n_lights + 1;
  while (len > 0) {
    const uint32_t half = len / 2;
    if (!(u <= P.sel_cdf[lo + half])) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return clampi((int)lo - 1, 0, (int)P.n_lights - 1);
}

// Light::SampleEmit (area_light.cpp:60-73, directional_light.cpp:60-79,
// env_light.cpp:105-147)
struct VcmEmit {
  f3 pos, dir, normal, radiance;
  float pdf_pos, pdf_dir;
};
__device__ VcmEmit vcm_sample_emit(const VcmParams &P, const DevLight &l, Pcg32 &rng) {
  VcmEmit r;
  r.pos = r.dir = r.normal = r.radiance = mk3(0.f, 0.f, 0.f);
  r.pdf_pos = r.pdf_dir = 0.f;
  const f3 center = f3a(P.world_c);
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
    r.dir = to_world(make_frame(ng), cosine_hemisphere(s0, s1, &r.pdf_dir));
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

// SampleScattering (vcm.cpp:409-451)
__device__ bool vcm_scatter(const VcmParams &P, VcmSub &st, Pcg32 &rng, const Bsdf &b,
                            const Inct &it) {
  f3 sam;
  sam.x = pcg_next_float(rng);
  sam.y = pcg_next_float(rng);
  sam.z = pcg_next_float(rng);
  BsdfSample s = bsdf_sample_all(b, neg(st.direction), sam);
  if (!s.valid) {
    s.f = mk3(0.f, 0.f, 0.f);
    s.pdf = 0.f;
    s.dir = mk3(0.f, 0.f, 0.f);
  }
  if (is_black(s.f)) return false;
  st.direction = s.dir;
  float dir_pdf = s.pdf;
  float rev_pdf = bsdf_pdf_all(b, neg(st.direction), s.dir);
  if (pcg_next_float(rng) > P.cont_prob) return false;
  dir_pdf *= P.cont_prob;
  rev_pdf *= P.cont_prob;
  const float cos_out = dot(it.ng, s.dir);
  st.dVC = (cos_out / dir_pdf) * (st.dVC * rev_pdf + st.dVCM + P.mis_vm);
  st.dVM = (cos_out / dir_pdf) * (st.dVM * rev_pdf + st.dVCM * P.mis_vc + 1.f);
  st.dVCM = 1.f / dir_pdf;
  st.specular_path &= 0u;
  st.origin = it.pos;
  st.throughput = mul(st.throughput, scale(s.f, cos_out / dir_pdf));
  return true;
}

// ThinLensCamera::ToRaster (thin_lens.cpp:85-119)
__device__ bool vcm_to_raster(const VcmParams &P, float u0, float u1, f3 pos, f3 *dir, float *rx,
                              float *ry) {
  const f3 cam_coord = xform_point(P.w2c, pos);
  f3 raster;
  if (P.cam.lens_radius > 0 && P.cam.focal_distance > 0) {
    f2 lens = disk_concentric(u0, u1);
    lens.x = P.cam.lens_radius * lens.x;
    lens.y = P.cam.lens_radius * lens.y;
    const f3 origin = mk3(lens.x, lens.y, 0.f);
    const f3 direction = normalize(sub(cam_coord, origin));
    const float ft = P.cam.focal_distance / direction.z;
    const f3 focus = add(origin, scale(direction, ft));
    raster = xform_point(P.c2r, focus);
    if (!(raster.x < P.res_x && raster.x >= 0.f && raster.y < P.res_y && raster.y >= 0.f))
      return false;
    *dir = sub(xform_point(P.cam.c2w, origin), pos);
  } else {
    raster = xform_point(P.w2r, pos);
    if (!(raster.x < P.res_x && raster.x >= 0.f && raster.y < P.res_y && raster.y >= 0.f))
      return false;
    *dir = sub(f3a(P.cam_pos), pos);
  }
  *rx = raster.x;
  *ry = raster.y;
  return !(dot(f3a(P.cam_dir), neg(*dir)) <= 0.f);
}

// ConnectToCamera (vcm.cpp:544-600)
template <int SPILL>
__device__ void vcm_connect_camera(const DevScene &sc, const VcmParams &P, const VcmState &S,
                                   const VcmSub &st, Pcg32 &rng, const Bsdf &b, const Inct &it,
                                   int *ln, float *le) {
  const float u0 = pcg_next_float(rng);
  const float u1 = pcg_next_float(rng);
  f3 dir;
  float rx, ry;
  if (!vcm_to_raster(P, u0, u1, it.pos, &dir, &rx, &ry)) return;
  const float distance = length(dir);
  dir = normalize(dir);
  const f3 f = bsdf_eval_all(b, dir, it.wr);
  if (is_black(f)) return;
  float rev_pdf = bsdf_pdf_all(b, it.wr, dir);
  rev_pdf *= P.cont_prob;
  const float cos_to_cam = dot(it.ng, dir);
  // ImageAreaToSurfaceArea (thin_lens.cpp:121-135)
  const float cos_at_camera = dot(f3a(P.cam_dir), neg(normalize(dir)));
  const float dimg = P.cam.focal_distance / cos_at_camera;
  const float i2s = dimg * dimg / cos_at_camera * fabsf(cos_to_cam) / (distance * distance);
  const float w_light = (i2s / P.light_count) * (P.mis_vm + st.dVCM + st.dVC * rev_pdf);
  const float mis = 1.f / (w_light + 1.f);
  const float s2i = 1.f / i2s;
  const f3 contrib = divs(mul(scale(st.throughput, mis), f), P.light_count * s2i);
  if (is_black(contrib)) return;
  if (vcm_occluded<SPILL>(sc, make_ray(it.pos, dir, AJG_EPS, distance * 0.9f), ln, le)) return;
  const size_t px = 3 * ((size_t)(int)ry * P.W + (size_t)(int)rx);
  atomicAdd(&S.splat[px], contrib.x);
  atomicAdd(&S.splat[px + 1], contrib.y);
  atomicAdd(&S.splat[px + 2], contrib.z);
}

template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_vcm_light(DevScene sc, VcmParams P, VcmState S) {
  AJG_VCM_LDS;
  const uint32_t i = blockIdx.x * kTravThreads + threadIdx.x;
  if (i >= P.n) return;
  Pcg32 rng;
  pcg_seed(rng, sample_seed(P.seed_base, i, 2u * P.iter));
  VcmSub st;
  // GenerateLightSample (vcm.cpp:670-701)
  const int idx = vcm_pick(P, rng);
  const DevLight &l = sc.lights[idx];
  const VcmEmit em = vcm_sample_emit(P, l, rng);
  st.throughput = scale(em.radiance, fabsf(dot(em.normal, em.dir)));
  st.origin = em.pos;
  st.direction = em.dir;
  const float emission_pdf = em.pdf_pos * em.pdf_dir * P.sel_pdf;
  st.throughput = divs(st.throughput, emission_pdf);
  st.path_length = 1;
  st.is_finite_light = l.type == kLightArea ? 1u : 0u;
  st.specular_path = 0;
  st.dVCM = 1.f / em.pdf_dir;
  if (l.type != AJG_LIGHT_DIRECTIONAL) {
    const float used_cos = st.is_finite_light ? fabsf(dot(em.normal, em.dir)) : 1.f;
    st.dVC = used_cos / emission_pdf;
  } else {
    st.dVC = 0.f;
  }
  st.dVM = st.dVC * P.mis_vc;
  uint32_t nv = 0;
  float4 *slot = S.verts + (size_t)i * P.vcap * kVtx;
  for (;; ++st.path_length) {
    const DevRay ray = make_ray(st.origin, st.direction, AJG_EPS, __int_as_float(0x7f800000));
    const THit h = vcm_closest<SPILL>(sc, ray, ln, le);
    if (h.slot < 0) break;
    const Inct it = post_intersect(sc, ray, h.t, h.slot, h.u, h.v);
    const Bsdf b = make_bsdf(sc.mats[it.material], it.ng, it.ns);
    (void)pcg_next_float(rng), (void)pcg_next_float(rng), (void)pcg_next_float(rng);
    if (st.path_length > 1 || st.is_finite_light) st.dVCM *= h.t * h.t;
    const float cos_in = fabsf(dot(it.ns, neg(ray.d)));
    st.dVCM /= fabsf(cos_in);
    st.dVC /= fabsf(cos_in);
    st.dVM /= fabsf(cos_in);
    if ((P.use_vc || P.use_vm) && nv < P.vcap) {
      float4 *v = slot + (size_t)nv * kVtx;
      v[0] = make_float4(it.pos.x, it.pos.y, it.pos.z, __uint_as_float(st.path_length));
      v[1] = make_float4(it.wr.x, it.wr.y, it.wr.z, __uint_as_float(it.material));
      v[2] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, st.dVCM);
      v[3] = make_float4(it.ns.x, it.ns.y, it.ns.z, st.dVC);
      v[4] = make_float4(it.ng.x, it.ng.y, it.ng.z, st.dVM);
      ++nv;
    }
    if (P.use_vc && (int)st.path_length + 1 >= P.min_len)
      vcm_connect_camera<SPILL>(sc, P, S, st, rng, b, it, ln, le);
    if ((int)st.path_length + 2 > P.max_len) break;
    if (!vcm_scatter(P, st, rng, b, it)) break;
  }
  S.n_verts[i] = nv;
}

// ordered keys of floats (monotonic in the value) for atomic min / max
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void __launch_bounds__(256)
k_vcm_clear(VcmParams P, VcmState S) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) {
    for (int k = 0; k < 3; ++k) {
      S.bbox[k] = 0xffffffffu;
      S.bbox[3 + k] = 0u;
    }
  }
  if (i <= P.n) S.cell_count[i] = 0u;
  if (i < P.n) S.cell_cursor[i] = 0u;
}

__global__ void __launch_bounds__(256)
k_vcm_bbox(VcmParams P, VcmState S) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
  if (i < P.n) {
    const uint32_t nv = S.n_verts[i];
    for (uint32_t k = 0; k < nv; ++k) {
      const float4 p = S.verts[((size_t)i * P.vcap + k) * kVtx];
      const uint32_t c[3] = {fkey(p.x), fkey(p.y), fkey(p.z)};
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], c[a]);
        mx[a] = max(mx[a], c[a]);
      }
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off > 0; off >>= 1) {
      mn[a] = min(mn[a], (uint32_t)__shfl_down((int)mn[a], off, 64));
      mx[a] = max(mx[a], (uint32_t)__shfl_down((int)mx[a], off, 64));
    }
  __shared__ uint32_t wk[4][6];  // one atomic set per block
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      wk[threadIdx.x >> 6][a] = mn[a];
      wk[threadIdx.x >> 6][3 + a] = mx[a];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      const uint32_t m = min(min(wk[0][a], wk[1][a]), min(wk[2][a], wk[3][a]));
      const uint32_t M = max(max(wk[0][3 + a], wk[1][3 + a]), max(wk[2][3 + a], wk[3][3 + a]));
      if (m != 0xffffffffu) atomicMin(&S.bbox[a], m);
      if (M != 0u) atomicMax(&S.bbox[3 + a], M);
    }
  }
}

struct VcmGrid {
  f3 mn, mx;
  float inv_cell;
  uint32_t n_cells;
  bool empty;
};
__device__ __forceinline__ VcmGrid vcm_grid(const VcmParams &P, const VcmState &S) {
  VcmGrid g;
  g.empty = S.bbox[0] == 0xffffffffu;
  g.mn = g.empty ? mk3(FLT_MAX, FLT_MAX, FLT_MAX)
                 : mk3(fkey_inv(S.bbox[0]), fkey_inv(S.bbox[1]), fkey_inv(S.bbox[2]));
  g.mx = g.empty ? mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX)
                 : mk3(fkey_inv(S.bbox[3]), fkey_inv(S.bbox[4]), fkey_inv(S.bbox[5]));
  g.inv_cell = 1.f / (P.radius * 2.f);
  g.n_cells = P.n;
  return g;
}
// HashGrid::GetCellIndex (hashgrid.h:144-164)
__device__ __forceinline__ uint32_t vcm_cell(const VcmGrid &g, int x, int y, int z) {
  const uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
  return ((ux * 73856093u) ^ (uy * 19349663u) ^ (uz * 83492791u)) % g.n_cells;
}
__device__ __forceinline__ uint32_t vcm_cell_of(const VcmGrid &g, f3 p) {
  const f3 d = sub(p, g.mn);
  return vcm_cell(g, (int)floorf(g.inv_cell * d.x), (int)floorf(g.inv_cell * d.y),
                  (int)floorf(g.inv_cell * d.z));
}

template <bool FILL>
__global__ void __launch_bounds__(256)
k_vcm_grid(VcmParams P, VcmState S) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  const VcmGrid g = vcm_grid(P, S);
  const uint32_t nv = S.n_verts[i];
  for (uint32_t k = 0; k < nv; ++k) {
    const uint32_t slot = i * P.vcap + k;
    const float4 p = S.verts[(size_t)slot * kVtx];
    const uint32_t c = vcm_cell_of(g, mk3(p.x, p.y, p.z));
    if (!FILL)
      atomicAdd(&S.cell_count[c], 1u);
    else
      S.indices[S.cell_start[c] + atomicAdd(&S.cell_cursor[c], 1u)] = slot;
  }
}

// Light::Sample as DirectIllumination reads it (area_light.cpp:37-58,
// env_light.cpp:68-103, directional_light.cpp:43-58)
struct VcmLSample {
  f3 ref, pos, normal, radiance;
  float pdf, emit_pdf;
};
__device__ VcmLSample vcm_light_sample(const DevScene &sc, const DevLight &l, f3 ref, Pcg32 &rng) {
  VcmLSample r;
  r.ref = r.pos = r.normal = r.radiance = mk3(0.f, 0.f, 0.f);
  r.pdf = r.emit_pdf = 0.f;
  const float radius = sc.radius;
  if (l.type == kLightArea) {
    f3 sam;
    sam.x = pcg_next_float(rng);
    sam.y = pcg_next_float(rng);
    sam.z = pcg_next_float(rng);
    f3 pos, ng;
    float pdf_area;
    area_geom_sample(l, sam, &pos, &ng, &pdf_area);
    if (dot(ng, sub(ref, pos)) <= 0) return r;
    const f3 i2r = sub(ref, pos);
    const float dist2 = dot(i2r, i2r);
    r.ref = ref;
    r.pos = pos;
    r.normal = ng;
    r.radiance = l.radiance;
    r.pdf = pdf_area * dist2 / fabsf(dot(ng, i2r));
    r.emit_pdf = pdf_area * dot(ng, normalize(i2r)) / AJG_PI;
  } else if (l.type == AJG_LIGHT_ENV) {
    const float first = pcg_next_float(rng);
    const float second = pcg_next_float(rng);
    f3 dir = mk3(0.f, 0.f, 0.f);
    const LightSample ls = env_sample(l.env, ref, second, first, radius, &dir);
    if (ls.pdf == 0.f && is_black(ls.radiance)) return r;
    r.ref = ref;
    r.pos = ls.pos;
    r.normal = neg(dir);
    r.radiance = ls.radiance;
    r.pdf = ls.pdf;
    r.emit_pdf = ls.pdf / (AJG_PI * radius * radius);
  } else {
    const float u0 = pcg_next_float(rng);
    const float u1 = pcg_next_float(rng);
    const LightSample ls = dir_sample(l, ref, u0, u1, radius);
    r.ref = ref;
    r.pos = ls.pos;
    r.radiance = ls.radiance;
    r.pdf = ls.pdf;
    r.emit_pdf = ls.pdf / (AJG_PI * radius * radius);
  }
  return r;
}

// DirectIllumination (vcm.cpp:604-668)
template <int SPILL>
__device__ f3 vcm_direct(const DevScene &sc, const VcmParams &P, const VcmSub &st, const Inct &it,
                         const Bsdf &b, Pcg32 &rng, int *ln, float *le) {
  const f3 zero = mk3(0.f, 0.f, 0.f);
  const int idx = vcm_pick(P, rng);
  const DevLight &l = sc.lights[idx];
  const float pick = P.sel_pdf;
  const VcmLSample sr = vcm_light_sample(sc, l, it.pos, rng);
  if (is_black(sr.radiance)) return zero;
  const f3 r2l = normalize(sub(sr.pos, sr.ref));
  const f3 bsdf_eval = bsdf_eval_all(b, r2l, it.wr);
  if (is_black(bsdf_eval)) return zero;
  float bsdf_pdf_w = bsdf_pdf_all(b, r2l, it.wr);
  float bsdf_rev_pdf_w = bsdf_pdf_all(b, it.wr, r2l);
  bsdf_pdf_w *= l.type == AJG_LIGHT_DIRECTIONAL ? 0.f : P.cont_prob;
  bsdf_rev_pdf_w *= P.cont_prob;
  const float w_light = bsdf_pdf_w / (pick * sr.pdf);
  const float cos_to_light = dot(it.ns, r2l);
  const float cos_at_light = dot(sr.normal, neg(r2l));
  const float w_camera = (sr.emit_pdf * cos_to_light / (sr.pdf * cos_at_light)) *
                         (P.mis_vm + st.dVCM + st.dVC * bsdf_rev_pdf_w);
  const float mis = 1.f / (w_light + 1.f + w_camera);
  const f3 contrib = scale(mul(sr.radiance, bsdf_eval), mis * cos_to_light / (pick * sr.pdf));
  if (is_black(contrib)) return zero;
  if (vcm_occluded<SPILL>(sc, make_ray(it.pos, r2l, AJG_EPS, length(sub(sr.pos, sr.ref))), ln, le))
    return zero;
  return contrib;
}

struct VcmVertex {
  f3 pos, in_dir, throughput, ns, ng;
  uint32_t path_length, material;
  float dVCM, dVC, dVM;
};
__device__ __forceinline__ VcmVertex vcm_vertex(const VcmState &S, uint32_t slot) {
  const float4 *v = S.verts + (size_t)slot * kVtx;
  const float4 a = v[0], b = v[1], c = v[2], d = v[3], e = v[4];
  VcmVertex r;
  r.pos = mk3(a.x, a.y, a.z);
  r.path_length = __float_as_uint(a.w);
  r.in_dir = mk3(b.x, b.y, b.z);
  r.material = __float_as_uint(b.w);
  r.throughput = mk3(c.x, c.y, c.z);
  r.dVCM = c.w;
  r.ns = mk3(d.x, d.y, d.z);
  r.dVC = d.w;
  r.ng = mk3(e.x, e.y, e.z);
  r.dVM = e.w;
  return r;
}

// ConnectVertices (vcm.cpp:474-542)
template <int SPILL>
__device__ f3 vcm_connect(const DevScene &sc, const VcmParams &P, const VcmSub &cs,
                          const VcmVertex &lv, const Inct &it, const Bsdf &b, int *ln, float *le) {
  const f3 zero = mk3(0.f, 0.f, 0.f);
  f3 direction = sub(lv.pos, it.pos);
  const float dist2 = dot(direction, direction);
  const float distance = length(direction);
  direction = normalize(direction);
  const f3 bsdf_eval = bsdf_eval_all(b, direction, it.wr);
  if (is_black(bsdf_eval)) return zero;
  float cam_dir_pdf = bsdf_pdf_all(b, direction, it.wr);
  float cam_rev_pdf = bsdf_pdf_all(b, it.wr, direction);
  cam_dir_pdf *= P.cont_prob;
  cam_rev_pdf *= P.cont_prob;
  const Bsdf lb = make_bsdf(sc.mats[lv.material], lv.ng, lv.ns);
  const f3 light_eval = bsdf_eval_all(lb, neg(direction), lv.in_dir);
  if (is_black(light_eval)) return zero;
  float light_dir_pdf = bsdf_pdf_all(lb, neg(direction), lv.in_dir);
  float light_rev_pdf = bsdf_pdf_all(lb, lv.in_dir, neg(direction));
  light_dir_pdf *= P.cont_prob;
  light_rev_pdf *= P.cont_prob;
  const float g = dot(lv.ns, neg(direction)) * dot(it.ns, direction) / dist2;
  if (g < 0.f) return zero;
  const float cam_dir_pdf_a = cam_dir_pdf * fabsf(dot(lv.ns, neg(direction))) / dist2;
  const float light_dir_pdf_a = light_dir_pdf * fabsf(dot(it.ns, direction)) / dist2;
  const float w_light = cam_dir_pdf_a * (P.mis_vm + lv.dVCM + lv.dVC * light_rev_pdf);
  const float w_camera = light_dir_pdf_a * (P.mis_vm + cs.dVCM + cs.dVC * cam_rev_pdf);
  const float mis = 1.f / (w_light + 1.f + w_camera);
  const f3 contrib = mul(scale(bsdf_eval, mis * g), light_eval);
  if (is_black(contrib)) return zero;
  if (vcm_occluded<SPILL>(sc, make_ray(it.pos, direction, AJG_EPS, distance * 0.9f), ln, le))
    return zero;
  return contrib;
}

// HashGrid::Process with RangeQuery::Process (hashgrid.h:72-136, vcm.cpp:88-126)
__device__ f3 vcm_merge(const VcmParams &P, const VcmState &S, const VcmGrid &g, const VcmSub &cs,
                        const Inct &it, const Bsdf &b) {
  f3 contrib = mk3(0.f, 0.f, 0.f);
  const f3 q = it.pos;
  const f3 dmin = sub(q, g.mn), dmax = sub(g.mx, q);
  if (minc(dmin) < 0.f || minc(dmax) < 0.f) return contrib;
  const f3 cp = scale(dmin, g.inv_cell);
  const f3 cf = mk3(floorf(cp.x), floorf(cp.y), floorf(cp.z));
  const int px = (int)cf.x, py = (int)cf.y, pz = (int)cf.z;
  const f3 fr = sub(cp, cf);
  const int pxo = px + (fr.x < 0.5f ? -1 : +1);
  const int pyo = py + (fr.y < 0.5f ? -1 : +1);
  const int pzo = pz + (fr.z < 0.5f ? -1 : +1);
  for (int j = 0; j < 8; ++j) {
    const uint32_t ci = vcm_cell(g, (j & 4) ? pxo : px, (j & 2) ? pyo : py, (j & 1) ? pzo : pz);
    for (uint32_t k = S.cell_start[ci], end = S.cell_start[ci + 1]; k < end; ++k) {
      const uint32_t slot = S.indices[k];
      const float4 p0 = S.verts[(size_t)slot * kVtx];
      const f3 d = sub(q, mk3(p0.x, p0.y, p0.z));
      if (!(dot(d, d) <= P.radius_sqr)) continue;
      const VcmVertex lv = vcm_vertex(S, slot);
      const uint32_t len = lv.path_length + cs.path_length;
      if (len > (uint32_t)P.max_len || (int)len < P.min_len) continue;
      const f3 bsdf_eval = bsdf_eval_all(b, lv.in_dir, it.wr);
      if (is_black(bsdf_eval)) continue;
      float cam_dir_pdf = bsdf_pdf_all(b, lv.in_dir, it.wr);
      float cam_rev_pdf = bsdf_pdf_all(b, it.wr, lv.in_dir);
      if (cam_dir_pdf == 0.f || cam_rev_pdf == 0.f) continue;
      cam_dir_pdf *= P.cont_prob;
      cam_rev_pdf *= P.cont_prob;
      const float w_light = lv.dVCM * P.mis_vc + lv.dVM * cam_dir_pdf;
      const float w_camera = cs.dVCM * P.mis_vc + cs.dVM * cam_rev_pdf;
      const float mis = 1.f / (w_light + 1.f + w_camera);
      contrib = add(contrib, mul(scale(bsdf_eval, mis), lv.throughput));
    }
  }
  return contrib;
}

// Geometry::Pdf of an area light, as PdfBdpt reads it (area_light.cpp:87-98)
__device__ __forceinline__ float vcm_area_pdf(const DevLight &l) {
  float apdf = l.kind == AJG_GEOM_CUBE ? l.cube_pdf0 : 1.f / l.area;
  if (l.two_sided) apdf = apdf * 0.5f;
  return apdf;
}

// Four waves per SIMD (at most 128 registers, the rest spilled): measured
// 1.6x faster than the compiler's budget of 221 registers (two waves); three
// waves are 1.2x slower than four, five to eight slower again.
template <int SPILL>
__global__ void __launch_bounds__(kTravThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_vcm_camera(DevScene sc, VcmParams P, VcmState S) {
  AJG_VCM_LDS;
  const uint32_t p = blockIdx.x * kTravThreads + threadIdx.x;
  if (p >= P.n) return;
  Pcg32 rng;
  pcg_seed(rng, sample_seed(P.seed_base, p, 2u * P.iter + 1u));
  const VcmGrid g = vcm_grid(P, S);
  // GenerateCameraSample (vcm.cpp:703-732)
  const int x = (int)p % (int)P.W, y = (int)p / (int)P.W;
  const float u = (float)x + pcg_next_float(rng);
  const float v = (float)y + pcg_next_float(rng);
  const float l0 = pcg_next_float(rng);
  const float l1 = pcg_next_float(rng);
  const DevRay primary = camera_ray(P.cam, u, v, l0, l1);
  const float cos_at_camera = dot(f3a(P.cam_dir), normalize(primary.d));
  const float dimg = P.cam.focal_distance / cos_at_camera;
  const float camera_pdf = dimg * dimg / cos_at_camera;
  VcmSub st;
  st.origin = primary.o;
  st.direction = primary.d;
  st.throughput = mk3(1.f, 1.f, 1.f);
  st.path_length = 1;
  st.is_finite_light = 0;
  st.specular_path = 1;
  st.dVCM = 1.f / camera_pdf;
  st.dVC = 0.f;
  st.dVM = 0.f;
  f3 color = mk3(0.f, 0.f, 0.f);
  for (;; ++st.path_length) {
    const DevRay ray = make_ray(st.origin, st.direction, AJG_EPS, __int_as_float(0x7f800000));
    const THit h = vcm_closest<SPILL>(sc, ray, ln, le);
    if (h.slot < 0) {
      if (sc.env_index >= 0 && (int)st.path_length >= P.min_len)
        color = add(color, mul(st.throughput, env_radiance(sc.lights[sc.env_index].env, ray.d)));
      break;
    }
    const Inct it = post_intersect(sc, ray, h.t, h.slot, h.u, h.v);
    const Bsdf b = make_bsdf(sc.mats[it.material], it.ng, it.ns);
    (void)pcg_next_float(rng), (void)pcg_next_float(rng), (void)pcg_next_float(rng);
    {
      const float cos_in = fabsf(dot(it.ns, neg(ray.d)));
      st.dVCM *= h.t * h.t;
      st.dVCM /= fabsf(cos_in);
      st.dVC /= fabsf(cos_in);
      st.dVM /= fabsf(cos_in);
    }
    if (it.light) {
      if ((int)st.path_length >= P.min_len) {
        const DevLight &l = sc.lights[it.light - 1];
        const f3 radiance = dot(it.ng, it.wr) > 0 ? l.radiance : mk3(0.f, 0.f, 0.f);
        if (!is_black(radiance)) {
          if (st.path_length == 1) {
            color = add(color, mul(st.throughput, radiance));
          } else if (P.use_vm && !P.use_vc) {
            if (st.specular_path) color = add(color, mul(st.throughput, radiance));
          } else {
            const float apdf = vcm_area_pdf(l);
            float direct_pdf_a = apdf;
            float emission_pdf_w = std_max(0.f, dot(it.ng, it.wr)) / AJG_PI;
            emission_pdf_w *= apdf;
            direct_pdf_a *= P.cont_prob;
            emission_pdf_w *= P.cont_prob;
            const float w_camera = direct_pdf_a * st.dVCM + emission_pdf_w * st.dVC;
            const float mis = 1.f / (1.f + w_camera);
            color = add(color, mul(scale(st.throughput, mis), radiance));
          }
        }
      }
      break;
    }
    if ((int)st.path_length >= P.max_len) break;
    if (P.use_vc && (int)st.path_length + 1 >= P.min_len)
      color = add(color, mul(st.throughput, vcm_direct<SPILL>(sc, P, st, it, b, rng, ln, le)));
    if (P.use_vc) {
      const uint32_t nv = S.n_verts[p];
      for (uint32_t k = 0; k < nv; ++k) {
        const VcmVertex lv = vcm_vertex(S, p * P.vcap + k);
        const int len = (int)(lv.path_length + 1 + st.path_length);
        if (len < P.min_len || len > P.max_len) continue;
        color = add(color, mul(mul(st.throughput, lv.throughput),
                               vcm_connect<SPILL>(sc, P, st, lv, it, b, ln, le)));
      }
    }
    if (P.use_vm && !g.empty)
      color = add(color, mul(scale(st.throughput, P.vm_norm), vcm_merge(P, S, g, st, it, b)));
    if (!vcm_scatter(P, st, rng, b, it)) break;
  }
  float *c = S.color + 3 * (size_t)p;
  c[0] += color.x;
  c[1] += color.y;
  c[2] += color.z;
}

__global__ void __launch_bounds__(256)
k_vcm_count_vertices(VcmParams P, VcmState S) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t v = i < P.n ? S.n_verts[i] : 0u;
  for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_down((int)v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(S.ctr, (unsigned long long)v);
}

// cofactor inverse of a column-major 4x4 in double
static void inv4(const double *a, double *inv) {
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] +
           a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] -
           a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] +
           a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] -
            a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] -
           a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] +
           a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] -
           a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] +
            a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] +
           a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] -
           a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] +
            a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] -
            a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] -
           a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] +
           a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] -
            a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] +
            a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  for (int i = 0; i < 16; ++i) inv[i] /= det;
}

// Matrix4::transformPoint / transformVector on the host (same order as the device)
static f3 hxform_point(const float *m, f3 p) {
  const float in[4] = {p.x, p.y, p.z, 1.f};
  float out[4];
  for (int r = 0; r < 4; ++r) {
    float acc = 0.f;
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return divs(mk3(out[0], out[1], out[2]), out[3]);
}
static f3 hxform_vector(const float *m, f3 p) {
  const float in[4] = {p.x, p.y, p.z, 0.f};
  float out[3];
  for (int r = 0; r < 3; ++r) {
    float acc = 0.f;
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return mk3(out[0], out[1], out[2]);
}

// Device buffers of a VCM render, kept on the scene and reused.
struct VcmBuffers {
  uint32_t n = 0, vcap = 0;
  std::vector<void *> bufs;
  VcmState st;
  void release() {
    for (void *p : bufs) (void)hipFree(p);
    bufs.clear();
    n = vcap = 0;
  }
  ~VcmBuffers() { release(); }
};

static int ensure_vcm(ajg_ctx *ctx, VcmBuffers *b, uint32_t n, uint32_t vcap) {
  if (b->n >= n && b->vcap >= vcap && !b->bufs.empty()) return AJG_OK;
  b->release();
  auto al = [&](void **p, size_t bytes, const char *what) {
    int r = dev_alloc(ctx, p, bytes, what);
    if (r == AJG_OK) b->bufs.push_back(*p);
    return r;
  };
  VcmState &s = b->st;
  const size_t N = n;
  int st = 0;
  if ((st = al((void **)&s.verts, sizeof(float4) * kVtx * N * vcap, "vcm light vertices")) ||
      (st = al((void **)&s.n_verts, 4 * N, "vcm vertex counts")) ||
      (st = al((void **)&s.bbox, 64, "vcm grid box")) ||
      (st = al((void **)&s.cell_count, 4 * (N + 1), "vcm grid")) ||
      (st = al((void **)&s.cell_start, 4 * (N + 1), "vcm grid")) ||
      (st = al((void **)&s.cell_cursor, 4 * N, "vcm grid")) ||
      (st = al((void **)&s.indices, 4 * N * vcap, "vcm grid indices")) ||
      (st = al((void **)&s.scan_tmp, 4 * scan_tmp_words((uint32_t)N + 1), "vcm scan")) ||
      (st = al((void **)&s.ctr, 256, "vcm counters"))) {
    b->release();
    return st;
  }
  b->n = n;
  b->vcap = vcap;
  return AJG_OK;
}

}  // namespace ajg

using namespace ajg;

void vcm_free(void *cache) { delete (VcmBuffers *)cache; }

static inline uint32_t vcm_cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

extern "C" int ajg_render_vcm(ajg_scene *sc, const ajg_camera *camera, const ajg_vcm_args *a,
                              float *color_sum, float *splat, ajg_vcm_stats *stats,
                              void *stream) {
  if (!sc || !camera || !a || !color_sum || !splat) return AJG_ERR_INVALID;
  ajg_ctx *ctx = sc->ctx;
  if (int st = injected(ctx)) return st;
  if (sc->insts) return fail(ctx, AJG_ERR_INVALID, "vcm: instanced scenes are not supported");
  const int W = (int)camera->resolution[0], H = (int)camera->resolution[1];
  if (W <= 0 || H <= 0 || (uint64_t)W * (uint64_t)H > (1ull << 24))
    return fail(ctx, AJG_ERR_INVALID, "vcm: film size must be positive and at most 2^24 pixels");
  if (a->iteration < 0 || a->max_path_length < 1 || a->max_path_length > 64 ||
      !(a->base_radius > 0.f) || !std::isfinite(a->radius_alpha))
    return fail(ctx, AJG_ERR_INVALID, "vcm: bad arguments");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, AJG_ERR_HIP, "hipSetDevice");
  hipStream_t s = (hipStream_t)stream;
  const int spill = spill_for_depth(sc->accel->depth);
  if (!spill) return fail(ctx, AJG_ERR_INTERNAL, "vcm: tree too deep for the traversal stack");

  VcmParams P;
  memset(&P, 0, sizeof(P));
  memcpy(P.cam.r2c, camera->raster2camera, sizeof(P.cam.r2c));
  memcpy(P.cam.c2w, camera->camera2world, sizeof(P.cam.c2w));
  P.cam.lens_radius = camera->lens_radius;
  P.cam.focal_distance = camera->focal_distance;
  {  // world2camera, camera2raster, world2raster (double, rounded to float)
    double c2w[16], r2c[16], r2w[16], t[16];
    for (int i = 0; i < 16; ++i) c2w[i] = camera->camera2world[i], r2c[i] = camera->raster2camera[i];
    inv4(c2w, t);
    for (int i = 0; i < 16; ++i) P.w2c[i] = (float)t[i];
    inv4(r2c, t);
    for (int i = 0; i < 16; ++i) P.c2r[i] = (float)t[i];
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r) {
        double acc = 0;
        for (int k = 0; k < 4; ++k) acc += c2w[k * 4 + r] * r2c[c * 4 + k];
        r2w[c * 4 + r] = acc;
      }
    inv4(r2w, t);
    for (int i = 0; i < 16; ++i) P.w2r[i] = (float)t[i];
  }
  const f3 cp = hxform_point(camera->camera2world, mk3(0.f, 0.f, 0.f));
  const f3 cd = normalize(hxform_vector(camera->camera2world, mk3(0.f, 0.f, -1.f)));
  P.cam_pos[0] = cp.x, P.cam_pos[1] = cp.y, P.cam_pos[2] = cp.z;
  P.cam_dir[0] = cd.x, P.cam_dir[1] = cd.y, P.cam_dir[2] = cd.z;
  P.res_x = camera->resolution[0];
  P.res_y = camera->resolution[1];
  P.W = (uint32_t)W;
  P.H = (uint32_t)H;
  P.n = P.W * P.H;
  P.use_vm = a->use_vm ? 1 : 0;
  P.use_vc = a->use_vc ? 1 : 0;
  P.min_len = a->min_path_length;
  P.max_len = a->max_path_length;
  P.cont_prob = a->cont_prob;
  P.seed_base = a->seed_base;
  P.vcap = (uint32_t)std::max(1, a->max_path_length - 1);
  const f3 wmn = ld3(sc->world), wmx = ld3(sc->world + 3);
  box3 wb;
  wb.mn = wmn;
  wb.mx = wmx;
  const f3 wc = box_center(wb);
  P.world_c[0] = wc.x, P.world_c[1] = wc.y, P.world_c[2] = wc.z;
  P.world_r = sc->radius;
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

  VcmBuffers *b = (VcmBuffers *)sc->vcm_cache;
  if (!b) {
    b = new VcmBuffers();
    sc->vcm_cache = b;
  }
  if (int st = ensure_vcm(ctx, b, P.n, P.vcap)) return st;
  VcmState S = b->st;
  S.color = color_sum;
  S.splat = splat;
  const DevScene ds = dev_scene(sc);
  hipEvent_t ev[5];
  for (auto &e : ev) e = nullptr;
  auto done = [&](int code) {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    return code;
  };
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) return done(fail(ctx, AJG_ERR_HIP, "vcm events"));
  const bool timing = (a->flags & AJG_RENDER_TIMING) != 0;
  float ms_light = 0.f, ms_grid = 0.f, ms_camera = 0.f;
  if (hipMemsetAsync(color_sum, 0, 12 * (size_t)P.n, s) != hipSuccess ||
      hipMemsetAsync(splat, 0, 12 * (size_t)P.n, s) != hipSuccess ||
      hipMemsetAsync(S.ctr, 0, 8, s) != hipSuccess)
    return done(fail(ctx, AJG_ERR_HIP, "vcm clear"));
  (void)hipEventRecord(ev[0], s);
  const dim3 gt(vcm_cdiv(P.n, kTravThreads)), g256(vcm_cdiv(P.n + 1, 256));
  const f3 wsize = sub(wmx, wmn);
  for (int iter = 0; iter < a->iteration && P.n_lights; ++iter) {
    if (cancelled(ctx)) {  // polled as each iteration is enqueued
      (void)hipStreamSynchronize(s);
      return done(fail(ctx, AJG_ERR_CANCELLED, "vcm cancelled"));
    }
    // RenderIter (vcm.cpp:161-200)
    P.iter = (uint32_t)iter;
    P.light_count = (float)(W * H);
    float radius = a->base_radius * minc(wsize);
    radius /= std::pow((float)(iter + 1), 0.5f * (1.f - a->radius_alpha));
    radius = std::max(radius, 1e-7f);
    P.radius = radius;
    P.radius_sqr = radius * radius;
    P.vm_norm = 1.f / (P.radius_sqr * AJG_PI * P.light_count);
    const float eta = (AJG_PI * P.radius_sqr) * P.light_count;
    P.mis_vm = a->use_vm ? eta : 0.f;
    P.mis_vc = a->use_vc ? 1.f / eta : 0.f;
    if (iter == a->iteration - 1 && hipMemsetAsync(S.ctr, 0, 8, s) != hipSuccess)
      return done(fail(ctx, AJG_ERR_HIP, "vcm clear"));
    hipLaunchKernelGGL(k_vcm_clear, g256, dim3(256), 0, s, P, S);
    if (timing) (void)hipEventRecord(ev[1], s);
    if (spill == kSpillS)
      hipLaunchKernelGGL(k_vcm_light<kSpillS>, gt, dim3(kTravThreads), 0, s, ds, P, S);
    else
      hipLaunchKernelGGL(k_vcm_light<kSpillL>, gt, dim3(kTravThreads), 0, s, ds, P, S);
    if (timing) (void)hipEventRecord(ev[2], s);
    if (P.use_vm) {
      hipLaunchKernelGGL(k_vcm_bbox, g256, dim3(256), 0, s, P, S);
      hipLaunchKernelGGL(k_vcm_grid<false>, g256, dim3(256), 0, s, P, S);
      scan_u32(S.cell_count, S.cell_start, P.n + 1, false, S.scan_tmp, s);
      hipLaunchKernelGGL(k_vcm_grid<true>, g256, dim3(256), 0, s, P, S);
    }
    if (timing) (void)hipEventRecord(ev[3], s);
    if (spill == kSpillS)
      hipLaunchKernelGGL(k_vcm_camera<kSpillS>, gt, dim3(kTravThreads), 0, s, ds, P, S);
    else
      hipLaunchKernelGGL(k_vcm_camera<kSpillL>, gt, dim3(kTravThreads), 0, s, ds, P, S);
    if (iter == a->iteration - 1)
      hipLaunchKernelGGL(k_vcm_count_vertices, g256, dim3(256), 0, s, P, S);
    if (timing) {
      (void)hipEventRecord(ev[4], s);
      if (hipEventSynchronize(ev[4]) != hipSuccess) return done(fail(ctx, AJG_ERR_HIP, "vcm timing"));
      float t0 = 0.f, t1 = 0.f, t2 = 0.f;
      (void)hipEventElapsedTime(&t0, ev[1], ev[2]);
      (void)hipEventElapsedTime(&t1, ev[2], ev[3]);
      (void)hipEventElapsedTime(&t2, ev[3], ev[4]);
      ms_light += t0;
      ms_grid += t1;
      ms_camera += t2;
    }
  }
  (void)hipEventRecord(ev[4], s);
  unsigned long long nv = 0;
  if (hipMemcpyAsync(&nv, S.ctr, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(ctx, AJG_ERR_HIP, "vcm render"));
  if (hipGetLastError() != hipSuccess) return done(fail(ctx, AJG_ERR_HIP, "vcm kernel launch"));
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->light_vertices = nv;
    stats->paths = (uint64_t)P.n * (uint64_t)std::max(a->iteration, 0);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[4]);
    stats->ms_total = ms;
    stats->ms_light = ms_light;
    stats->ms_grid = ms_grid;
    stats->ms_camera = ms_camera;
  }
  return done(AJG_OK);
}
