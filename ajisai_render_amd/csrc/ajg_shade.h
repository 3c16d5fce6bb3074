// ajg_shade.h — device shading code of the wavefront path tracer: PCG32
// sampler, warps, shading frames, the Lambert (AggregateBSDF + Diffuse) and
// Disney BSDFs, the directional and environment lights and the thin-lens
// camera. Each function restates the reference function cited next to it,
// keeping its operation order (and so its FP32 results, up to the device's
// transcendental functions).
#pragma once

#include "../../include/ajisai_gpu.h"
#include "ajg_common.h"

namespace ajg {

#define AJG_PI 3.141592654f      // Constants<float>::pi (math/constants.h:30)
#define AJG_PI2 1.570796327f
#define AJG_PI4 0.785398163f
#define AJG_INV_PI (1.f / AJG_PI)
#define AJG_EPS 1e-4f             // Ray::Eps (core/ray.h:41)
#define AJG_FLT_EPSILON 1.19209290e-07f

// ---------------------------------------------------------------- PCG32
// random.h:30-69 and RandomSampler (sampler/random_sampler.cpp:28-45)
struct Pcg32 {
  uint64_t state;
};
__device__ __forceinline__ uint32_t pcg_next_u32(Pcg32 &r) {
  const uint64_t x0 = r.state;
  const unsigned rot = (unsigned)(x0 >> 59);
  r.state = x0 * 6364136223846793005ull + 1442695040888963407ull;
  uint64_t x = x0 ^ (x0 >> 18);
  const uint32_t v = (uint32_t)(x >> 27);
  return v >> rot | v << ((-rot) & 31);
}
__device__ __forceinline__ void pcg_seed(Pcg32 &r, uint64_t seed) {
  r.state = seed + 1442695040888963407ull;
  (void)pcg_next_u32(r);
}
__device__ __forceinline__ float pcg_next_float(Pcg32 &r) {
  return __uint_as_float((pcg_next_u32(r) >> 9) | 0x3f800000u) - 1.0f;
}
// Counter-based per-sample seed: splitmix64 of (base, traced pixel, sample).
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t sample_seed(uint32_t base, uint32_t pixel, uint32_t s) {
  return splitmix64(((uint64_t)base << 32) ^ splitmix64(((uint64_t)pixel << 32) | s));
}

// ---------------------------------------------------------------- warps
struct f2 {
  float x, y;
};
// squareToUniformDiskConcentric (warp.h:29-45)
__device__ __forceinline__ f2 disk_concentric(float sx, float sy) {
  const float ox = 2.f * sx - 1.f, oy = 2.f * sy - 1.f;
  float r, phi;
  if (ox == 0 && oy == 0) return f2{0.f, 0.f};
  if (fabsf(ox) > fabsf(oy)) {
    r = ox;
    phi = AJG_PI4 * (oy / ox);
  } else {
    r = oy;
    phi = AJG_PI2 - AJG_PI4 * (ox / oy);
  }
  return f2{r * cosf(phi), r * sinf(phi)};
}
// SquareToCosineHemisphereSample (warp.h:58-70): z = 0 -> 1e-10
__device__ __forceinline__ f3 cosine_hemisphere(float sx, float sy, float *pdf) {
  const f2 p = disk_concentric(sx, sy);
  float r = 0.f;
  r += p.x * p.x;
  r += p.y * p.y;
  float z = sqrtf(std_max(0.0f, 1 - r));
  if (z == 0) z = 1e-10f;
  *pdf = z * AJG_INV_PI;
  return mk3(p.x, p.y, z);
}
// UniformSampleCone / UniformConePDF (warp.h:72-104)
__device__ __forceinline__ f3 uniform_cone(float u0, float u1, float cos_max) {
  const float c = (1.f - u0) + u0 * cos_max;
  const float s = sqrtf(std_max(0.f, 1.f - c * c));
  const float phi = u1 * 2 * AJG_PI;
  return mk3(cosf(phi) * s, sinf(phi) * s, c);
}
__device__ __forceinline__ float uniform_cone_pdf(float cos_max) {
  if (cos_max == 1.0f) return 1.0f;
  return 1.0f / (2.0f * AJG_PI * (1.0f - cos_max));
}

// CoordinateSystem (coordinate_system.h:37-52), Duff et al. 2017
struct Frame {
  f3 dx, dy, dz;
};
__host__ __device__ __forceinline__ Frame make_frame(f3 n) {
  Frame f;
  f.dz = n;
  const float sign = copysignf(1.f, n.z);
  const float a = -1.f / (sign + n.z);
  const float b = n.x * n.y * a;
  f.dx = mk3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  f.dy = mk3(b, sign + n.y * n.y * a, -n.y);
  return f;
}
__device__ __forceinline__ f3 to_local(const Frame &f, f3 v) {
  return mk3(dot(v, f.dx), dot(v, f.dy), dot(v, f.dz));
}
__device__ __forceinline__ f3 to_world(const Frame &f, f3 v) {
  return add(add(scale(f.dx, v.x), scale(f.dy, v.y)), scale(f.dz, v.z));
}

// ------------------------------------------------------------- helpers
// CoefficientSpectrum::Luminance (math/spectrum.h:80-82), in double
__host__ __device__ __forceinline__ float luminance(f3 s) {
  return (float)(0.2126 * (double)s.x + 0.7152 * (double)s.y + 0.0722 * (double)s.z);
}
__device__ __forceinline__ bool is_black(f3 s) { return maxc(s) <= 0; }
// std::lerp (C++20, libstdc++ __lerp)
__device__ __forceinline__ float std_lerp(float a, float b, float t) {
  if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
  if (t == 1) return b;
  const float x = a + t * (b - a);
  return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}
// Lerp template of material/disney.cpp:42-45
__device__ __forceinline__ float lerp_t(float a, float b, float t) { return a * (1 - t) + b * t; }
__device__ __forceinline__ f3 lerp_t3(f3 a, f3 b, float t) {
  return add(scale(a, 1 - t), scale(b, t));
}
// PowerHeuristic (renderer/helper.h:38-42)
__device__ __forceinline__ float power_heuristic(float a, float b) {
  a *= a;
  b *= b;
  return a / (a + b);
}

// ----------------------------------------------------------------- BSDF
struct DevMaterial {
  uint32_t type;
  float base_color[3];
  float subsurface, metallic, specular, specular_tint, roughness, anisotropic, sheen,
      sheen_tint, clearcoat, clearcoat_gloss;
};

struct Bsdf {
  DevMaterial m;
  f3 ng, ns;
  Frame gf, sf;
};

struct BsdfSample {
  f3 dir, f;
  float pdf;
  bool valid;
};

__device__ __forceinline__ Bsdf make_bsdf(const DevMaterial &m, f3 ng, f3 ns) {
  Bsdf b;
  b.m = m;
  b.ng = ng;
  b.ns = ns;
  b.gf = make_frame(ng);
  b.sf = make_frame(ns);
  return b;
}
__device__ __forceinline__ f3 bsdf_albedo(const Bsdf &b) { return ld3(b.m.base_color); }
__device__ __forceinline__ BsdfSample invalid_sample() {
  BsdfSample s;
  s.dir = mk3(0, 0, 0);
  s.f = mk3(0, 0, 0);
  s.pdf = 0.f;
  s.valid = false;
  return s;
}
// BSDF::CauseBlackFringes (bsdf.h:99-106)
__device__ __forceinline__ bool black_fringes(const Bsdf &b, f3 w) {
  return (dot(b.ng, w) > 0) != (dot(b.ns, w) > 0);
}
// CorrectShadingNormal (bsdf.h:42-46)
__device__ __forceinline__ float correct_shading_normal(f3 ng, f3 ns, f3 wi) {
  const float denom = fabsf(dot(ng, wi));
  return denom == 0 ? 1.f : fabsf(dot(ns, wi) / denom);
}

// Lambert: AggregateBSDF with one DiffuseComponent of weight 1
// (aggregate_bsdf.cpp:43-131, diffuse_component.cpp:28-64, diffuse.cpp:35-48),
// black-fringe fallback of bsdf.cpp:30-63.
__device__ __forceinline__ BsdfSample lambert_sample(const Bsdf &b, f3 wo, f3 sam) {
  if (black_fringes(b, wo)) {
    if (dot(b.ng, wo) <= 0.f) return invalid_sample();
    float pdf;
    const f3 lwi = cosine_hemisphere(sam.x, sam.y, &pdf);
    if (pdf < AJG_FLT_EPSILON) return invalid_sample();
    const f3 wi = normalize(to_world(b.gf, lwi));
    const float factor = correct_shading_normal(b.ng, b.ns, wi);
    BsdfSample r;
    r.dir = wi;
    r.f = scale(divs(bsdf_albedo(b), AJG_PI), factor);
    r.pdf = pdf;
    r.valid = true;
    return r;
  }
  const f3 lwo = normalize(to_local(b.sf, wo));
  if (lwo.z <= 0.f) return invalid_sample();
  float pdf;
  f3 lwi = cosine_hemisphere(sam.y, sam.z, &pdf);
  if ((double)pdf < 1e-3) return invalid_sample();
  lwi = normalize(lwi);
  const f3 wi = to_world(b.sf, lwi);
  const float factor = correct_shading_normal(b.ng, b.ns, wi);
  BsdfSample r;
  r.dir = wi;
  r.f = scale(divs(bsdf_albedo(b), AJG_PI), factor);
  r.pdf = pdf * 1.f;
  r.valid = true;
  return r;
}
__device__ __forceinline__ f3 lambert_eval(const Bsdf &b, f3 wi, f3 wo) {
  if (black_fringes(b, wo)) {
    if (dot(b.ng, wi) <= 0.f || dot(b.ng, wo) <= 0.f) return mk3(0, 0, 0);
    return divs(bsdf_albedo(b), AJG_PI);
  }
  const f3 lwi = normalize(to_local(b.sf, wi));
  const f3 lwo = normalize(to_local(b.sf, wo));
  if (lwi.z <= 0.f || lwo.z <= 0.f) return mk3(0, 0, 0);
  return add(mk3(0.f, 0.f, 0.f), divs(bsdf_albedo(b), AJG_PI));
}
__device__ __forceinline__ float lambert_pdf(const Bsdf &b, f3 wi, f3 wo) {
  if (black_fringes(b, wo)) {
    if ((dot(b.ng, wi) > 0.f) != (dot(b.ng, wo) > 0.f)) return 0.f;
    const f3 local = normalize(to_local(b.gf, wi));
    return local.z * AJG_INV_PI;
  }
  const f3 lwi = normalize(to_local(b.sf, wi));
  const f3 lwo = normalize(to_local(b.sf, wo));
  if (lwi.z <= 0.f || lwo.z <= 0.f) return 0.f;
  float weight_sum = 0.f, pdf = 0.f;
  weight_sum += 1.f;
  pdf += 1.f * (lwi.z * AJG_INV_PI);
  return weight_sum > 0 ? pdf / weight_sum : 0.f;
}

// ---- Disney (material/disney.cpp:58-473)
__device__ __forceinline__ float tan_theta(f3 v) {
  const float temp = 1 - v.z * v.z;
  if (temp <= 0.0f) return 0.0f;
  return sqrtf(temp) / v.z;
}
__device__ __forceinline__ float tan_theta2(f3 v) {
  const float tmp = 1.f - v.z * v.z;
  if (tmp <= 0.f) return 0.f;
  return tmp / (v.z * v.z);
}
__device__ __forceinline__ f3 reflect(f3 vi, f3 n) { return add(vi, scale(n, 2 * dot(neg(vi), n))); }
__device__ __forceinline__ float fresnel_conductor(float cosi, float eta, float k) {
  const float tmp = (eta * eta + k * k) * cosi * cosi;
  const float rparl2 = (tmp - (2.f * eta * cosi) + 1) / (tmp + (2.f * eta * cosi) + 1);
  const float tmp_f = eta * eta + k * k;
  const float rperp2 =
      (tmp_f - (2.f * eta * cosi) + cosi * cosi) / (tmp_f + (2.f * eta * cosi) + cosi * cosi);
  return 0.5f * (rparl2 + rperp2);
}
__device__ __forceinline__ float fresnel_schlick(const DevMaterial &m, float cos_d, float nr) {
  const float omc = 1.f - cos_d;
  const float sq = omc * omc;
  const float fresnel = nr + (1.f - nr) * sq * sq * omc;
  const float cond = fresnel_conductor(cos_d, 0.4f, 1.6f);
  return std_lerp(fresnel, cond, m.metallic);
}
__device__ __forceinline__ float fresnel_schlick_coat(float cos_d) {
  const float omc = 1.0f - cos_d;
  const float sq = omc * omc;
  return 0.04f + (1.0f - 0.04f) * sq * sq * omc;
}
__device__ __forceinline__ float ggx_d(f3 wh, float alpha) {
  if (wh.z <= 0.f) return 0.f;
  const float t2 = tan_theta2(wh);
  const float c2 = wh.z * wh.z;
  const float root = alpha / (c2 * (alpha * alpha + t2));
  return AJG_INV_PI * (root * root);
}
__device__ __forceinline__ float smith_g(f3 v, f3 wh, float alpha) {
  const float tt = fabsf(tan_theta(v));
  if (tt == 0.0f) return 1.0f;
  if (dot(v, wh) * v.z <= 0) return 0.0f;
  const float root = alpha * tt;
  return 2.0f / (1.0f + sqrtf(1.0f + root * root));
}
__device__ __forceinline__ float ggx_g(f3 wi, f3 wo, f3 wh, float alpha) {
  return smith_g(wo, wh, alpha) * smith_g(wi, wh, alpha);
}
__device__ __forceinline__ float ggx_pdf_vn(f3 wi, f3 h, float alpha) {
  const float d = ggx_d(h, alpha);
  return smith_g(wi, h, alpha) * fabsf(dot(wi, h)) * d / (fabsf(wi.z) + 1e-4f);
}
// ImportanceSampleGGXVisibleNormalUnit (disney.cpp:346-412); the rational
// fit is evaluated in double precision like the reference's double literals.
__device__ __forceinline__ f2 ggx_vn_unit(float theta_i, float u1, float u2) {
  if (theta_i < 1e-4f) {
    const float r = sqrtf(std_max(u1 / ((1 - u1) + 1e-6f), 0.0f));
    const float sp = sinf(2 * AJG_PI * u2), cp = cosf(2 * AJG_PI * u2);
    return f2{r * cp, r * sp};
  }
  const float tan_i = tanf(theta_i);
  const float a = 1 / tan_i;
  const float g1 = 2.0f / (1.0f + sqrtf(std_max(1.0f + 1.0f / (a * a), 0.0f)));
  float A = 2.0f * u1 / g1 - 1.0f;
  if (fabsf(A) == 1) A -= (A >= 0.0f ? 1.0f : -1.0f) * 1e-4f;
  const float temp = 1.0f / (A * A - 1.0f);
  const float B = tan_i;
  const float D = sqrtf(std_max(B * B * temp * temp - (A * A - B * B) * temp, 0.0f));
  const float sx1 = B * temp - D;
  const float sx2 = B * temp + D;
  f2 slope;
  slope.x = (A < 0.0f || sx2 > 1.0f / tan_i) ? sx1 : sx2;
  float S;
  if (u2 > 0.5f) {
    S = 1.0f;
    u2 = 2.0f * (u2 - 0.5f);
  } else {
    S = -1.0f;
    u2 = 2.0f * (0.5f - u2);
  }
  const double w = (double)u2;
  const float z = (float)((w * (w * (w * (-0.365728915865723) + 0.790235037209296) -
                               0.424965825137544) +
                          0.000152998850436920) /
                         (w * (w * (w * (w * 0.169507819808272 - 0.397203533833404) -
                                    0.232500544458471) +
                               1) -
                          0.539825872510702));
  slope.y = S * z * sqrtf(1.0f + slope.x * slope.x);
  return slope;
}
// GGXSampleVisibleNormal (disney.cpp:314-344)
__device__ __forceinline__ f3 ggx_sample_vn(f3 wi_in, float u1, float u2, float alpha) {
  const f3 wi = normalize(mk3(alpha * wi_in.x, alpha * wi_in.y, wi_in.z));
  float theta = 0, phi = 0;
  if (wi.z < 0.99999f) {
    theta = acosf(wi.z);
    phi = atan2f(wi.y, wi.x);
  }
  const float sp = sinf(phi), cp = cosf(phi);
  const f2 s = ggx_vn_unit(theta, u1, u2);
  f2 r = {cp * s.x - sp * s.y, sp * s.x + cp * s.y};
  r.x *= alpha;
  r.y *= alpha;
  const float norm = 1.f / sqrtf(r.x * r.x + r.y * r.y + 1.0f);
  return mk3(-r.x * norm, -r.y * norm, norm);
}

__device__ __forceinline__ float disney_diffuse(const DevMaterial &m, f3 wi, f3 wo, float idh,
                                                float rough) {
  if (m.metallic == 1.f) return 0.f;
  const float oml = 1.f - fabsf(wi.z);
  const float sql = oml * oml;
  const float omv = 1.f - fabsf(wo.z);
  const float sqv = omv * omv;
  const float fd90 = 0.5f + 2.f * idh * rough;
  return AJG_INV_PI * (1.f + (fd90 - 1.f) * sql * sql * oml) *
         (1.f + (fd90 - 1.f) * sqv * sqv * omv);
}
__device__ __forceinline__ float disney_subsurface(const DevMaterial &m, f3 wi, f3 wo, float idh,
                                                   float rough) {
  if (m.subsurface == 0.0f) return 0.0f;
  const float oml = 1.0f - fabsf(wi.z);
  const float sql = oml * oml;
  const float omv = 1.0f - fabsf(wo.z);
  const float sqv = omv * omv;
  const float fss90 = idh * idh * rough;
  const float S = (1.0f + (fss90 - 1.0f) * sql * sql * oml) *
                  (1.0f + (fss90 - 1.0f) * sqv * sqv * omv);
  return AJG_INV_PI * 1.25f * (S * (1.0f / (fabsf(wo.z) + fabsf(wi.z)) - 0.5f) + 0.5f);
}
__device__ __forceinline__ float disney_specular(const DevMaterial &m, f3 wi, f3 wo, f3 wh,
                                                 float odh, float rough) {
  if (wi.z * wo.z <= 0.f) return 0.f;
  const float D = ggx_d(wh, rough * rough);
  if (D == 0.f) return 0.f;
  const float nr = std_lerp(0.f, 0.08f, m.specular);
  const float F = fresnel_schlick(m, odh, nr);
  const float rg = (0.5f + 0.5f * rough);
  const float G = ggx_g(wi, wo, wh, rg * rg);
  return F * D * G / (4.f * fabsf(wi.z) * fabsf(wo.z));
}
__device__ __forceinline__ float disney_clearcoat(const DevMaterial &m, f3 wi, f3 wo, f3 wh,
                                                  float idh, float gloss) {
  if (m.clearcoat == 0.0f) return 0.0f;
  const float rough = lerp_t(0.005f, 0.1f, gloss);
  const float D = ggx_d(wh, rough);
  if (D == 0.0f) return 0.0f;
  const float F = fresnel_schlick_coat(idh);
  const float G = ggx_g(wo, wi, wh, 0.25f);
  return m.clearcoat * D * F * G / (4.0f * fabsf(wi.z) * fabsf(wo.z));
}
// DisneyBSDF::Eval (disney.cpp:148-184)
__device__ __forceinline__ f3 disney_eval(const Bsdf &b, f3 wi, f3 wo) {
  const DevMaterial &m = b.m;
  const f3 lwi = normalize(to_local(b.sf, wi));
  const f3 lwo = normalize(to_local(b.sf, wo));
  if (fabsf(lwi.z) < AJG_FLT_EPSILON || fabsf(lwo.z) < AJG_FLT_EPSILON) return mk3(0, 0, 0);
  const float rough = clamp_r(m.roughness, 0.02f, 1.f);
  const f3 alb = bsdf_albedo(b);
  const float lum = luminance(alb);
  const f3 untinted = mk3(lum, lum, lum);
  f3 spec_albedo = lerp_t3(lerp_t3(alb, untinted, 1.0f - m.specular_tint), alb, m.metallic);
  const f3 sheen_albedo = lerp_t3(untinted, alb, m.sheen_tint);
  const f3 wh = normalize(add(lwo, lwi));
  const float odh = dot(lwo, wh);
  const float idh = dot(lwi, wh);
  const float omo = 1.f - odh;
  spec_albedo = lerp_t3(spec_albedo, untinted, omo * omo * omo);
  const float F = fresnel_schlick(m, odh, 0.0f);
  const f3 sheen_term = scale(sheen_albedo, F * m.sheen);
  const float dl = lerp_t(disney_diffuse(m, lwi, lwo, idh, rough),
                          disney_subsurface(m, lwi, lwo, idh, rough), m.subsurface);
  const f3 base = scale(add(scale(alb, dl), sheen_term), 1.0f - m.metallic);
  const f3 spec = scale(spec_albedo, disney_specular(m, lwi, lwo, wh, odh, rough));
  const float cc = disney_clearcoat(m, lwi, lwo, wh, idh, m.clearcoat_gloss);
  return add(add(base, spec), mk3(cc, cc, cc));
}
// DisneyBSDF::Pdf (disney.cpp:186-223); the coat term uses the world-space
// wo as the reference does.
__device__ __forceinline__ float disney_pdf(const Bsdf &b, f3 wi, f3 wo) {
  const DevMaterial &m = b.m;
  const f3 lwi = normalize(to_local(b.sf, wi));
  const f3 lwo = normalize(to_local(b.sf, wo));
  const f3 wh = normalize(add(lwo, lwi));
  if (maxc(wh) <= 0.f) return 0.f;
  const float rough = clamp_r(m.roughness, 0.02f, 1.f);
  const float mpdf = ggx_pdf_vn(lwo, wh, rough * rough);
  float pdf = 0.f;
  const float dwhdwi = 1.f / (4.f * fabsf(dot(lwi, wh)));
  const float spec_pdf = mpdf * dwhdwi;
  const float nr = std_lerp(0.f, 0.08f, m.specular);
  const float odh = dot(lwo, wh);
  const float prob_spec = fresnel_schlick(m, odh, nr);
  pdf += spec_pdf * prob_spec;
  pdf += fabsf(lwi.z) * AJG_INV_PI * (1.f - prob_spec);
  if (m.clearcoat > 0.f) {
    const float cw = m.clearcoat / (m.clearcoat + luminance(bsdf_albedo(b)));
    const float fc = fresnel_schlick_coat(fabsf(wo.z));
    const float pc = (fc * cw) / (fc * cw + (1 - fc) * (1 - cw));
    const float coat_rough = std_lerp(0.005f, 0.10f, m.clearcoat_gloss);
    const float chp = ggx_pdf_vn(wo, wh, coat_rough);
    const float coat_pdf = chp * dwhdwi;
    pdf *= 1.0f - pc;
    pdf += coat_pdf * pc;
  }
  return pdf;
}
// DisneyBSDF::Sample (disney.cpp:81-146)
__device__ __forceinline__ BsdfSample disney_sample(const Bsdf &b, f3 wo, f3 sam) {
  const DevMaterial &m = b.m;
  const f3 lwo = normalize(to_local(b.sf, wo));
  f3 lwi = mk3(0, 0, 0);
  bool sample_coat = false;
  float rs0 = sam.x, rs1 = sam.y, rs2 = sam.z;
  const float rough = clamp_r(m.roughness, 0.02f, 1.0f);
  float alpha = rough * rough;
  if (m.clearcoat > 0.f) {
    const float cw = m.clearcoat / (m.clearcoat + luminance(bsdf_albedo(b)));
    const float fc = fresnel_schlick_coat(fabsf(lwo.z));
    const float pc = (fc * cw) / (fc * cw + (1.f - fc) * (1.f - cw));
    if (sam.y < pc) {
      sample_coat = true;
      rs1 /= pc;
      alpha = std_lerp(0.005f, 0.1f, m.clearcoat_gloss);
    } else {
      rs1 = (sam.y - pc) / (1.f - pc);
    }
  }
  // The coat and the base lobe draw their half vector with the same sampler
  // and the same random numbers, only alpha differs: one call serves both
  // (one inlined copy, and no lane waits through the other lobe's draw).
  const f3 lwh = ggx_sample_vn(lwo, rs0, rs1, alpha);
  if (sample_coat) {
    lwi = reflect(neg(lwo), lwh);
  } else {
    const float nr = std_lerp(0.0f, 0.08f, m.specular);
    const float odh = dot(lwo, lwh);
    const float prob_spec = fresnel_schlick(m, odh, nr);
    if (rs2 <= prob_spec) {
      lwi = reflect(neg(lwo), lwh);
    } else {
      float unused;
      lwi = cosine_hemisphere(rs0, rs1, &unused);
      if (lwo.z < 0.0f) lwi.z *= -1.0f;
    }
  }
  if (lwi.z <= 0.0f) return invalid_sample();
  BsdfSample r;
  r.dir = to_world(b.sf, lwi);
  r.f = disney_eval(b, r.dir, wo);
  r.pdf = disney_pdf(b, r.dir, wo);
  r.valid = true;
  return r;
}

__device__ __forceinline__ BsdfSample bsdf_sample_all(const Bsdf &b, f3 wo, f3 sam) {
  return b.m.type == AJG_MAT_DISNEY ? disney_sample(b, wo, sam) : lambert_sample(b, wo, sam);
}
__device__ __forceinline__ f3 bsdf_eval_all(const Bsdf &b, f3 wi, f3 wo) {
  return b.m.type == AJG_MAT_DISNEY ? disney_eval(b, wi, wo) : lambert_eval(b, wi, wo);
}
__device__ __forceinline__ float bsdf_pdf_all(const Bsdf &b, f3 wi, f3 wo) {
  return b.m.type == AJG_MAT_DISNEY ? disney_pdf(b, wi, wo) : lambert_pdf(b, wi, wo);
}

// ---------------------------------------------------------------- lights
struct DevEnv {
  uint32_t w, h;
  const float *tex;        // w*h*3
  const float *cond_cdf;   // h*(w+1)
  const float *cond_func;  // h*w
  const float *cond_int;   // h
  const float *marg_cdf;   // h+1
  const float *marg_func;  // h
  float marg_int;
  float rot, scaler;
};

// Light type of an emissive primitive's AreaLight (the public light types
// are directional and environment).
static constexpr uint32_t kLightArea = 2;

struct DevLight {
  uint32_t type;
  f3 radiance;
  Frame fr;
  float angle_cos;
  DevEnv env;
  // area light: the triangles its geometry's Sample draws from (9 floats
  // each, in the order Sample indexes them), the geometry kind, the TwoSided
  // wrapper, surface area, the quad's first-triangle probability and the
  // cube's pdf of its first triangle (what Cube::Pdf returns)
  const float *stris;
  uint32_t kind, two_sided, n_stri;
  float area, prob, cube_pdf0;
};

static constexpr int kMaxLights = 4;

// HDRTexture::LinearSampleImpl (texture2d/hdrtex.cpp:60-89), Image clamp
// (core/image.h:38-48)
__device__ __forceinline__ f3 tex_texel(const DevEnv &e, int x, int y) {
  x = clampi(x, 0, (int)e.w - 1);
  y = clampi(y, 0, (int)e.h - 1);
  const float *p = e.tex + 3 * ((size_t)x + (size_t)y * e.w);
  return mk3(p[0], p[1], p[2]);
}
__host__ __device__ __forceinline__ void tex_coords(uint32_t w, uint32_t h, float u, float v,
                                                    float *cx, float *cy) {
  const float dimx = (float)w, dimy = (float)h;
  *cx = clamp_r(u * dimx - 0.5f, 0.f, dimx - 1.f);
  *cy = clamp_r(v * dimy - 0.5f, 0.f, dimy - 1.f);
}
__device__ __forceinline__ f3 tex_sample(const DevEnv &e, float u, float v) {
  float cx, cy;
  tex_coords(e.w, e.h, u, v, &cx, &cy);
  const float x0 = floorf(cx), y0 = floorf(cy), x1 = ceilf(cx), y1 = ceilf(cy);
  const float tx = cx - x0, ty = cy - y0;
  const f3 u0v0 = tex_texel(e, (int)x0, (int)y0);
  const f3 u1v1 = tex_texel(e, (int)x1, (int)y1);
  const f3 u0v1 = tex_texel(e, (int)x0, (int)y1);
  const f3 u1v0 = tex_texel(e, (int)x1, (int)y0);
  return add(scale(add(scale(u0v0, 1 - tx), scale(u1v0, tx)), 1 - ty),
             scale(add(scale(u0v1, 1 - tx), scale(u1v1, tx)), ty));
}

#define AJG_INV_2PI (1.f / (2.f * AJG_PI))
#define AJG_TWO_PI (2.f * AJG_PI)

__device__ __forceinline__ void env_uv(const DevEnv &e, f3 d, float *uo, float *vo,
                                       float *theta_o) {
  const f3 dir = normalize(d);
  const float theta = acosf(clamp_r(dir.y, -1.f, 1.f));
  float u = (AJG_PI + atan2f(dir.z, dir.x)) * AJG_INV_2PI;
  float v = theta * AJG_INV_PI;
  u += e.rot / 180 * AJG_PI * AJG_INV_2PI;
  v += 0.f;
  if (u > 1.f) u -= 1.f;
  if (u < 0.f) u += 1.f;
  *uo = u;
  *vo = v;
  *theta_o = theta;
}
// EnvLight::Radiance (env_light.cpp:50-66)
__device__ __forceinline__ f3 env_radiance(const DevEnv &e, f3 d) {
  float u, v, th;
  env_uv(e, d, &u, &v, &th);
  return scale(tex_sample(e, u, v), e.scaler);
}
// Distribution2D::Pdf (distribution.h:136-146) / EnvLight::Pdf (:152-170)
__device__ __forceinline__ float env_pdf(const DevEnv &e, f3 d) {
  float u, v, th;
  env_uv(e, d, &u, &v, &th);
  const float sin_t = sinf(th);
  const int iu = clampi((int)(u * (float)e.w), 0, (int)e.w - 1);
  const int iv = clampi((int)(v * (float)e.h), 0, (int)e.h - 1);
  float p2;
  if (e.cond_int[iv] * e.marg_int == 0.0f)
    p2 = 0.f;
  else
    p2 = (e.cond_func[(size_t)iv * e.w + iu] * e.marg_func[iv]) / (e.cond_int[iv] * e.marg_int);
  return p2 / (2.f * AJG_PI * AJG_PI * sin_t);
}
// Distribution::SampleContinuous (distribution.h:64-81): first cdf[k] >= u
__device__ __forceinline__ float distrib_sample(const float *cdf, const float *func,
                                                float func_int, uint32_t n, float u,
                                                float *pdf, int *off) {
  uint32_t lo = 0, len = n + 1;
  while (len > 0) {
    const uint32_t half = len / 2;
    if (!(u <= cdf[lo + half])) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const int offset = clampi((int)lo - 1, 0, (int)n - 1);
  if (off) *off = offset;
  float du = u - cdf[offset];
  if ((cdf[offset + 1] - cdf[offset]) > 0) du /= (cdf[offset + 1] - cdf[offset]);
  *pdf = (func_int > 0) ? func[offset] / func_int : 0;
  return (offset + du) / (float)n;
}

struct LightSample {
  f3 pos, radiance;
  float pdf;
};

// EnvLight::Sample (env_light.cpp:68-103); u0 is SampleContinuous' first
// argument (u), u1 its second (v).
__device__ __forceinline__ LightSample env_sample(const DevEnv &e, f3 ref, float u0, float u1,
                                                  float radius, f3 *dir_out = nullptr) {
  LightSample r;
  r.pos = mk3(0, 0, 0);
  r.radiance = mk3(0, 0, 0);
  r.pdf = 0.f;
  float pdf_v, pdf_u;
  int iv;
  const float sv = distrib_sample(e.marg_cdf, e.marg_func, e.marg_int, e.h, u1, &pdf_v, &iv);
  const float su = distrib_sample(e.cond_cdf + (size_t)iv * (e.w + 1),
                                  e.cond_func + (size_t)iv * e.w, e.cond_int[iv], e.w, u0,
                                  &pdf_u, nullptr);
  float pdf = pdf_u * pdf_v;
  if (pdf == 0.f) return r;
  float u = su;
  const float v = sv;
  u -= e.rot / 180 * AJG_PI * AJG_INV_2PI;
  if (u > 1.f) u -= 1.f;
  if (u < 0.f) u += 1.f;
  const float phi = u * AJG_TWO_PI;
  const float theta = v * AJG_PI;
  const float sin_t = sinf(theta);
  pdf = sin_t != 0.f ? pdf / (2.f * AJG_PI * AJG_PI * sin_t) : 0.f;
  const f3 dir = mk3(-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi));
  if (dir_out) *dir_out = dir;
  r.pos = add(ref, scale(dir, 2.f * radius));
  r.radiance = tex_sample(e, u, v);
  r.pdf = pdf;
  return r;
}
// DirectionalLight::Sample (directional_light.cpp:43-58)
__device__ __forceinline__ LightSample dir_sample(const DevLight &l, f3 ref, float u0, float u1,
                                                  float radius) {
  LightSample r;
  const f3 ref2light = to_world(l.fr, uniform_cone(u0, u1, l.angle_cos));
  r.pos = add(ref, scale(ref2light, 2.f * radius));
  r.radiance = l.radiance;
  r.pdf = uniform_cone_pdf(l.angle_cos);
  return r;
}

// AreaLight::Sample (area_light.cpp:37-58) over its geometry's Sample
// (quad.cpp:129-160, cube.cpp:145-172, mesh.cpp:251-282) and the TwoSided
// wrapper (two_sided.cpp:62-99), as the reference writes them: the mesh
// draws the vertices indices[t], [t+1], [t+2] for triangle number t (the
// triangles in `stris` are those), the cube reports the pdf of the triangle
// it picked, and the solid-angle pdf divides dist^2 by |n . (ref - pos)|.
// A sample whose emitting side faces away from ref has pdf 0.
// area_geom_sample is the geometry's Sample(pdf, sam) with the TwoSided
// wrapper: a point, its (flipped) normal and the area pdf.
__device__ __forceinline__ void area_geom_sample(const DevLight &l, f3 sam, f3 *pos, f3 *n,
                                                 float *pdf_area) {
  bool back = false;
  if (l.two_sided) {
    back = sam.x < 0.5f;
    sam.x = back ? 2 * sam.x : 2 * (sam.x - 0.5f);
  }
  float sx = sam.x, sy = sam.y;
  if (sx + sy > 1.f) {
    sx = 1.f - sx;
    sy = 1.f - sy;
  }
  const float *t;
  bool cube_pdf = false;
  if (l.kind == AJG_GEOM_QUAD) {
    t = sam.z < l.prob ? l.stris : l.stris + 9;
  } else {
    const int cnt = (int)l.n_stri;
    const int idx = min((int)(sam.z * (float)cnt), cnt - 1);
    t = l.stris + 9 * (size_t)idx;
    cube_pdf = l.kind == AJG_GEOM_CUBE;
  }
  const f3 a = mk3(t[0], t[1], t[2]), b = mk3(t[3], t[4], t[5]), c = mk3(t[6], t[7], t[8]);
  float pdf = cube_pdf ? 1.f / (12.f * length(cross(sub(b, a), sub(c, a))) * 0.5f) : 1.f / l.area;
  *pos = add(add(scale(a, (1 - sx) - sy), scale(b, sx)), scale(c, sy));
  f3 ng = normalize(cross(sub(b, a), sub(c, a)));
  if (l.two_sided) {
    pdf = pdf * 0.5f;
    if (back) ng = scale(ng, -1.f);
  }
  *n = ng;
  *pdf_area = pdf;
}
__device__ __forceinline__ LightSample area_sample(const DevLight &l, f3 ref, f3 sam) {
  LightSample r;
  f3 ng;
  float pdf;
  area_geom_sample(l, sam, &r.pos, &ng, &pdf);
  const f3 i2r = sub(ref, r.pos);
  if (dot(ng, i2r) <= 0) {
    r.radiance = mk3(0.f, 0.f, 0.f);
    r.pdf = 0.f;
    return r;
  }
  const float dist2 = dot(i2r, i2r);
  r.pdf = pdf * dist2 / fabsf(dot(ng, i2r));
  r.radiance = l.radiance;
  return r;
}

// AreaLight::Pdf (area_light.cpp:67-78) with Geometry::Pdf (quad.cpp:163-166,
// cube.cpp:175-183, mesh.cpp:290-293, two_sided.cpp:107-115).
__device__ __forceinline__ float area_pdf(const DevLight &l, f3 ref, f3 pos, f3 n) {
  if (dot(sub(ref, pos), n) <= 0.f) return 0.f;
  float apdf = l.kind == AJG_GEOM_CUBE ? l.cube_pdf0 : 1.f / l.area;
  if (l.two_sided) apdf = apdf * 0.5f;
  const f3 l2r = sub(ref, pos);
  const float dist2 = dot(l2r, l2r);
  const float abscos = fabsf(dot(l2r, n));
  return apdf * (dist2 / abscos);
}

// ---------------------------------------------------------------- camera
// Matrix4::transformPoint / transformVector (math/matrix4.h:115-122) with
// the column-major product order of rectangular_matrix.h:199-212.
__device__ __forceinline__ f3 xform_point(const float *m, f3 p) {
  const float in[4] = {p.x, p.y, p.z, 1.f};
  float out[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return divs(mk3(out[0], out[1], out[2]), out[3]);
}
__device__ __forceinline__ f3 xform_vector(const float *m, f3 p) {
  const float in[4] = {p.x, p.y, p.z, 0.f};
  float out[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return mk3(out[0], out[1], out[2]);
}

struct DevRay {
  f3 o, d;
  float tmin, tmax;
};
// Ray constructor (core/ray.h:33-44)
__device__ __forceinline__ DevRay make_ray(f3 o, f3 d, float tmin, float tmax) {
  DevRay r;
  const f3 dn = normalize(d);
  r.o = add(o, scale(dn, AJG_EPS));
  r.d = normalize(d);
  r.tmin = tmin;
  r.tmax = tmax;
  return r;
}

struct DevCamera {
  float r2c[16];
  float c2w[16];
  float lens_radius, focal_distance;
};
// ThinLensCamera::GenerateRay (camera/thin_lens.cpp:62-83)
__device__ __forceinline__ DevRay camera_ray(const DevCamera &cam, float rx, float ry, float l0,
                                             float l1) {
  const f3 cc = xform_point(cam.r2c, mk3(rx, ry, 1.f));
  f3 o = mk3(0.f, 0.f, 0.f);
  f3 d = normalize(cc);
  if (cam.lens_radius > 0 && cam.focal_distance > 0) {
    f2 lens = disk_concentric(l0, l1);
    lens.x *= cam.lens_radius;
    lens.y *= cam.lens_radius;
    const float ft = cam.focal_distance / d.z;
    const f3 focus = add(o, scale(d, ft));
    o = mk3(lens.x, lens.y, 0.f);
    d = normalize(sub(focus, o));
  }
  return make_ray(xform_point(cam.c2w, o), xform_vector(cam.c2w, d), AJG_EPS,
                  __int_as_float(0x7f800000));
}

}  // namespace ajg
