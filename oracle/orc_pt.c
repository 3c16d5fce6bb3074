/*
 * Path tracer of the oracle — TEST INFRASTRUCTURE (see orc.h).
 * Restates PathTracing::TraceMIS (ajisai/core/renderer/pt.cpp:98-227) with
 * its helpers (renderer/helper.h), the Lambert (AggregateBSDF + Diffuse) and
 * Disney BSDFs, directional and environment lights, the thin-lens camera,
 * the PCG32 sampler and the box-filtered film.
 *
 * Parity status of the Disney BSDF: its pieces are pinned by identities the
 * reference's formulas satisfy (GGX D normalisation, the visible-normal
 * identity of Smith G1 with D, the visible-normal sampler's density, the
 * conductor Fresnel) and Eval by closed forms at normal incidence and in
 * retro-reflection (tests/test_oracle_disney.py); sample/pdf consistency, MC
 * integrals, reciprocity and the world-space wo of the coat pdf
 * (disney.cpp:208-220) are tested in tests/test_oracle_pt.py. No reference
 * output exists, so values between those configurations rest on the pinned
 * pieces.
 *
 * Environment sampling draws its two numbers as the GCC build of the
 * reference evaluates SampleContinuous(Next1D(), Next1D(), ...)
 * (env_light.cpp:71): arguments right to left, so the first draw feeds v.
 * This order is an assumption about the compiler and is unpinned; it does
 * not change any expectation.
 */
#include <pthread.h>
#include <stdlib.h>

#include "orc.h"
#include "orc_math.h"

#define ORC_PI 3.141592654f       /* Constants<float>::pi (constants.h:30) */
#define ORC_PI2 1.570796327f      /* pi2 */
#define ORC_PI4 0.785398163f      /* pi4 */
#define ORC_INV_PI (1.f / ORC_PI)
#define ORC_EPS 1e-4f             /* Ray::Eps (ray.h:41) */

/* ---- PCG32 (random.h:30-69) -------------------------------------------- */
static const uint64_t kPcgMult = 6364136223846793005ull;
static const uint64_t kPcgInc = 1442695040888963407ull;

void orc_pcg32_seed(orc_pcg32 *r, uint64_t seed) {
  r->state = seed + kPcgInc;
  (void)orc_pcg32_next_u32(r);
}
uint32_t orc_pcg32_next_u32(orc_pcg32 *r) {
  uint64_t x = r->state;
  unsigned rot = (unsigned)(x >> 59);
  r->state = x * kPcgMult + kPcgInc;
  x ^= x >> 18;
  uint32_t v = (uint32_t)(x >> 27);
  return v >> rot | v << ((-rot) & 31);
}
float orc_pcg32_next_float(orc_pcg32 *r) {
  uint32_t u = (orc_pcg32_next_u32(r) >> 9) | 0x3f800000u;
  float f;
  memcpy(&f, &u, 4);
  return f - 1.0f;
}
void orc_pcg32_floats(uint64_t seed, uint32_t n, float *out) {
  orc_pcg32 r;
  orc_pcg32_seed(&r, seed);
  for (uint32_t i = 0; i < n; ++i) out[i] = orc_pcg32_next_float(&r);
}

/* Per-sample stream seed (the build's counter-based mode): splitmix64 of
 * (seed_base, pixel, sample). */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static uint64_t sample_seed(uint32_t base, uint32_t pixel, uint32_t s) {
  return splitmix64(((uint64_t)base << 32) ^ splitmix64(((uint64_t)pixel << 32) | s));
}

/* ---- warps (warp.h:29-112) and frames (coordinate_system.h:37-52) ----- */
typedef struct v2 {
  float x, y;
} v2;

static v2 disk_concentric(v2 s) {
  v2 off = {2.f * s.x - 1.f, 2.f * s.y - 1.f};
  float r, phi;
  if (off.x == 0 && off.y == 0) {
    v2 z = {0.f, 0.f};
    return z;
  } else if (fabsf(off.x) > fabsf(off.y)) {
    r = off.x;
    phi = ORC_PI4 * (off.y / off.x);
  } else {
    r = off.y;
    phi = ORC_PI2 - ORC_PI4 * (off.x / off.y);
  }
  v2 out = {r * cosf(phi), r * sinf(phi)};
  return out;
}

static v3 cosine_hemisphere(v2 s, float *pdf) {
  v2 p = disk_concentric(s);
  float r = 0.f;
  r += p.x * p.x;
  r += p.y * p.y;
  float one_minus = 1 - r;
  float z = sqrtf(std_max(0.0f, one_minus));
  if (z == 0) z = 1e-10f;
  if (pdf) *pdf = z * ORC_INV_PI;
  return V3(p.x, p.y, z);
}

static v3 uniform_cone(v2 u, float cos_max) {
  float cos_t = (1.f - u.x) + u.x * cos_max;
  float a = 1.f - cos_t * cos_t;
  float sin_t = sqrtf(std_max(0.f, a));
  float phi = u.y * 2 * ORC_PI;
  return V3(cosf(phi) * sin_t, sinf(phi) * sin_t, cos_t);
}
static float uniform_cone_pdf(float cos_max) {
  if (cos_max == 1.0f) return 1.0f;
  return 1.0f / (2.0f * ORC_PI * (1.0f - cos_max));
}

typedef struct frame {
  v3 dx, dy, dz;
} frame;

static frame make_frame(v3 n) {
  frame f;
  f.dz = n;
  float sign = copysignf(1.f, n.z);
  float a = -1.f / (sign + n.z);
  float b = n.x * n.y * a;
  f.dx = V3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  f.dy = V3(b, sign + n.y * n.y * a, -n.y);
  return f;
}
static v3 to_local(const frame *f, v3 v) {
  return V3(vdot(v, f->dx), vdot(v, f->dy), vdot(v, f->dz));
}
static v3 to_world(const frame *f, v3 v) {
  return vadd(vadd(vscale(f->dx, v.x), vscale(f->dy, v.y)), vscale(f->dz, v.z));
}

/* ---- small helpers ------------------------------------------------------ */
static inline float luminance(v3 s) { /* spectrum.h:80-82, in double */
  return (float)(0.2126 * (double)s.x + 0.7152 * (double)s.y +
                 0.0722 * (double)s.z);
}
static inline int is_black(v3 s) { return vmaxc(s) <= 0; }
static inline float clampf_r(float v, float lo, float hi) { /* std::clamp */
  return v < lo ? lo : (hi < v ? hi : v);
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* std::lerp, C++20 (libstdc++ __lerp) */
static float std_lerp(float a, float b, float t) {
  if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
  if (t == 1) return b;
  const float x = a + t * (b - a);
  return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}
/* Lerp template of disney.cpp:42-45 */
static inline float lerp_d(float a, float b, float t) { return a * (1 - t) + b * t; }
static inline v3 lerp_v(v3 a, v3 b, float t) {
  return vadd(vscale(a, 1 - t), vscale(b, t));
}
static inline float power_heuristic(float a, float b) { /* helper.h:38-42 */
  a *= a;
  b *= b;
  return a / (a + b);
}

/* ---- BSDFs -------------------------------------------------------------- */
typedef struct bsdf {
  const orc_material *m;
  v3 ng, ns;
  frame gframe, sframe;
} bsdf;

typedef struct bsdf_sample {
  v3 dir, f;
  float pdf;
  int valid;
} bsdf_sample;

static bsdf make_bsdf(const orc_material *m, v3 ng, v3 ns) {
  bsdf b;
  b.m = m;
  b.ng = ng;
  b.ns = ns;
  b.gframe = make_frame(ng);
  b.sframe = make_frame(ns);
  return b;
}
static v3 albedo(const bsdf *b) { return vload(b->m->base_color); }

/* BSDF black-fringe handling (bsdf.h:99-106, bsdf.cpp:30-63) */
static int black_fringes(const bsdf *b, v3 w) {
  return (vdot(b->ng, w) > 0) != (vdot(b->ns, w) > 0);
}
static float correct_shading_normal(v3 ng, v3 ns, v3 wi) { /* bsdf.h:42-46 */
  const float denom = fabsf(vdot(ng, wi));
  return denom == 0 ? 1.f : fabsf(vdot(ns, wi) / denom);
}

static bsdf_sample invalid_sample(void) {
  bsdf_sample s = {{0, 0, 0}, {0, 0, 0}, 0.f, 0};
  return s;
}

/* Lambert through AggregateBSDF (aggregate_bsdf.cpp:43-131) with a single
 * DiffuseComponent of weight 1 (diffuse.cpp:35-48, diffuse_component.cpp). */
static bsdf_sample lambert_sample(const bsdf *b, v3 wo, v3 sam) {
  if (black_fringes(b, wo)) {
    if (vdot(b->ng, wo) <= 0.f) return invalid_sample();
    float pdf;
    v2 s2 = {sam.x, sam.y};
    v3 lwi = cosine_hemisphere(s2, &pdf);
    if (pdf < FLT_EPSILON) return invalid_sample();
    v3 wi = vnormalize(to_world(&b->gframe, lwi));
    float factor = correct_shading_normal(b->ng, b->ns, wi);
    bsdf_sample r = {wi, vscale(vdivs(albedo(b), ORC_PI), factor), pdf, 1};
    return r;
  }
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  if (lwo.z <= 0.f) return invalid_sample();
  /* Distribution over the single weight always selects component 0. */
  if (lwo.z < 0) return invalid_sample();
  float pdf;
  v2 s2 = {sam.y, sam.z};
  v3 lwi = cosine_hemisphere(s2, &pdf);
  if ((double)pdf < 1e-3) return invalid_sample();
  v3 f = vdivs(albedo(b), ORC_PI); /* coef_ */
  lwi = vnormalize(lwi);
  pdf *= 1.f;
  v3 wi = to_world(&b->sframe, lwi);
  float factor = correct_shading_normal(b->ng, b->ns, wi);
  bsdf_sample r = {wi, vscale(f, factor), pdf, 1};
  return r;
}
static v3 lambert_eval(const bsdf *b, v3 wi, v3 wo) {
  if (black_fringes(b, wo)) {
    if (vdot(b->ng, wi) <= 0.f || vdot(b->ng, wo) <= 0.f) return V3(0, 0, 0);
    return vdivs(albedo(b), ORC_PI);
  }
  v3 lwi = vnormalize(to_local(&b->sframe, wi));
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  if (lwi.z <= 0.f || lwo.z <= 0.f) return V3(0, 0, 0);
  v3 ret = V3(0.f, 0.f, 0.f);
  ret = vadd(ret, vdivs(albedo(b), ORC_PI));
  return ret;
}
static float lambert_pdf(const bsdf *b, v3 wi, v3 wo) {
  if (black_fringes(b, wo)) {
    if ((vdot(b->ng, wi) > 0.f) != (vdot(b->ng, wo) > 0.f)) return 0.f;
    v3 local = vnormalize(to_local(&b->gframe, wi));
    return local.z * ORC_INV_PI;
  }
  v3 lwi = vnormalize(to_local(&b->sframe, wi));
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  if (lwi.z <= 0.f || lwo.z <= 0.f) return 0.f;
  float weight_sum = 0, pdf = 0;
  weight_sum += 1.f;
  pdf += 1.f * (lwi.z * ORC_INV_PI);
  return weight_sum > 0 ? pdf / weight_sum : 0;
}

/* ---- Disney (material/disney.cpp) -------------------------------------- */
static float tan_theta(v3 v) {
  float temp = 1 - v.z * v.z;
  if (temp <= 0.0f) return 0.0f;
  return sqrtf(temp) / v.z;
}
static float tan_theta2(v3 v) {
  float tmp = 1.f - v.z * v.z;
  if (tmp <= 0.f) return 0.f;
  return tmp / (v.z * v.z);
}
static v3 reflect(v3 vi, v3 n) { return vadd(vi, vscale(n, 2 * vdot(vneg(vi), n))); }

static float fresnel_conductor(float cosi, float eta, float k) {
  float tmp = (eta * eta + k * k) * cosi * cosi;
  float rparl2 = (tmp - (2.f * eta * cosi) + 1) / (tmp + (2.f * eta * cosi) + 1);
  float tmp_f = eta * eta + k * k;
  float rperp2 = (tmp_f - (2.f * eta * cosi) + cosi * cosi) /
                 (tmp_f + (2.f * eta * cosi) + cosi * cosi);
  return 0.5f * (rparl2 + rperp2);
}
static float fresnel_schlick(const orc_material *m, float cos_d, float nr) {
  float omc = 1.f - cos_d;
  float sq = omc * omc;
  float fresnel = nr + (1.f - nr) * sq * sq * omc;
  float cond = fresnel_conductor(cos_d, 0.4f, 1.6f);
  return std_lerp(fresnel, cond, m->metallic);
}
static float fresnel_schlick_coat(float cos_d) {
  float omc = 1.0f - cos_d;
  float sq = omc * omc;
  return 0.04f + (1.0f - 0.04f) * sq * sq * omc;
}
static float ggx_d(v3 wh, float alpha) {
  if (wh.z <= 0.f) return 0.f;
  float t2 = tan_theta2(wh);
  float c2 = wh.z * wh.z;
  float root = alpha / (c2 * (alpha * alpha + t2));
  return ORC_INV_PI * (root * root);
}
static float smith_g(v3 v, v3 wh, float alpha) {
  const float tt = fabsf(tan_theta(v));
  if (tt == 0.0f) return 1.0f;
  if (vdot(v, wh) * v.z <= 0) return 0.0f;
  const float root = alpha * tt;
  return 2.0f / (1.0f + sqrtf(1.0f + root * root));
}
static float ggx_g(v3 wi, v3 wo, v3 wh, float alpha) {
  return smith_g(wo, wh, alpha) * smith_g(wi, wh, alpha);
}
static float ggx_pdf_vn(v3 wi, v3 h, float alpha) {
  float d = ggx_d(h, alpha);
  return smith_g(wi, h, alpha) * fabsf(vdot(wi, h)) * d / (fabsf(wi.z) + 1e-4f);
}
static v2 ggx_vn_unit(float theta_i, float u1, float u2) {
  v2 slope;
  if (theta_i < 1e-4f) {
    float a = u1 / ((1 - u1) + 1e-6f);
    float r = sqrtf(std_max(a, 0.0f));
    float sp = sinf(2 * ORC_PI * u2), cp = cosf(2 * ORC_PI * u2);
    slope.x = r * cp;
    slope.y = r * sp;
    return slope;
  }
  float tan_i = tanf(theta_i);
  float a = 1 / tan_i;
  float g1a = 1.0f + 1.0f / (a * a);
  float g1 = 2.0f / (1.0f + sqrtf(std_max(g1a, 0.0f)));
  float A = 2.0f * u1 / g1 - 1.0f;
  if (fabsf(A) == 1) A -= (A >= 0.0f ? 1.0f : -1.0f) * 1e-4f;
  float temp = 1.0f / (A * A - 1.0f);
  float B = tan_i;
  float dd = B * B * temp * temp - (A * A - B * B) * temp;
  float D = sqrtf(std_max(dd, 0.0f));
  float sx1 = B * temp - D;
  float sx2 = B * temp + D;
  slope.x = (A < 0.0f || sx2 > 1.0f / tan_i) ? sx1 : sx2;
  float S;
  if (u2 > 0.5f) {
    S = 1.0f;
    u2 = 2.0f * (u2 - 0.5f);
  } else {
    S = -1.0f;
    u2 = 2.0f * (0.5f - u2);
  }
  double w = (double)u2;
  float z = (float)((w * (w * (w * (-0.365728915865723) + 0.790235037209296) -
                          0.424965825137544) +
                     0.000152998850436920) /
                    (w * (w * (w * (w * 0.169507819808272 - 0.397203533833404) -
                               0.232500544458471) +
                          1) -
                     0.539825872510702));
  slope.y = S * z * sqrtf(1.0f + slope.x * slope.x);
  return slope;
}
static v3 ggx_sample_vn(v3 wi_in, float u1, float u2, float *pdf, float alpha) {
  v3 wi = vnormalize(V3(alpha * wi_in.x, alpha * wi_in.y, wi_in.z));
  float theta = 0, phi = 0;
  if (wi.z < 0.99999f) {
    theta = acosf(wi.z);
    phi = atan2f(wi.y, wi.x);
  }
  float sp = sinf(phi), cp = cosf(phi);
  v2 s = ggx_vn_unit(theta, u1, u2);
  v2 r = {cp * s.x - sp * s.y, sp * s.x + cp * s.y};
  r.x *= alpha;
  r.y *= alpha;
  float norm = 1.f / sqrtf(r.x * r.x + r.y * r.y + 1.0f);
  v3 h = V3(-r.x * norm, -r.y * norm, norm);
  *pdf = ggx_pdf_vn(wi_in, h, alpha);
  return h;
}

static float disney_diffuse_term(const orc_material *m, v3 wi, v3 wo, float idh,
                                 float rough) {
  if (m->metallic == 1.f) return 0.f;
  float oml = 1.f - fabsf(wi.z);
  float sql = oml * oml;
  float omv = 1.f - fabsf(wo.z);
  float sqv = omv * omv;
  float fd90 = 0.5f + 2.f * idh * rough;
  return ORC_INV_PI * (1.f + (fd90 - 1.f) * sql * sql * oml) *
         (1.f + (fd90 - 1.f) * sqv * sqv * omv);
}
static float disney_subsurface_term(const orc_material *m, v3 wi, v3 wo,
                                    float idh, float rough) {
  if (m->subsurface == 0.0f) return 0.0f;
  float oml = 1.0f - fabsf(wi.z);
  float sql = oml * oml;
  float omv = 1.0f - fabsf(wo.z);
  float sqv = omv * omv;
  float fss90 = idh * idh * rough;
  float S = (1.0f + (fss90 - 1.0f) * sql * sql * oml) *
            (1.0f + (fss90 - 1.0f) * sqv * sqv * omv);
  return ORC_INV_PI * 1.25f *
         (S * (1.0f / (fabsf(wo.z) + fabsf(wi.z)) - 0.5f) + 0.5f);
}
static float disney_specular_term(const orc_material *m, v3 wi, v3 wo, v3 wh,
                                  float odh, float rough) {
  if (wi.z * wo.z <= 0.f) return 0.f;
  float D = ggx_d(wh, rough * rough);
  if (D == 0.f) return 0.f;
  float nr = std_lerp(0.f, 0.08f, m->specular);
  float F = fresnel_schlick(m, odh, nr);
  float rg = (0.5f + 0.5f * rough);
  float G = ggx_g(wi, wo, wh, rg * rg);
  return F * D * G / (4.f * fabsf(wi.z) * fabsf(wo.z));
}
static float disney_clearcoat_term(const orc_material *m, v3 wi, v3 wo, v3 wh,
                                   float idh, float gloss) {
  if (m->clearcoat == 0.0f) return 0.0f;
  float rough = lerp_d(0.005f, 0.1f, gloss);
  float D = ggx_d(wh, rough);
  if (D == 0.0f) return 0.0f;
  float F = fresnel_schlick_coat(idh);
  float G = ggx_g(wo, wi, wh, 0.25f);
  return m->clearcoat * D * F * G / (4.0f * fabsf(wi.z) * fabsf(wo.z));
}

static v3 disney_eval(const bsdf *b, v3 wi, v3 wo) {
  const orc_material *m = b->m;
  v3 lwi = vnormalize(to_local(&b->sframe, wi));
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  if (fabsf(lwi.z) < FLT_EPSILON || fabsf(lwo.z) < FLT_EPSILON) return V3(0, 0, 0);
  float rough = clampf_r(m->roughness, 0.02f, 1.f);
  v3 alb = albedo(b);
  float lum = luminance(alb);
  v3 untinted = V3(lum, lum, lum);
  v3 spec_albedo = lerp_v(lerp_v(alb, untinted, 1.0f - m->specular_tint), alb,
                          m->metallic);
  v3 sheen_albedo = lerp_v(untinted, alb, m->sheen_tint);
  v3 wh = vnormalize(vadd(lwo, lwi));
  float odh = vdot(lwo, wh);
  float idh = vdot(lwi, wh);
  float omo = 1.f - odh;
  spec_albedo = lerp_v(spec_albedo, untinted, omo * omo * omo);
  float F = fresnel_schlick(m, odh, 0.0f);
  v3 sheen_term = vscale(sheen_albedo, F * m->sheen);
  float dl = lerp_d(disney_diffuse_term(m, lwi, lwo, idh, rough),
                    disney_subsurface_term(m, lwi, lwo, idh, rough),
                    m->subsurface);
  v3 base = vscale(vadd(vscale(alb, dl), sheen_term), 1.0f - m->metallic);
  v3 spec = vscale(spec_albedo, disney_specular_term(m, lwi, lwo, wh, odh, rough));
  float cc = disney_clearcoat_term(m, lwi, lwo, wh, idh, m->clearcoat_gloss);
  return vadd(vadd(base, spec), V3(cc, cc, cc));
}

static float disney_pdf(const bsdf *b, v3 wi, v3 wo) {
  const orc_material *m = b->m;
  v3 lwi = vnormalize(to_local(&b->sframe, wi));
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  v3 wh = vnormalize(vadd(lwo, lwi));
  if (vmaxc(wh) <= 0.f) return 0.f;
  float rough = clampf_r(m->roughness, 0.02f, 1.f);
  float mpdf = ggx_pdf_vn(lwo, wh, rough * rough);
  float pdf = 0.f;
  float dwhdwi = 1.f / (4.f * fabsf(vdot(lwi, wh)));
  float spec_pdf = mpdf * dwhdwi;
  float nr = std_lerp(0.f, 0.08f, m->specular);
  float odh = vdot(lwo, wh);
  float prob_spec = fresnel_schlick(m, odh, nr);
  pdf += spec_pdf * prob_spec;
  pdf += fabsf(lwi.z) * ORC_INV_PI * (1.f - prob_spec);
  if (m->clearcoat > 0.f) {
    float cw = m->clearcoat / (m->clearcoat + luminance(albedo(b)));
    float fc = fresnel_schlick_coat(fabsf(wo.z)); /* world-space wo */
    float pc = (fc * cw) / (fc * cw + (1 - fc) * (1 - cw));
    float coat_rough = std_lerp(0.005f, 0.10f, m->clearcoat_gloss);
    float chp = ggx_pdf_vn(wo, wh, coat_rough); /* world-space wo */
    float coat_pdf = chp * dwhdwi;
    pdf *= 1.0f - pc;
    pdf += coat_pdf * pc;
  }
  return pdf;
}

static bsdf_sample disney_sample(const bsdf *b, v3 wo, v3 sam) {
  const orc_material *m = b->m;
  v3 lwo = vnormalize(to_local(&b->sframe, wo));
  v3 lwi = V3(0, 0, 0), lwh;
  int sample_coat = 0;
  float rs[3] = {sam.x, sam.y, sam.z};
  if (m->clearcoat > 0.f) {
    float cw = m->clearcoat / (m->clearcoat + luminance(albedo(b)));
    float fc = fresnel_schlick_coat(fabsf(lwo.z));
    float pc = (fc * cw) / (fc * cw + (1.f - fc) * (1.f - cw));
    if (sam.y < pc) {
      sample_coat = 1;
      rs[1] /= pc;
      float coat_rough = std_lerp(0.005f, 0.1f, m->clearcoat_gloss);
      float coat_pdf;
      lwh = ggx_sample_vn(lwo, rs[0], rs[1], &coat_pdf, coat_rough);
      lwi = reflect(vneg(lwo), lwh);
    } else {
      sample_coat = 0;
      rs[1] = (sam.y - pc) / (1.f - pc);
    }
  }
  if (!sample_coat) {
    float mpdf;
    float rough = clampf_r(m->roughness, 0.02f, 1.0f);
    lwh = ggx_sample_vn(lwo, rs[0], rs[1], &mpdf, rough * rough);
    float nr = std_lerp(0.0f, 0.08f, m->specular);
    float odh = vdot(lwo, lwh);
    float prob_spec = fresnel_schlick(m, odh, nr);
    if (rs[2] <= prob_spec) {
      lwi = reflect(vneg(lwo), lwh);
    } else {
      v2 s2 = {rs[0], rs[1]};
      lwi = cosine_hemisphere(s2, NULL);
      if (lwo.z < 0.0f) lwi.z *= -1.0f;
    }
  }
  if (lwi.z <= 0.0f) return invalid_sample();
  v3 wi = to_world(&b->sframe, lwi);
  bsdf_sample r = {wi, disney_eval(b, wi, wo), disney_pdf(b, wi, wo), 1};
  return r;
}

static bsdf_sample bsdf_sample_all(const bsdf *b, v3 wo, v3 sam) {
  return b->m->type == ORC_MAT_DISNEY ? disney_sample(b, wo, sam)
                                      : lambert_sample(b, wo, sam);
}
static v3 bsdf_eval_all(const bsdf *b, v3 wi, v3 wo) {
  return b->m->type == ORC_MAT_DISNEY ? disney_eval(b, wi, wo)
                                      : lambert_eval(b, wi, wo);
}
static float bsdf_pdf_all(const bsdf *b, v3 wi, v3 wo) {
  return b->m->type == ORC_MAT_DISNEY ? disney_pdf(b, wi, wo)
                                      : lambert_pdf(b, wi, wo);
}

/* ---- Distribution / Distribution2D (utility/distribution.h:34-149) ----- */
typedef struct distrib {
  uint32_t n;
  float *func, *cdf;
  float func_int;
} distrib;

static void distrib_init(distrib *d, const float *f, uint32_t n, int fp64) {
  d->n = n;
  d->func = (float *)malloc(sizeof(float) * n);
  d->cdf = (float *)malloc(sizeof(float) * (n + 1));
  memcpy(d->func, f, sizeof(float) * n);
  float inv = 1.f / (float)n;
  d->cdf[0] = 0.f;
  if (fp64) {
    double acc = 0.0;
    for (uint32_t i = 1; i <= n; ++i) {
      acc += (double)d->func[i - 1] * (double)inv;
      d->cdf[i] = (float)acc;
    }
  } else {
    for (uint32_t i = 1; i <= n; ++i) d->cdf[i] = d->cdf[i - 1] + d->func[i - 1] * inv;
  }
  d->func_int = d->cdf[n];
  if (d->func_int > 0.f) {
    for (uint32_t i = 0; i <= n; ++i) d->cdf[i] /= d->func_int;
  } else {
    for (uint32_t i = 0; i <= n; ++i) d->cdf[i] = (float)i / (float)n;
  }
}
static void distrib_free(distrib *d) {
  free(d->func);
  free(d->cdf);
}
/* first index k in [0, n] with cdf[k] >= u (upper_bound with '<='), or n+1 */
static uint32_t first_ge(const float *cdf, uint32_t count, float u) {
  uint32_t lo = 0, len = count;
  while (len > 0) {
    uint32_t half = len / 2;
    if (!(u <= cdf[lo + half])) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}
static float distrib_sample_cont(const distrib *d, float u, float *pdf, int *off) {
  int k = (int)first_ge(d->cdf, d->n + 1, u);
  int offset = clampi(k - 1, 0, (int)d->n - 1);
  if (off) *off = offset;
  float du = u - d->cdf[offset];
  if ((d->cdf[offset + 1] - d->cdf[offset]) > 0) du /= (d->cdf[offset + 1] - d->cdf[offset]);
  if (pdf) *pdf = (d->func_int > 0) ? d->func[offset] / d->func_int : 0;
  return (offset + du) / (float)d->n;
}

/* ---- scene -------------------------------------------------------------- */
enum { ORC_LIGHT_AREA = 2 }; /* emissive primitive (internal light type) */

typedef struct light_rt {
  uint32_t type;
  v3 dir, radiance;
  /* area light: the geometry's Sample triangles in the order the geometry's
   * Sample indexes them (9 floats each), its kind and TwoSided wrapper,
   * surface area, the quad's first-triangle probability, the cube's pdf of
   * its first triangle */
  uint32_t kind, two_sided, n_stri;
  float *stris;
  float area, prob, cube_pdf0;
  frame fr;
  float angle_cos;
  /* env */
  uint32_t w, h;
  const float *tex;
  float rot, scaler;
  distrib *cond; /* h rows */
  distrib marg;
  box3 aabb;
} light_rt;

struct orc_scene {
  int *prim_light; /* per primitive: its area light, or -1 */
  uint32_t n_tri;
  float *tri;
  orc_node *nodes;
  uint32_t *perm;
  uint32_t *tri_prim, *tri_local;
  orc_primitive *prims;
  uint32_t n_prims;
  orc_material *mats;
  uint32_t n_mats;
  light_rt *lights;
  uint32_t n_lights;
  int env_index;
  box3 world;
};

/* HDRTexture::LinearSampleImpl (hdrtex.cpp:60-89) with Image clamping
 * (image.h:38-48). */
static v3 tex_texel(const light_rt *l, int x, int y) {
  x = clampi(x, 0, (int)l->w - 1);
  y = clampi(y, 0, (int)l->h - 1);
  return vload(l->tex + 3 * ((size_t)x + (size_t)y * l->w));
}
static v3 tex_sample(const light_rt *l, float u, float v) {
  float dimx = (float)l->w, dimy = (float)l->h;
  float cx = u * dimx - 0.5f, cy = v * dimy - 0.5f;
  cx = clampf_r(cx, 0.f, dimx - 1.f);
  cy = clampf_r(cy, 0.f, dimy - 1.f);
  float x0 = floorf(cx), y0 = floorf(cy), x1 = ceilf(cx), y1 = ceilf(cy);
  float tx = cx - x0, ty = cy - y0;
  v3 u0v0 = tex_texel(l, (int)x0, (int)y0);
  v3 u1v1 = tex_texel(l, (int)x1, (int)y1);
  v3 u0v1 = tex_texel(l, (int)x0, (int)y1);
  v3 u1v0 = tex_texel(l, (int)x1, (int)y0);
  return vadd(vscale(vadd(vscale(u0v0, 1 - tx), vscale(u1v0, tx)), 1 - ty),
              vscale(vadd(vscale(u0v1, 1 - tx), vscale(u1v1, tx)), ty));
}

static const float kInv2Pi = 1.f / (2.f * ORC_PI);
static const float kTwoPi = 2.f * ORC_PI;

static void env_uv(const light_rt *l, v3 d, float *uo, float *vo, float *theta_o) {
  v3 dir = vnormalize(d);
  float theta = acosf(clampf_r(dir.y, -1.f, 1.f));
  float u = (ORC_PI + atan2f(dir.z, dir.x)) * kInv2Pi;
  float v = theta * ORC_INV_PI;
  u += l->rot / 180 * ORC_PI * kInv2Pi;
  v += 0.f;
  if (u > 1.f) u -= 1.f;
  if (u < 0.f) u += 1.f;
  *uo = u;
  *vo = v;
  *theta_o = theta;
}
/* EnvLight::Radiance (env_light.cpp:50-66) */
static v3 env_radiance(const light_rt *l, v3 d) {
  float u, v, th;
  env_uv(l, d, &u, &v, &th);
  return vscale(tex_sample(l, u, v), l->scaler);
}
/* Distribution2D::Pdf (distribution.h:136-146) */
static float distrib2d_pdf(const light_rt *l, float u, float v) {
  int iu = clampi((int)(u * (float)l->w), 0, (int)l->w - 1);
  int iv = clampi((int)(v * (float)l->h), 0, (int)l->h - 1);
  if (l->cond[iv].func_int * l->marg.func_int == 0.0f) return 0.f;
  return (l->cond[iv].func[iu] * l->marg.func[iv]) /
         (l->cond[iv].func_int * l->marg.func_int);
}
/* EnvLight::Pdf (env_light.cpp:152-170) */
static float env_pdf(const light_rt *l, v3 d) {
  float u, v, th;
  env_uv(l, d, &u, &v, &th);
  float sin_t = sinf(th);
  return distrib2d_pdf(l, u, v) / (2.f * ORC_PI * ORC_PI * sin_t);
}

typedef struct light_sample {
  v3 pos, radiance;
  float pdf;
} light_sample;

/* EnvLight::Sample (env_light.cpp:68-103) from two uniform numbers (u0 is
 * the number passed as SampleContinuous' first argument). */
static light_sample env_sample_uv(const light_rt *l, v3 ref, float u0, float u1,
                                  v3 *dir_out) {
  light_sample r = {{0, 0, 0}, {0, 0, 0}, 0.f};
  float pdfs[2];
  int iv;
  float sv = distrib_sample_cont(&l->marg, u1, &pdfs[1], &iv);
  float su = distrib_sample_cont(&l->cond[iv], u0, &pdfs[0], NULL);
  float pdf = pdfs[0] * pdfs[1];
  if (dir_out) *dir_out = V3(0, 0, 0);
  if (pdf == 0.f) return r;
  float u = su, v = sv;
  u -= l->rot / 180 * ORC_PI * kInv2Pi;
  if (u > 1.f) u -= 1.f;
  if (u < 0.f) u += 1.f;
  float phi = u * kTwoPi;
  float theta = v * ORC_PI;
  float sin_t = sinf(theta);
  pdf = sin_t != 0.f ? pdf / (2.f * ORC_PI * ORC_PI * sin_t) : 0.f;
  v3 dir = V3(-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi));
  float radius = vlength(vsub(l->aabb.mx, box_center(l->aabb)));
  r.pos = vadd(ref, vscale(dir, 2.f * radius));
  r.radiance = tex_sample(l, u, v);
  r.pdf = pdf;
  if (dir_out) *dir_out = dir;
  return r;
}

/* DirectionalLight::Sample (directional_light.cpp:43-58) */
static light_sample dir_sample(const light_rt *l, v3 ref, v2 u) {
  light_sample r;
  v3 ref2light = to_world(&l->fr, uniform_cone(u, l->angle_cos));
  float pdf = uniform_cone_pdf(l->angle_cos);
  float radius = vlength(vsub(l->aabb.mx, box_center(l->aabb)));
  r.pos = vadd(ref, vscale(ref2light, 2.f * radius));
  r.radiance = l->radiance;
  r.pdf = pdf;
  return r;
}

static int v_is_point(const orc_mesh_view *mv) { (void)mv; return 0; }

static inline uint32_t idx_word(const orc_mesh_view *mv, uint32_t local, int k) {
  uint32_t v;
  memcpy(&v, (const char *)mv->index_buffer + mv->index_offset +
                 (size_t)local * mv->index_stride + 4 * k, 4);
  return v;
}
static inline v3 view_pos(const orc_mesh_view *mv, uint32_t i) {
  float f[3];
  memcpy(f, (const char *)mv->position_buffer + mv->position_offset +
                (size_t)i * mv->position_stride, 12);
  return V3(f[0], f[1], f[2]);
}

/* ---- area lights (AreaLight, area_light.cpp:27-85; Geometry::Sample and
 * Pdf of quad.cpp:129-171, cube.cpp:145-191, mesh.cpp:251-297, TwoSided
 * two_sided.cpp:62-115). Restated as written, including:
 *  - MeshGeo::Sample indexes indices_ with the triangle number itself
 *    (indices_[idx], [idx+1], [idx+2], idx < triangle count), so its
 *    "triangles" are consecutive index entries, and its pdf is 1/area;
 *  - the Cube picks one of its 12 triangles uniformly with the pdf of the
 *    picked triangle, and its Pdf uses the first triangle;
 *  - the solid-angle pdf divides dist^2 by |n . (ref - pos)| (not by the
 *    cosine), here and in Pdf. */
static void area_light_init(light_rt *l, const orc_primitive *pr, const float *flat) {
  l->type = ORC_LIGHT_AREA;
  l->radiance = vload(pr->emission);
  l->kind = pr->kind;
  l->two_sided = pr->two_sided;
  const uint32_t n = (uint32_t)pr->view.index_size;
  if (pr->kind == ORC_GEOM_QUAD) {
    const v3 a = view_pos(&pr->view, 0), b = view_pos(&pr->view, 1),
             c = view_pos(&pr->view, 2), d = view_pos(&pr->view, 3);
    l->n_stri = 2;
    l->stris = (float *)malloc(sizeof(float) * 18);
    vstore(l->stris, a), vstore(l->stris + 3, b), vstore(l->stris + 6, c);
    vstore(l->stris + 9, a), vstore(l->stris + 12, c), vstore(l->stris + 15, d);
    const float abc = vlength(vcross(vsub(b, a), vsub(c, a))) * 0.5f;
    const float acd = vlength(vcross(vsub(c, a), vsub(d, a))) * 0.5f;
    l->area = abc + acd;
    l->prob = abc / l->area;
  } else if (pr->kind == ORC_GEOM_CUBE) {
    l->n_stri = n;
    l->stris = (float *)malloc(sizeof(float) * 9 * n);
    for (uint32_t t = 0; t < n; ++t)   /* logical (a, b, c): the flattened order reversed */
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) l->stris[9 * t + 3 * k + c] = flat[9 * t + 3 * (2 - k) + c];
    const float *t0 = l->stris;
    l->cube_pdf0 = 1.f / (12.f * vlength(vcross(vsub(vload(t0 + 3), vload(t0)),
                                                 vsub(vload(t0 + 6), vload(t0)))) * 0.5f);
  } else {
    l->n_stri = n;
    l->stris = (float *)malloc(sizeof(float) * 9 * n);
    for (uint32_t t = 0; t < n; ++t)
      for (int k = 0; k < 3; ++k) {
        const uint32_t flat_i = t + (uint32_t)k;  /* the index array read at the triangle number */
        const uint32_t w = idx_word(&pr->view, flat_i / 3, flat_i % 3);
        vstore(l->stris + 9 * t + 3 * k, view_pos(&pr->view, w));
      }
    float area = 0.f;
    for (uint32_t t = 0; t < n; ++t) {
      const float *p = flat + 9 * (size_t)t;
      const v3 e0 = vsub(vload(p + 3), vload(p)), e1 = vsub(vload(p + 6), vload(p));
      area = (float)((double)area + 0.5 * (double)vlength(vcross(e0, e1)));
    }
    l->area = area;
  }
}

/* Geometry::Sample of the light's primitive (with the TwoSided wrapper) */
static void area_geom_sample(const light_rt *l, v3 sam, v3 *pos, v3 *ng, float *pdf) {
  int back = 0;
  if (l->two_sided) {
    back = sam.x < 0.5f;
    sam.x = back ? 2 * sam.x : 2 * (sam.x - 0.5f);
  }
  v3 sn = sam;
  if (sn.x + sn.y > 1.f) {
    sn.x = 1.f - sn.x;
    sn.y = 1.f - sn.y;
  }
  const float *t;
  if (l->kind == ORC_GEOM_QUAD) {
    t = sam.z < l->prob ? l->stris : l->stris + 9;
    *pdf = 1.f / l->area;
  } else {
    const int cnt = (int)l->n_stri;
    int idx = (int)(sam.z * (float)cnt);
    if (idx > cnt - 1) idx = cnt - 1;
    t = l->stris + 9 * (size_t)idx;
    if (l->kind == ORC_GEOM_CUBE)
      *pdf = 1.f / (12.f * vlength(vcross(vsub(vload(t + 3), vload(t)),
                                           vsub(vload(t + 6), vload(t)))) * 0.5f);
    else
      *pdf = 1.f / l->area;
  }
  const v3 a = vload(t), b = vload(t + 3), c = vload(t + 6);
  *pos = vadd(vadd(vscale(a, (1 - sn.x) - sn.y), vscale(b, sn.x)), vscale(c, sn.y));
  *ng = vnormalize(vcross(vsub(b, a), vsub(c, a)));
  if (l->two_sided) {
    *pdf = *pdf * 0.5f;
    if (back) *ng = vscale(*ng, -1.f);
  }
}

/* AreaLight::Pdf(ref, pos, normal) (area_light.cpp:67-78) */
static float area_light_pdf(const light_rt *l, v3 ref, v3 pos, v3 n) {
  if (vdot(vsub(ref, pos), n) <= 0.f) return 0.f;
  float apdf = l->kind == ORC_GEOM_CUBE ? l->cube_pdf0 : 1.f / l->area;
  if (l->two_sided) apdf = apdf * 0.5f;
  const v3 l2r = vsub(ref, pos);
  const float dist2 = vdot(l2r, l2r);
  const float abscos = fabsf(vdot(l2r, n));
  return apdf * (dist2 / abscos);
}

orc_scene *orc_scene_create(const orc_primitive *prims, uint32_t n_prims,
                            const orc_material *mats, uint32_t n_mats,
                            const orc_light *lights, uint32_t n_lights,
                            uint32_t search_radius, int32_t env_cdf_fp64) {
  (void)v_is_point;
  orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
  orc_mesh_view *views = (orc_mesh_view *)malloc(sizeof(orc_mesh_view) * (n_prims ? n_prims : 1));
  for (uint32_t i = 0; i < n_prims; ++i) views[i] = prims[i].view;
  float box[6];
  s->n_tri = orc_flatten(views, n_prims, NULL, box);
  s->tri = (float *)malloc(sizeof(float) * 9 * (s->n_tri ? s->n_tri : 1));
  orc_flatten(views, n_prims, s->tri, box);
  free(views);
  s->world.mn = vload(box);
  s->world.mx = vload(box + 3);
  s->tri_prim = (uint32_t *)malloc(sizeof(uint32_t) * (s->n_tri ? s->n_tri : 1));
  s->tri_local = (uint32_t *)malloc(sizeof(uint32_t) * (s->n_tri ? s->n_tri : 1));
  uint32_t t = 0;
  for (uint32_t p = 0; p < n_prims; ++p)
    for (size_t i = 0; i < prims[p].view.index_size; ++i, ++t) {
      s->tri_prim[t] = p;
      s->tri_local[t] = (uint32_t)i;
    }
  if (s->n_tri > 0) {
    s->nodes = (orc_node *)malloc(sizeof(orc_node) * (2 * s->n_tri - 1));
    s->perm = (uint32_t *)malloc(sizeof(uint32_t) * s->n_tri);
    if (orc_ploc_build(s->tri, s->n_tri, box, search_radius, s->nodes, s->perm) < 0) {
      orc_scene_destroy(s);
      return NULL;
    }
  }
  s->prims = (orc_primitive *)malloc(sizeof(orc_primitive) * (n_prims ? n_prims : 1));
  memcpy(s->prims, prims, sizeof(orc_primitive) * n_prims);
  s->n_prims = n_prims;
  s->mats = (orc_material *)malloc(sizeof(orc_material) * (n_mats ? n_mats : 1));
  memcpy(s->mats, mats, sizeof(orc_material) * n_mats);
  s->n_mats = n_mats;
  /* DefaultScene (default_scene.cpp:34-48): emissive primitives' area lights
   * first, in primitive order, then the scene's lights */
  uint32_t n_area = 0;
  s->prim_light = (int *)malloc(sizeof(int) * (n_prims ? n_prims : 1));
  for (uint32_t p = 0; p < n_prims; ++p)
    s->prim_light[p] = is_black(vload(prims[p].emission)) ? -1 : (int)n_area++;
  s->n_lights = n_area + n_lights;
  s->lights = (light_rt *)calloc(s->n_lights ? s->n_lights : 1, sizeof(light_rt));
  s->env_index = -1;
  for (uint32_t p = 0, t0 = 0; p < n_prims; t0 += (uint32_t)prims[p].view.index_size, ++p) {
    if (s->prim_light[p] < 0) continue;
    light_rt *l = &s->lights[s->prim_light[p]];
    area_light_init(l, &prims[p], s->tri + 9 * (size_t)t0);
  }
  for (uint32_t i0 = 0; i0 < n_lights; ++i0) {
    const uint32_t i = n_area + i0;
    light_rt *l = &s->lights[i];
    l->type = lights[i0].type;
    l->aabb = s->world; /* Light::Process(world_bound) (default_scene.cpp:47) */
    if (l->type == ORC_LIGHT_DIRECTIONAL) {
      l->dir = vnormalize(vload(lights[i0].dir));
      l->fr = make_frame(l->dir);
      l->angle_cos = cosf(lights[i0].angle);
      l->radiance = vload(lights[i0].radiance);
    } else {
      s->env_index = (int)i;
      l->w = lights[i0].env_w;
      l->h = lights[i0].env_h;
      l->tex = lights[i0].env_texels;
      l->rot = lights[i0].rot;
      l->scaler = lights[i0].scaler;
      /* EnvLight ctor (env_light.cpp:29-48) */
      float *wts = (float *)malloc(sizeof(float) * l->w * l->h);
      for (uint32_t y = 0; y < l->h; ++y) {
        float v = ((float)y + 0.5f) / (float)l->h;
        float sin_t = sinf(ORC_PI * ((float)y + 0.5f) / (float)l->h);
        for (uint32_t x = 0; x < l->w; ++x) {
          float u = ((float)x + 0.5f) / (float)l->w;
          wts[x + l->w * y] = luminance(tex_sample(l, u, v)) * sin_t;
        }
      }
      l->cond = (distrib *)malloc(sizeof(distrib) * l->h);
      float *mf = (float *)malloc(sizeof(float) * l->h);
      for (uint32_t y = 0; y < l->h; ++y) {
        distrib_init(&l->cond[y], wts + (size_t)y * l->w, l->w, env_cdf_fp64);
        mf[y] = l->cond[y].func_int;
      }
      distrib_init(&l->marg, mf, l->h, env_cdf_fp64);
      free(mf);
      free(wts);
    }
  }
  return s;
}

void orc_scene_destroy(orc_scene *s) {
  if (!s) return;
  for (uint32_t i = 0; i < s->n_lights; ++i) {
    light_rt *l = &s->lights[i];
    if (l->type == ORC_LIGHT_ENV && l->cond) {
      for (uint32_t y = 0; y < l->h; ++y) distrib_free(&l->cond[y]);
      free(l->cond);
      distrib_free(&l->marg);
    }
  }
  for (uint32_t i = 0; i < s->n_lights; ++i)
    if (s->lights[i].type == ORC_LIGHT_AREA) free(s->lights[i].stris);
  free(s->lights);
  free(s->prim_light);
  free(s->tri);
  free(s->nodes);
  free(s->perm);
  free(s->tri_prim);
  free(s->tri_local);
  free(s->prims);
  free(s->mats);
  free(s);
}
uint32_t orc_scene_tri_count(const orc_scene *s) { return s->n_tri; }
const orc_node *orc_scene_nodes(const orc_scene *s) { return s->nodes; }
const uint32_t *orc_scene_perm(const orc_scene *s) { return s->perm; }
const float *orc_scene_tris(const orc_scene *s) { return s->tri; }
void orc_scene_world_bound(const orc_scene *s, float *b) {
  vstore(b, s->world.mn);
  vstore(b + 3, s->world.mx);
}

int orc_scene_env_tables(const orc_scene *s, uint32_t li, float *cond_cdf,
                         float *cond_func_int, float *marg_cdf,
                         float *marg_func_int) {
  if (li >= s->n_lights || s->lights[li].type != ORC_LIGHT_ENV) return -1;
  const light_rt *l = &s->lights[li];
  for (uint32_t y = 0; y < l->h; ++y) {
    memcpy(cond_cdf + (size_t)y * (l->w + 1), l->cond[y].cdf, sizeof(float) * (l->w + 1));
    cond_func_int[y] = l->cond[y].func_int;
  }
  memcpy(marg_cdf, l->marg.cdf, sizeof(float) * (l->h + 1));
  *marg_func_int = l->marg.func_int;
  return 0;
}
void orc_env_sample(const orc_scene *s, uint32_t li, float u0, float u1,
                    float *dir3, float *pdf, float *rgb) {
  v3 dir;
  light_sample ls = env_sample_uv(&s->lights[li], V3(0, 0, 0), u0, u1, &dir);
  vstore(dir3, dir);
  *pdf = ls.pdf;
  vstore(rgb, ls.radiance);
}
float orc_env_pdf(const orc_scene *s, uint32_t li, const float *d) {
  return env_pdf(&s->lights[li], vload(d));
}
void orc_env_radiance(const orc_scene *s, uint32_t li, const float *d, float *rgb) {
  vstore(rgb, env_radiance(&s->lights[li], vload(d)));
}

void orc_bsdf_sample(const orc_material *m, const float *ng, const float *ns,
                     const float *wo, const float *sam3, float *wi, float *f,
                     float *pdf, int *valid) {
  bsdf b = make_bsdf(m, vload(ng), vload(ns));
  bsdf_sample r = bsdf_sample_all(&b, vload(wo), vload(sam3));
  vstore(wi, r.dir);
  vstore(f, r.f);
  *pdf = r.pdf;
  *valid = r.valid;
}
/* n samples of one BSDF at one wo (sam3 and wi are n x 3; valid is n). */
void orc_bsdf_sample_batch(const orc_material *m, const float *ng, const float *ns,
                           const float *wo, uint32_t n, const float *sam3, float *wi,
                           int *valid) {
  bsdf b = make_bsdf(m, vload(ng), vload(ns));
  for (uint32_t i = 0; i < n; ++i) {
    bsdf_sample r = bsdf_sample_all(&b, vload(wo), vload(sam3 + 3 * (size_t)i));
    vstore(wi + 3 * (size_t)i, r.dir);
    valid[i] = r.valid;
  }
}
void orc_bsdf_eval(const orc_material *m, const float *ng, const float *ns,
                   const float *wi, const float *wo, float *f) {
  bsdf b = make_bsdf(m, vload(ng), vload(ns));
  vstore(f, bsdf_eval_all(&b, vload(wi), vload(wo)));
}
float orc_bsdf_pdf(const orc_material *m, const float *ng, const float *ns,
                   const float *wi, const float *wo) {
  bsdf b = make_bsdf(m, vload(ng), vload(ns));
  return bsdf_pdf_all(&b, vload(wi), vload(wo));
}
/* Microfacet pieces of the Disney BSDF, exported for the tests that pin them
 * (GGX D, Smith G1 and the visible-normal pdf, disney.cpp:289-299, 350-360,
 * 435-453; the visible-normal sampler, :301-348, 390-412; the conductor
 * Fresnel, :427-436). */
void orc_ggx_parts(uint32_t n, const float *v, const float *h, float alpha, float *d, float *g1,
                   float *pdf_vn) {
  for (uint32_t i = 0; i < n; ++i) {
    const v3 vv = vload(v + 3 * (size_t)i), hh = vload(h + 3 * (size_t)i);
    d[i] = ggx_d(hh, alpha);
    g1[i] = smith_g(vv, hh, alpha);
    pdf_vn[i] = ggx_pdf_vn(vv, hh, alpha);
  }
}
void orc_ggx_sample_vn_batch(uint32_t n, const float *v, const float *u2, float alpha, float *h,
                             float *pdf) {
  for (uint32_t i = 0; i < n; ++i)
    vstore(h + 3 * (size_t)i, ggx_sample_vn(vload(v + 3 * (size_t)i), u2[2 * (size_t)i],
                                            u2[2 * (size_t)i + 1], &pdf[i], alpha));
}
float orc_fresnel_conductor(float cosi, float eta, float k) { return fresnel_conductor(cosi, eta, k); }

void orc_square_to_disk(const float *s2, float *out2) {
  v2 s = {s2[0], s2[1]};
  v2 r = disk_concentric(s);
  out2[0] = r.x;
  out2[1] = r.y;
}
void orc_cosine_hemisphere(const float *s2, float *out3, float *pdf) {
  v2 s = {s2[0], s2[1]};
  vstore(out3, cosine_hemisphere(s, pdf));
}
void orc_frame(const float *n, float *dx, float *dy) {
  frame f = make_frame(vload(n));
  vstore(dx, f.dx);
  vstore(dy, f.dy);
}

/* ---- camera (thin_lens.cpp:32-83, matrix4.h:119-157) ------------------- */
static v3 xform_point(const float *m, v3 p) {
  float in[4] = {p.x, p.y, p.z, 1.f}, out[4];
  for (int r = 0; r < 4; ++r) {
    float acc = 0.f;
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return vdivs(V3(out[0], out[1], out[2]), out[3]);
}
static v3 xform_vector(const float *m, v3 p) {
  float in[4] = {p.x, p.y, p.z, 0.f}, out[3];
  for (int r = 0; r < 3; ++r) {
    float acc = 0.f;
    for (int c = 0; c < 4; ++c) acc += m[c * 4 + r] * in[c];
    out[r] = acc;
  }
  return V3(out[0], out[1], out[2]);
}
/* Ray constructor (ray.h:33-44) */
static orc_ray make_ray(v3 o, v3 d, float tmin, float tmax) {
  orc_ray r;
  v3 dn = vnormalize(d);
  vstore(r.o, vadd(o, vscale(dn, ORC_EPS)));
  vstore(r.d, vnormalize(d));
  r.tmin = tmin;
  r.tmax = tmax;
  return r;
}
void orc_camera_ray(const orc_camera *cam, float rx, float ry,
                    const float *lens2, orc_ray *out) {
  v3 cc = xform_point(cam->raster2camera, V3(rx, ry, 1.f));
  v3 o = V3(0.f, 0.f, 0.f);
  v3 d = vnormalize(cc);
  if (cam->lens_radius > 0 && cam->focal_distance > 0) {
    v2 s = {lens2[0], lens2[1]};
    v2 lens = disk_concentric(s);
    lens.x *= cam->lens_radius;
    lens.y *= cam->lens_radius;
    float ft = cam->focal_distance / d.z;
    v3 focus = vadd(o, vscale(d, ft));
    o = V3(lens.x, lens.y, 0.f);
    d = vnormalize(vsub(focus, o));
  }
  *out = make_ray(xform_point(cam->camera2world, o),
                  xform_vector(cam->camera2world, d), ORC_EPS, INFINITY);
}

static void inv4d(const double *m, double *inv) {
  /* cofactor inverse of a column-major 4x4 */
  double a[16];
  for (int i = 0; i < 16; ++i) a[i] = m[i];
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
  double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  for (int i = 0; i < 16; ++i) inv[i] /= det;
}
static void mul4d(const double *a, const double *b, double *o) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double acc = 0;
      for (int k = 0; k < 4; ++k) acc += a[k * 4 + r] * b[c * 4 + k];
      o[c * 4 + r] = acc;
    }
}
static void norm3d(double *v) {
  double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= l;
  v[1] /= l;
  v[2] /= l;
}
static void cross3d(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - b[1] * a[2];
  o[1] = a[2] * b[0] - b[2] * a[0];
  o[2] = a[0] * b[1] - b[0] * a[1];
}
/* Camera matrices in double, rounded to float: camera2world = from({right,
 * realUp, backward}, eye) (matrix4.h:145-153); camera2screen =
 * perspectiveProjection(fov, aspect, 1, 1000) (matrix4.h:155-170);
 * screen2raster as thin_lens.cpp:43-46. */
void orc_camera_setup(orc_camera *cam, const float *pos, const float *look_at,
                      const float *up, float fov_deg, float res_x, float res_y,
                      float lens_radius, float focal_distance) {
  double eye[3] = {pos[0], pos[1], pos[2]};
  double back[3] = {pos[0] - (double)look_at[0], pos[1] - (double)look_at[1],
                    pos[2] - (double)look_at[2]};
  norm3d(back);
  double upd[3] = {up[0], up[1], up[2]}, right[3], rup[3];
  cross3d(upd, back, right);
  norm3d(right);
  cross3d(back, right, rup);
  norm3d(rup);
  double c2w[16] = {right[0], right[1], right[2], 0, rup[0], rup[1], rup[2], 0,
                    back[0],  back[1],  back[2],  0, eye[0], eye[1], eye[2], 1};
  double fov = (double)fov_deg * 3.14159265358979323846 / 180.0;
  double aspect = (double)res_x / (double)res_y;
  double t = tan(fov * 0.5);
  double sx = 2.0 / (2.0 * t * aspect), sy = -(2.0 / (2.0 * t));
  double zs = 1.0 / (1000.0 - 1.0), m22 = 1.0 * zs, m32 = m22 * 1000.0;
  double c2s[16] = {sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, m22, -1, 0, 0, m32, 0};
  /* screen2raster = S(res) * S(0.5,-0.5,1) * T(1,-1,0) */
  double s2r[16] = {res_x * 0.5, 0, 0, 0, 0, res_y * -0.5, 0, 0, 0, 0, 1, 0,
                    res_x * 0.5 * 1.0, res_y * -0.5 * -1.0, 0, 1};
  double c2s_inv[16], s2r_inv[16], r2c[16];
  inv4d(c2s, c2s_inv);
  inv4d(s2r, s2r_inv);
  mul4d(c2s_inv, s2r_inv, r2c);
  for (int i = 0; i < 16; ++i) {
    cam->raster2camera[i] = (float)r2c[i];
    cam->camera2world[i] = (float)c2w[i];
  }
  cam->resolution[0] = res_x;
  cam->resolution[1] = res_y;
  cam->lens_radius = lens_radius;
  cam->focal_distance = focal_distance;
}

/* ---- TraceMIS (pt.cpp:98-227) ------------------------------------------ */
typedef struct sampler {
  orc_pcg32 rng;
} sampler;
static inline float next1(sampler *s) { return orc_pcg32_next_float(&s->rng); }

typedef struct inct {
  v3 pos, ng, ns, wr;
} inct;

static int scene_intersect(const orc_scene *s, const orc_ray *r, orc_hit *h,
                           uint64_t *q) {
  (*q)++;
  if (s->n_tri == 0) {
    h->t = INFINITY;
    h->id = 0xFFFFFFFFu;
    return 0;
  }
  orc_trace_closest(s->nodes, s->perm, s->tri, r, 1, h);
  return h->id != 0xFFFFFFFFu;
}
static int scene_occlude(const orc_scene *s, const orc_ray *r, uint64_t *q) {
  uint8_t o = 0;
  (*q)++;
  if (s->n_tri == 0) return 0;
  orc_trace_any(s->nodes, s->perm, s->tri, r, 1, &o);
  return o;
}

/* MT normal + PostIntersect (ploc_aggregate.cpp:199-211; quad.cpp:73-90,
 * cube.cpp:96-108, mesh.cpp:183-204, two_sided.cpp:44-53). */
static inct post_intersect(const orc_scene *s, const orc_ray *r, const orc_hit *h) {
  inct it;
  v3 o = vload(r->o), d = vload(r->d);
  const float *p = s->tri + 9 * (size_t)h->id;
  v3 a = vload(p), b = vload(p + 3), c = vload(p + 6);
  v3 n = vnormalize(vcross(vsub(b, a), vsub(c, a)));
  const orc_primitive *pr = &s->prims[s->tri_prim[h->id]];
  uint32_t local = s->tri_local[h->id];
  it.pos = vadd(o, vscale(d, h->t));
  it.wr = vneg(d);
  if (pr->kind == ORC_GEOM_QUAD) {
    it.ng = n;
    it.ns = n;
  } else {
    uint32_t i, j, k;
    if (pr->kind == ORC_GEOM_CUBE) { /* std::tuple stored in reverse order */
      i = idx_word(&pr->view, local, 2);
      j = idx_word(&pr->view, local, 1);
      k = idx_word(&pr->view, local, 0);
    } else {
      i = idx_word(&pr->view, local, 0);
      j = idx_word(&pr->view, local, 1);
      k = idx_word(&pr->view, local, 2);
    }
    v3 pi = view_pos(&pr->view, i), pj = view_pos(&pr->view, j),
       pk = view_pos(&pr->view, k);
    it.ng = vnormalize(vcross(vsub(pj, pi), vsub(pk, pi)));
    if (pr->kind == ORC_GEOM_MESH) {
      const float *N = pr->normal_and_vs;
      float w = 1 - h->u - h->v;
      it.ns = vadd(vadd(vscale(vload(N + 4 * (size_t)i), w),
                        vscale(vload(N + 4 * (size_t)j), h->u)),
                   vscale(vload(N + 4 * (size_t)k), h->v));
    } else {
      it.ns = it.ng;
    }
  }
  if (pr->two_sided) {
    if (vdot(it.ng, it.wr) < 0) {
      it.ng = vscale(it.ng, -1.f);
      it.ns = vscale(it.ns, -1.f);
    }
  }
  return it;
}

static v3 eps_offset(const inct *it, v3 dir) { /* intersection.h:42-47 */
  if (vdot(dir, it->ng) > 0.f) return vadd(it->pos, vscale(it->ng, ORC_EPS));
  return vsub(it->pos, vscale(it->ng, ORC_EPS));
}

/* MISSampleLight / MISSampleEnvLight (helper.h:73-133) */
static v3 mis_sample_light(const orc_scene *s, const light_rt *l, const inct *it,
                           const bsdf *b, sampler *smp, uint64_t *q) {
  light_sample ls;
  v3 zero = V3(0, 0, 0);
  if (l->type == ORC_LIGHT_AREA) {
    /* MISSampleAreaLight (helper.h:44-71) with AreaLight::Sample */
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    v3 pos, ng;
    float pdf_area;
    area_geom_sample(l, sam, &pos, &ng, &pdf_area);
    if (vdot(ng, vsub(it->pos, pos)) <= 0) return zero;
    const v3 i2r = vsub(it->pos, pos);
    const float dist2 = vdot(i2r, i2r);
    const float pdf = pdf_area * dist2 / fabsf(vdot(ng, i2r));
    if (is_black(l->radiance) || pdf <= 0.f) return zero;
    const float len = vlength(vsub(pos, it->pos)) - ORC_EPS;
    if (len < ORC_EPS) return zero;
    const v3 to_light = vnormalize(vsub(pos, it->pos));
    orc_ray sr = make_ray(it->pos, to_light, ORC_EPS, 0.99f * len);
    if (scene_occlude(s, &sr, q)) return zero;
    v3 bf = bsdf_eval_all(b, to_light, it->wr);
    if (is_black(bf)) return zero;
    v3 f = vscale(vmul(l->radiance, bf), fabsf(vdot(to_light, it->ns)));
    float bpdf = bsdf_pdf_all(b, to_light, it->wr);
    return vscale(vdivs(f, pdf), power_heuristic(pdf, bpdf));
  }
  if (l->type == ORC_LIGHT_ENV) {
    float first = next1(smp);
    float second = next1(smp);
    ls = env_sample_uv(l, it->pos, second, first, NULL);
  } else {
    v2 u;
    u.x = next1(smp);
    u.y = next1(smp);
    ls = dir_sample(l, it->pos, u);
  }
  if (is_black(ls.radiance) || ls.pdf <= 0.f) return zero;
  const float len = vlength(vsub(ls.pos, it->pos)) - ORC_EPS;
  if (len < ORC_EPS) return zero;
  const v3 to_light = vnormalize(vsub(ls.pos, it->pos));
  orc_ray sr = make_ray(it->pos, to_light, ORC_EPS, 0.99f * len);
  if (scene_occlude(s, &sr, q)) return zero;
  v3 bf = bsdf_eval_all(b, to_light, it->wr);
  if (is_black(bf)) return zero;
  v3 normal = vnormalize(it->ns);
  v3 f = vscale(vmul(ls.radiance, bf), fabsf(vdot(to_light, normal)));
  float bpdf = bsdf_pdf_all(b, to_light, it->wr);
  return vscale(vdivs(f, ls.pdf), power_heuristic(ls.pdf, bpdf));
}

/* MISSampleBsdf (helper.h:209-265) */
static v3 mis_sample_bsdf(const orc_scene *s, const inct *it, const bsdf *b,
                          sampler *smp, uint64_t *q) {
  v3 zero = V3(0, 0, 0);
  v3 sam;
  sam.x = next1(smp);
  sam.y = next1(smp);
  sam.z = next1(smp);
  bsdf_sample bs = bsdf_sample_all(b, it->wr, sam);
  if (!bs.valid) bs.f = zero, bs.pdf = 0.f;
  if (is_black(bs.f) || bs.pdf == 0.f) return zero;
  bs.dir = vnormalize(bs.dir);
  orc_ray r = make_ray(eps_offset(it, bs.dir), bs.dir, ORC_EPS, INFINITY);
  orc_hit h;
  int hit = scene_intersect(s, &r, &h, q);
  if (!hit) {
    v3 env = zero;
    if (s->env_index >= 0) {
      const light_rt *l = &s->lights[s->env_index];
      v3 d = vload(r.d);
      v3 le = env_radiance(l, d);
      if (is_black(le)) return zero;
      v3 f = vscale(vmul(le, bs.f), fabsf(vdot(it->ns, d)));
      float lpdf = env_pdf(l, d);
      env = vadd(env, vscale(vdivs(f, bs.pdf), power_heuristic(bs.pdf, lpdf)));
    }
    return env;
  }
  const int li = s->prim_light[s->tri_prim[h.id]];
  if (li < 0) return zero;
  const light_rt *l = &s->lights[li];
  inct hi = post_intersect(s, &r, &h);
  if (!(vdot(hi.ng, hi.wr) > 0)) return zero;  /* AreaLight::Radiance */
  if (is_black(l->radiance)) return zero;
  const v3 d = vload(r.d);
  v3 f = vscale(vmul(l->radiance, bs.f), fabsf(vdot(it->ns, d)));
  const float lpdf = area_light_pdf(l, vload(r.o), hi.pos, hi.ng);
  return vscale(vdivs(f, bs.pdf), power_heuristic(bs.pdf, lpdf));
}

typedef struct path_stats {
  uint64_t queries;
} path_stats;

static v3 trace_mis(const orc_scene *s, const orc_pt_args *a, orc_ray ray,
                    sampler *smp, uint64_t *q) {
  v3 L = V3(0, 0, 0), beta = V3(1.f, 1.f, 1.f);
  for (int bounce = 0; bounce < a->max_bounces; bounce++) {
    if (bounce > a->min_bounces) {
      if (next1(smp) > a->cont_prob) return L;
      beta = vdivs(beta, a->cont_prob);
    }
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, q)) {
      if (bounce == 0 && s->env_index >= 0)
        L = vadd(L, env_radiance(&s->lights[s->env_index], vload(ray.d)));
      return L;
    }
    inct it = post_intersect(s, &ray, &h);
    if (bounce == 0) {  /* emitted radiance at the first hit (pt.cpp:188-194) */
      const int li = s->prim_light[s->tri_prim[h.id]];
      if (li >= 0 && vdot(it.ng, it.wr) > 0)
        L = vadd(L, vmul(beta, s->lights[li].radiance));
    }
    uint32_t mi = s->prims[s->tri_prim[h.id]].material;
    bsdf b = make_bsdf(&s->mats[mi], it.ng, it.ns);
    v3 direct = V3(0, 0, 0);
    for (uint32_t li = 0; li < s->n_lights; ++li)
      direct = vadd(direct, vmul(beta, mis_sample_light(s, &s->lights[li], &it, &b, smp, q)));
    direct = vadd(direct, vmul(beta, mis_sample_bsdf(s, &it, &b, smp, q)));
    L = vadd(L, vscale(direct, 1.f));
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    bsdf_sample bs = bsdf_sample_all(&b, it.wr, sam);
    if (!bs.valid) bs.f = V3(0, 0, 0), bs.pdf = 0.f;
    if (is_black(bs.f) || bs.pdf < 0.f) return L;
    beta = vmul(beta, vdivs(vscale(bs.f, fabsf(vdot(it.ng, bs.dir))), bs.pdf));
    ray = make_ray(eps_offset(&it, bs.dir), bs.dir, ORC_EPS, INFINITY);
  }
  return L;
}

/* PathTracing::Trace (pt.cpp:229-311), selected when use_mis is false
 * (pt.cpp:85-87): BSDF sampling only. An escaped ray adds the environment's
 * radiance unweighted by the path throughput, at every bounce (pt.cpp:
 * 245-249, as the reference writes it); every hit on an emitter's front adds
 * throughput x emission (pt.cpp:289-293); the direct-light loop is empty and
 * draws nothing (pt.cpp:296-301). */
static v3 trace_plain(const orc_scene *s, const orc_pt_args *a, orc_ray ray,
                      sampler *smp, uint64_t *q) {
  v3 L = V3(0, 0, 0), beta = V3(1.f, 1.f, 1.f);
  for (int bounce = 0; bounce < a->max_bounces; bounce++) {
    if (bounce > a->min_bounces) {
      if (next1(smp) > a->cont_prob) return L;
      beta = vdivs(beta, a->cont_prob);
    }
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, q)) {
      if (s->env_index >= 0) L = vadd(L, env_radiance(&s->lights[s->env_index], vload(ray.d)));
      return L;
    }
    inct it = post_intersect(s, &ray, &h);
    const int li = s->prim_light[s->tri_prim[h.id]];
    if (li >= 0 && vdot(it.ng, it.wr) > 0) L = vadd(L, vmul(beta, s->lights[li].radiance));
    uint32_t mi = s->prims[s->tri_prim[h.id]].material;
    bsdf b = make_bsdf(&s->mats[mi], it.ng, it.ns);
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    bsdf_sample bs = bsdf_sample_all(&b, it.wr, sam);
    if (!bs.valid) bs.f = V3(0, 0, 0), bs.pdf = 0.f;
    if (is_black(bs.f) || bs.pdf < 0.f) return L;
    beta = vmul(beta, vdivs(vscale(bs.f, fabsf(vdot(it.ng, bs.dir))), bs.pdf));
    ray = make_ray(eps_offset(&it, bs.dir), bs.dir, ORC_EPS, INFINITY);
  }
  return L;
}

static void render_pixel(const orc_scene *s, const orc_camera *cam,
                         const orc_pt_args *a, uint32_t x, uint32_t y,
                         uint32_t W, uint32_t Wp, double *sum, double *sum_sq,
                         uint32_t *count, float *per_sample, uint64_t *q) {
  sampler smp;
  uint32_t p = x + y * Wp;
  if (a->seed_mode == 0) orc_pcg32_seed(&smp.rng, (uint64_t)(int64_t)(int)(x + y * W));
  double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
  for (int si = 0; si < a->spp; ++si) {
    if (a->seed_mode != 0) orc_pcg32_seed(&smp.rng, sample_seed(a->seed_base, p, (uint32_t)si));
    const float u = (float)x + next1(&smp);
    const float v = (float)y + next1(&smp);
    float lens[2];
    lens[0] = next1(&smp);
    lens[1] = next1(&smp);
    orc_ray r;
    orc_camera_ray(cam, u, v, lens, &r);
    v3 L = a->use_mis ? trace_mis(s, a, r, &smp, q) : trace_plain(s, a, r, &smp, q);
    if (per_sample) vstore(per_sample + 3 * ((size_t)p * a->spp + si), L);
    acc[0] += L.x;
    acc[1] += L.y;
    acc[2] += L.z;
    acc2[0] += (double)L.x * L.x;
    acc2[1] += (double)L.y * L.y;
    acc2[2] += (double)L.z * L.z;
  }
  for (int k = 0; k < 3; ++k) {
    sum[3 * (size_t)p + k] = acc[k];
    sum_sq[3 * (size_t)p + k] = acc2[k];
  }
  count[p] = (uint32_t)a->spp;
}

static void traced_dims(const orc_camera *cam, const orc_pt_args *a, uint32_t *W,
                        uint32_t *H, uint32_t *Wp, uint32_t *Hp) {
  *W = (uint32_t)(int)cam->resolution[0];
  *H = (uint32_t)(int)cam->resolution[1];
  uint32_t ts = (uint32_t)a->tile_size;
  *Wp = (*W + ts - 1) / ts * ts;
  *Hp = (*H + ts - 1) / ts * ts;
}

uint64_t orc_render(const orc_scene *s, const orc_camera *cam,
                    const orc_pt_args *a, uint32_t pix_begin, uint32_t pix_end,
                    double *sum, double *sum_sq, uint32_t *count,
                    float *per_sample) {
  uint32_t W, H, Wp, Hp;
  traced_dims(cam, a, &W, &H, &Wp, &Hp);
  uint64_t q = 0;
  uint32_t n = Wp * Hp;
  if (pix_end > n) pix_end = n;
  for (uint32_t p = pix_begin; p < pix_end; ++p)
    render_pixel(s, cam, a, p % Wp, p / Wp, W, Wp, sum, sum_sq, count, per_sample, &q);
  return q;
}

typedef struct mt_job {
  const orc_scene *s;
  const orc_camera *cam;
  const orc_pt_args *a;
  uint32_t begin, end, W, Wp;
  double *sum, *sum_sq;
  uint32_t *count;
  uint32_t next; /* guarded by the atomic builtins */
  uint64_t queries;
} mt_job;

static void *mt_worker(void *arg) {
  mt_job *j = (mt_job *)arg;
  uint64_t q = 0;
  for (;;) {
    uint32_t p0 = __atomic_fetch_add(&j->next, 64u, __ATOMIC_RELAXED);
    if (p0 >= j->end) break;
    uint32_t p1 = p0 + 64 < j->end ? p0 + 64 : j->end;
    for (uint32_t p = p0; p < p1; ++p)
      render_pixel(j->s, j->cam, j->a, p % j->Wp, p / j->Wp, j->W, j->Wp, j->sum,
                   j->sum_sq, j->count, NULL, &q);
  }
  __atomic_fetch_add(&j->queries, q, __ATOMIC_RELAXED);
  return NULL;
}

uint64_t orc_render_mt(const orc_scene *s, const orc_camera *cam,
                       const orc_pt_args *a, uint32_t pix_begin,
                       uint32_t pix_end, double *sum, double *sum_sq,
                       uint32_t *count, int n_threads) {
  uint32_t W, H, Wp, Hp;
  traced_dims(cam, a, &W, &H, &Wp, &Hp);
  uint32_t n = Wp * Hp;
  if (pix_end > n) pix_end = n;
  mt_job j = {s, cam, a, pix_begin, pix_end, W, Wp, sum, sum_sq, count, pix_begin, 0};
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, mt_worker, &j);
  for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
  free(th);
  return j.queries;
}

/* Film resolve: a sample at traced pixel p splats weight 1 into the pixels
 * {p-1, p} x {p-1, p} (box radius 0.5, film.cpp:27-48); film writes clamp to
 * the edge (image.h:38-48); I = radiance / weight (film.h:131-152). */
/* Filtered film (film.cpp:27-63 Tile::AddSample + Film::MergeTile, image.h:38-48
 * clamped access, film.h:130-147 WriteImage without splats): every traced
 * pixel p (whole tiles) adds filter-weighted sums into j in
 * [ceil(p - 0.5 - r), floor(p - 0.5 + r)] per axis, clamped to the film;
 * box_filter.cpp:28-42 (weight 1) or gaussian_filter.cpp:28-50
 * (max(0, exp(-alpha d^2) - exp(-alpha r^2)) per axis, product). The sums of
 * a pixel's samples are weighted at once (the reference weights sample by
 * sample; same value up to float rounding). */
static float filter_axis(int type, float d, float r, float alpha) {
  if (type == 0) return 1.f;
  const float ev = expf(-alpha * r * r);
  return std_max(0.f, (float)expf(-alpha * d * d) - ev);
}
void orc_film_resolve_filter(uint32_t w, uint32_t h, uint32_t ts, int type, float rx, float ry,
                             float alpha, const double *sum, const uint32_t *count, float *rgb) {
  uint32_t Wp = (w + ts - 1) / ts * ts, Hp = (h + ts - 1) / ts * ts;
  float *rad = (float *)calloc((size_t)w * h * 3, sizeof(float));
  float *wt = (float *)calloc((size_t)w * h, sizeof(float));
  for (uint32_t py = 0; py < Hp; ++py)
    for (uint32_t px = 0; px < Wp; ++px) {
      const size_t p = (size_t)px + (size_t)py * Wp;
      const int x0 = (int)ceilf((float)px - 0.5f - rx), x1 = (int)floorf((float)px - 0.5f + rx);
      const int y0 = (int)ceilf((float)py - 0.5f - ry), y1 = (int)floorf((float)py - 0.5f + ry);
      for (int i = y0; i <= y1; ++i)
        for (int j = x0; j <= x1; ++j) {
          const float fw = filter_axis(type, (float)j - (float)px + 0.5f, rx, alpha) *
                           filter_axis(type, (float)i - (float)py + 0.5f, ry, alpha);
          const size_t q = (size_t)clampi(j, 0, (int)w - 1) + (size_t)clampi(i, 0, (int)h - 1) * w;
          wt[q] += fw * (float)count[p];
          for (int k = 0; k < 3; ++k) rad[3 * q + k] += fw * (float)sum[3 * p + k];
        }
    }
  for (size_t q = 0; q < (size_t)w * h; ++q)
    for (int k = 0; k < 3; ++k) rgb[3 * q + k] = wt[q] != 0 ? rad[3 * q + k] / wt[q] : 0.f;
  free(rad);
  free(wt);
}

/* ImageWriter::Write pixel path (image.cpp:32-62, image.h:73-84,
 * color.h:54-64): linear RGB -> sRGB (12.92 x below 0.0031308, else
 * 1.055 x^(1/2.4) - 0.055, per component), clamp to [0, 1],
 * round(v * 255.5) in double clamped to [0, 255]; flip = rows bottom to top
 * (stbi_flip_vertically_on_write). NaN components become 0. */
void orc_image_srgb8(uint32_t w, uint32_t h, const float *rgb, int flip, uint8_t *out) {
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const uint32_t sy = flip ? h - 1 - y : y;
      for (int k = 0; k < 3; ++k) {
        const float v = rgb[3 * ((size_t)x + (size_t)sy * w) + k];
        float s = v > 0.0031308f ? 1.055f * powf(v, 1.0f / 2.4f) - 0.055f : v * 12.92f;
        s = s < 0.f ? 0.f : (s > 1.f ? 1.f : s);
        if (s != s) s = 0.f;
        double r = round((double)s * 255.5);
        out[3 * ((size_t)x + (size_t)y * w) + k] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
      }
    }
}

void orc_film_resolve(uint32_t w, uint32_t h, uint32_t ts, const double *sum,
                      const uint32_t *count, float *rgb) {
  uint32_t Wp = (w + ts - 1) / ts * ts, Hp = (h + ts - 1) / ts * ts;
  double *rad = (double *)calloc((size_t)w * h * 3, sizeof(double));
  double *wt = (double *)calloc((size_t)w * h, sizeof(double));
  for (uint32_t py = 0; py < Hp; ++py)
    for (uint32_t px = 0; px < Wp; ++px) {
      size_t p = (size_t)px + (size_t)py * Wp;
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          int qx = clampi((int)px + dx, 0, (int)w - 1);
          int qy = clampi((int)py + dy, 0, (int)h - 1);
          size_t q = (size_t)qx + (size_t)qy * w;
          for (int k = 0; k < 3; ++k) rad[3 * q + k] += sum[3 * p + k];
          wt[q] += count[p];
        }
    }
  for (size_t q = 0; q < (size_t)w * h; ++q)
    for (int k = 0; k < 3; ++k)
      rgb[3 * q + k] = wt[q] != 0 ? (float)(rad[3 * q + k] / wt[q]) : 0.f;
  free(rad);
  free(wt);
}

/* ---- SPPM (SPPMRenderer, sppm.cpp:35-303) -------------------------------
 * Restated as written: every iteration reseeds each pixel's sampler with
 * x + y*W (sppm.cpp:222), so the visible points and their direct light are
 * the same in every iteration and the direct light is accumulated once per
 * iteration; visible points register in the hash-grid cells [min, max) of
 * their radius box (the upper cell excluded, photon_mapping.h:88-90); the
 * entry index keeps the reference's low bits (z & 2 on bit 1,
 * photon_mapping.h:133-138); the next iteration's grid cell is 1.05 x the
 * largest radius among valid visible points before the update
 * (sppm.cpp:263-271). Photon i of iteration k draws from its own PCG32
 * stream seeded with mix(photon_seed_base, i, k): the reference passes the
 * renderer's one shared sampler to every photon thread (sppm.cpp:255-260),
 * whose draws are not reproducible. */
typedef struct sppm_px {
  int valid;
  v3 pos, wr, beta;
  v3 beta_traced; /* the throughput as traced (UpdatePixel zeroes beta) */
  bsdf b;
  v3 direct, tau;
  float radius, n;
  float phi[3];
  int32_t m;
} sppm_px;

/* TracerVP (sppm.cpp:35-106) */
static void sppm_tracer_vp(const orc_scene *s, int max_depth, orc_ray ray, sampler *smp,
                           sppm_px *px, uint64_t *q) {
  v3 beta = V3(1.f, 1.f, 1.f);
  px->valid = 0;
  for (int depth = 0; depth < max_depth; ++depth) {
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, q)) {
      if (depth == 0 && s->env_index >= 0)
        px->direct = vadd(px->direct,
                          vmul(beta, env_radiance(&s->lights[s->env_index], vload(ray.d))));
      return;
    }
    inct it = post_intersect(s, &ray, &h);
    uint32_t mi = s->prims[s->tri_prim[h.id]].material;
    bsdf b = make_bsdf(&s->mats[mi], it.ng, it.ns);
    if (depth == 0) {
      const int li = s->prim_light[s->tri_prim[h.id]];
      if (li >= 0 && vdot(it.ng, it.wr) > 0)
        px->direct = vadd(px->direct, vmul(beta, s->lights[li].radiance));
    }
    v3 sum_di = V3(0, 0, 0);
    for (uint32_t li = 0; li < s->n_lights; ++li)
      sum_di = vadd(sum_di, vmul(beta, mis_sample_light(s, &s->lights[li], &it, &b, smp, q)));
    sum_di = vadd(sum_di, vmul(beta, mis_sample_bsdf(s, &it, &b, smp, q)));
    px->direct = vadd(px->direct, vscale(sum_di, 1.f / 1));
    /* HasDiffuseComponent: Lambert always, Disney when metallic < 1
     * (disney.cpp:79) */
    const int diffuse = b.m->type == ORC_MAT_DIFFUSE || b.m->metallic < 1.f;
    if (diffuse || depth == max_depth - 1) {
      if (!(isfinite(beta.x) && isfinite(beta.y) && isfinite(beta.z))) return;
      px->valid = 1;
      px->pos = it.pos;
      px->beta = beta;
      px->beta_traced = beta;
      px->b = b;
      px->wr = it.wr;
      return;
    }
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    bsdf_sample bs = bsdf_sample_all(&b, it.wr, sam);
    if (!bs.valid) bs.f = V3(0, 0, 0), bs.pdf = 0.f;
    if (is_black(bs.f) || bs.pdf < 0.f) return;
    beta = vmul(beta, vdivs(vscale(bs.f, fabsf(vdot(it.ng, bs.dir))), bs.pdf));
    ray = make_ray(eps_offset(&it, bs.dir), bs.dir, ORC_EPS, INFINITY);
  }
}

/* VisiblePointSearcher (photon_mapping.h:70-151): 40960 entries. */
#define SPPM_ENTRIES 40960u
typedef struct sppm_grid {
  box3 bound;
  float side;
  uint32_t *start; /* SPPM_ENTRIES + 1 */
  uint32_t *list;
} sppm_grid;

static void sppm_to_grid(const sppm_grid *g, v3 p, int *c) {
  const v3 hd = vsub(p, g->bound.mn);
  c[0] = (int)(std_max(hd.x, 0.f) / g->side);
  c[1] = (int)(std_max(hd.y, 0.f) / g->side);
  c[2] = (int)(std_max(hd.z, 0.f) / g->side);
}
static uint64_t hash_combine(uint64_t a, uint64_t b) { /* utility/hash.h:31-36 */
  return a ^ (b + 0x9e3779b97f4a7c15ull + (a << 6) + (a >> 2));
}
static uint32_t sppm_entry(int x, int y, int z) {
  /* std::hash<int> is the value itself (sign-extended to size_t) */
  const uint64_t low3 = (uint64_t)((x & 1) | ((y & 1) << 1) | (z & 2));
  const uint64_t h = hash_combine(hash_combine((uint64_t)(int64_t)x, (uint64_t)(int64_t)y),
                                  (uint64_t)(int64_t)z);
  return (uint32_t)(((h << 3) % SPPM_ENTRIES) | low3);
}

static void sppm_vp_cells(const sppm_grid *g, const sppm_px *p, int *mn, int *mx) {
  const v3 r = V3(p->radius, p->radius, p->radius);
  sppm_to_grid(g, vsub(p->pos, r), mn);
  sppm_to_grid(g, vadd(p->pos, r), mx);
}

static void sppm_grid_build(sppm_grid *g, const sppm_px *px, uint32_t n) {
  memset(g->start, 0, sizeof(uint32_t) * (SPPM_ENTRIES + 1));
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t i = 0; i < n; ++i) {
      if (!px[i].valid) continue;
      int mn[3], mx[3];
      sppm_vp_cells(g, &px[i], mn, mx);
      for (int z = mn[2]; z < mx[2]; ++z)
        for (int y = mn[1]; y < mx[1]; ++y)
          for (int x = mn[0]; x < mx[0]; ++x) {
            const uint32_t e = sppm_entry(x, y, z);
            if (pass == 0)
              g->start[e + 1]++;
            else
              g->list[g->start[e]++] = i;
          }
    }
    if (pass == 0) {
      uint32_t total = 0;
      for (uint32_t e = 1; e <= SPPM_ENTRIES; ++e) g->start[e] += g->start[e - 1];
      total = g->start[SPPM_ENTRIES];
      g->list = (uint32_t *)realloc(g->list, sizeof(uint32_t) * (total ? total : 1));
    }
  }
  /* the fill pass advanced start[e] to the end of entry e: shift back */
  for (uint32_t e = SPPM_ENTRIES; e > 0; --e) g->start[e] = g->start[e - 1];
  g->start[0] = 0;
}

/* VisiblePointSearcher::AddPhoton (photon_mapping.h:103-121) */
static void atomic_addf(float *p, float v) {
  float old, nw;
  __atomic_load(p, &old, __ATOMIC_RELAXED);
  do {
    nw = old + v;
  } while (!__atomic_compare_exchange(p, &old, &nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

/* shared: other threads deposit into the same pixels (atomic adds; the sums
 * then depend on the order photons arrive in) */
static int64_t sppm_add_photon(const sppm_grid *g, sppm_px *px, v3 pos, v3 phi, v3 wr,
                               int shared) {
  int c[3];
  sppm_to_grid(g, pos, c);
  const uint32_t e = sppm_entry(c[0], c[1], c[2]);
  int64_t n = 0;
  for (uint32_t k = g->start[e]; k < g->start[e + 1]; ++k) {
    sppm_px *p = &px[g->list[k]];
    const v3 d = vsub(p->pos, pos);
    if (vdot(d, d) > p->radius * p->radius) continue;
    const v3 dphi = vmul(phi, bsdf_eval_all(&p->b, wr, p->wr));
    if (!(isfinite(dphi.x) && isfinite(dphi.y) && isfinite(dphi.z))) continue;
    if (shared) {
      atomic_addf(&p->phi[0], dphi.x);
      atomic_addf(&p->phi[1], dphi.y);
      atomic_addf(&p->phi[2], dphi.z);
      __atomic_fetch_add(&p->m, 1, __ATOMIC_RELAXED);
    } else {
      p->phi[0] += dphi.x;
      p->phi[1] += dphi.y;
      p->phi[2] += dphi.z;
      p->m++;
    }
    ++n;
  }
  return n;
}

typedef struct light_emit {
  v3 pos, dir, normal, radiance;
  float pdf_pos, pdf_dir;
} light_emit;

/* Light::SampleEmit: AreaLight (area_light.cpp:60-73), DirectionalLight
 * (directional_light.cpp:60-79), EnvLight (env_light.cpp:105-147). */
static light_emit light_sample_emit(const light_rt *l, sampler *smp) {
  light_emit r;
  memset(&r, 0, sizeof(r));
  if (l->type == ORC_LIGHT_AREA) {
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    v3 pos, ng;
    float pdf_pos;
    area_geom_sample(l, sam, &pos, &ng, &pdf_pos);
    v2 s2;
    s2.x = next1(smp);
    s2.y = next1(smp);
    float pdf_dir;
    const v3 lwi = cosine_hemisphere(s2, &pdf_dir);
    const frame f = make_frame(ng);
    r.pos = pos;
    r.dir = to_world(&f, lwi);
    r.normal = ng;
    r.radiance = l->radiance;
    r.pdf_pos = pdf_pos;
    r.pdf_dir = pdf_dir;
  } else if (l->type == ORC_LIGHT_DIRECTIONAL) {
    v2 c;
    c.x = next1(smp);
    c.y = next1(smp);
    const v3 ref2light = to_world(&l->fr, uniform_cone(c, l->angle_cos));
    const float pdf_dir = uniform_cone_pdf(l->angle_cos);
    const v3 center = box_center(l->aabb);
    const float radius = vlength(vsub(l->aabb.mx, center));
    v2 d2;
    d2.x = next1(smp);
    d2.y = next1(smp);
    const v2 disk = disk_concentric(d2);
    const frame dc = make_frame(ref2light);
    r.pos = vadd(center, vscale(vadd(vadd(vscale(dc.dx, disk.x), vscale(dc.dy, disk.y)), ref2light),
                                radius));
    r.dir = vneg(ref2light);
    r.normal = vneg(vnormalize(ref2light));
    r.radiance = l->radiance;
    r.pdf_pos = 1.f / (ORC_PI * radius * radius);
    r.pdf_dir = pdf_dir;
  } else {
    /* SampleContinuous(Next1D(), Next1D()): arguments drawn right to left */
    const float second = next1(smp);
    const float first = next1(smp);
    float pdfs[2];
    int iv;
    const float v = distrib_sample_cont(&l->marg, second, &pdfs[1], &iv);
    float u = distrib_sample_cont(&l->cond[iv], first, &pdfs[0], NULL);
    float pdf = pdfs[0] * pdfs[1];
    if (pdf == 0.f) return r;
    u -= l->rot / 180 * ORC_PI * kInv2Pi;
    if (u > 1.f) u -= 1.f;
    if (u < 0.f) u += 1.f;
    const float phi = u * kTwoPi;
    const float theta = v * ORC_PI;
    const float sin_t = sinf(theta);
    pdf = sin_t != 0.f ? pdf / (2.f * ORC_PI * ORC_PI * sin_t) : 0.f;
    v3 dir = V3(-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi));
    dir = vneg(dir);
    const v3 center = box_center(l->aabb);
    const float radius = vlength(vsub(l->aabb.mx, center));
    v2 d2;
    d2.x = next1(smp);
    d2.y = next1(smp);
    const v2 disk = disk_concentric(d2);
    const frame dc = make_frame(vneg(dir));
    r.pos = vadd(center, vscale(vsub(vadd(vscale(dc.dx, disk.x), vscale(dc.dy, disk.y)), dir),
                                radius));
    r.dir = dir;
    r.normal = vnormalize(dir);
    r.radiance = tex_sample(l, u, v);
    r.pdf_pos = 1.f / (ORC_PI * radius * radius);
    r.pdf_dir = pdf;
  }
  return r;
}

/* Test hook: n draws of Light::SampleEmit from light li, draw i from PCG32
 * seeded with sample_seed(seed, i, 0). */
void orc_light_emit_batch(const orc_scene *s, uint32_t li, uint32_t seed, uint32_t n,
                          float *pos, float *dir, float *normal, float *radiance,
                          float *pdf_pos, float *pdf_dir) {
  for (uint32_t i = 0; i < n; ++i) {
    sampler smp;
    orc_pcg32_seed(&smp.rng, sample_seed(seed, i, 0));
    const light_emit em = light_sample_emit(&s->lights[li], &smp);
    vstore(pos + 3 * (size_t)i, em.pos);
    vstore(dir + 3 * (size_t)i, em.dir);
    vstore(normal + 3 * (size_t)i, em.normal);
    vstore(radiance + 3 * (size_t)i, em.radiance);
    pdf_pos[i] = em.pdf_pos;
    pdf_dir[i] = em.pdf_dir;
  }
}

/* TracerPhoton (sppm.cpp:108-143); the light is picked by DefaultScene::
 * SampleLight (default_scene.cpp:76-85) from a Distribution over equal
 * weights (FP32 prefix sums, distribution.h:40-58). */
static int64_t sppm_tracer_photon(const orc_scene *s, const orc_sppm_args *a,
                                  const distrib *sel, float sel_pdf, const sppm_grid *g,
                                  sppm_px *px, sampler *smp, uint64_t *q, int shared) {
  if (s->n_lights == 0) return 0;
  int idx;
  (void)distrib_sample_cont(sel, next1(smp), NULL, &idx);
  const light_rt *l = &s->lights[idx];
  const light_emit em = light_sample_emit(l, smp);
  if (is_black(em.radiance) || em.pdf_dir == 0.f || em.pdf_pos == 0.f) return 0;
  v3 coef = vdivs(vscale(em.radiance, fabsf(vdot(em.normal, em.dir))),
                  sel_pdf * em.pdf_pos * em.pdf_dir);
  orc_ray ray = make_ray(em.pos, em.dir, ORC_EPS, INFINITY);
  int64_t n = 0;
  for (int depth = 1; depth <= a->photon_max_depth; ++depth) {
    if (depth > a->photon_min_depth) {
      if (next1(smp) > a->photon_cont_prob) return n;
      coef = vdivs(coef, a->photon_cont_prob);
    }
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, q)) return n;
    inct it = post_intersect(s, &ray, &h);
    if (depth > 1) n += sppm_add_photon(g, px, it.pos, coef, it.wr, shared);
    uint32_t mi = s->prims[s->tri_prim[h.id]].material;
    bsdf b = make_bsdf(&s->mats[mi], it.ng, it.ns);
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    bsdf_sample bs = bsdf_sample_all(&b, it.wr, sam);
    if (!bs.valid) bs.f = V3(0, 0, 0), bs.pdf = 0.f;
    if (is_black(bs.f) || bs.pdf < 0.f) return n;
    coef = vmul(coef, vscale(vdivs(bs.f, bs.pdf), fabsf(vdot(it.ng, bs.dir))));
    ray = make_ray(eps_offset(&it, bs.dir), bs.dir, ORC_EPS, INFINITY);
  }
  return n;
}

/* Test hook: the grid of visible points (pos, radius) with cell side `side`
 * over a bound starting at bmin, and, per visible point, how many of the
 * photon positions AddPhoton's lookup and distance test match to it. */
void orc_sppm_grid_probe(const float *bmin, float side, uint32_t n_vp, const float *vp_pos,
                         const float *vp_r, uint32_t n_ph, const float *ph_pos, uint32_t *counts) {
  sppm_px *px = (sppm_px *)calloc(n_vp ? n_vp : 1, sizeof(sppm_px));
  for (uint32_t i = 0; i < n_vp; ++i) {
    px[i].valid = 1;
    px[i].pos = vload(vp_pos + 3 * (size_t)i);
    px[i].radius = vp_r[i];
    counts[i] = 0;
  }
  sppm_grid g;
  memset(&g, 0, sizeof(g));
  g.bound.mn = vload(bmin);
  g.side = side;
  g.start = (uint32_t *)malloc(sizeof(uint32_t) * (SPPM_ENTRIES + 1));
  sppm_grid_build(&g, px, n_vp);
  for (uint32_t k = 0; k < n_ph; ++k) {
    const v3 pos = vload(ph_pos + 3 * (size_t)k);
    int c[3];
    sppm_to_grid(&g, pos, c);
    const uint32_t e = sppm_entry(c[0], c[1], c[2]);
    for (uint32_t j = g.start[e]; j < g.start[e + 1]; ++j) {
      const sppm_px *p = &px[g.list[j]];
      const v3 d = vsub(p->pos, pos);
      if (vdot(d, d) > p->radius * p->radius) continue;
      counts[g.list[j]]++;
    }
  }
  free(g.start);
  free(g.list);
  free(px);
}

/* UpdatePixel (sppm.cpp:145-167) */
static void sppm_update(float alpha, sppm_px *p) {
  if (p->m > 0) {
    const float new_n = p->n + alpha * (float)p->m;
    const float new_r = p->radius * sqrtf(new_n / (p->n + (float)p->m));
    const v3 phi = V3(p->phi[0], p->phi[1], p->phi[2]);
    p->tau = vdivs(vscale(vadd(p->tau, vmul(p->beta, phi)), new_r * new_r),
                   p->radius * p->radius);
    p->n = new_n;
    p->radius = new_r;
    p->m = 0;
    p->phi[0] = p->phi[1] = p->phi[2] = 0.f;
  }
  p->beta = V3(0, 0, 0);
}

/* Test hook: one pixel through `iters` UpdatePixel calls, iteration k
 * bringing m[k] photons of flux phi[3k..3k+2] at throughput beta (3);
 * writes the radius, N and tau (3) after each call (out: iters x 5). */
void orc_sppm_update_probe(float alpha, float r0, const float *beta, uint32_t iters,
                           const uint32_t *m, const float *phi, float *out) {
  sppm_px p;
  memset(&p, 0, sizeof(p));
  p.valid = 1;
  p.radius = r0;
  for (uint32_t k = 0; k < iters; ++k) {
    p.beta = vload(beta);
    p.m = m[k];
    p.phi[0] = phi[3 * k];
    p.phi[1] = phi[3 * k + 1];
    p.phi[2] = phi[3 * k + 2];
    sppm_update(alpha, &p);
    out[5 * k] = p.radius;
    out[5 * k + 1] = p.n;
    out[5 * k + 2] = p.tau.x;
    out[5 * k + 3] = p.tau.y;
    out[5 * k + 4] = p.tau.z;
  }
}

/* One SPPM phase over items [0, total) split across threads: the visible
 * points of every pixel, or one iteration's photons. */
typedef struct sppm_job {
  const orc_scene *s;
  const orc_camera *cam;
  const orc_sppm_args *a;
  const distrib *sel;
  float sel_pdf;
  const sppm_grid *g;
  sppm_px *px;
  uint32_t W, it, total, phase, shared;
  uint32_t next;    /* guarded by the atomic builtins */
  int64_t deposits; /* ditto */
  uint64_t q;       /* ditto */
} sppm_job;

static void sppm_vp_pixel(const sppm_job *j, uint32_t i, uint64_t *q) {
  const uint32_t x = i % j->W, y = i / j->W;
  sampler smp;
  orc_pcg32_seed(&smp.rng, (uint64_t)(int64_t)(int)(x + y * j->W));
  const float u = (float)x + next1(&smp);
  const float v = (float)y + next1(&smp);
  float lens[2];
  lens[0] = next1(&smp);
  lens[1] = next1(&smp);
  orc_ray r;
  orc_camera_ray(j->cam, u, v, lens, &r);
  sppm_tracer_vp(j->s, j->a->forward_max_depth, r, &smp, &j->px[i], q);
}

static void *sppm_worker(void *arg) {
  sppm_job *j = (sppm_job *)arg;
  uint64_t q = 0;
  int64_t dep = 0;
  for (;;) {
    const uint32_t i0 = __atomic_fetch_add(&j->next, 256u, __ATOMIC_RELAXED);
    if (i0 >= j->total) break;
    const uint32_t i1 = i0 + 256 < j->total ? i0 + 256 : j->total;
    for (uint32_t i = i0; i < i1; ++i) {
      if (j->phase == 0) {
        sppm_vp_pixel(j, i, &q);
      } else {
        sampler smp;
        orc_pcg32_seed(&smp.rng, sample_seed(j->a->photon_seed_base, i, j->it));
        dep += sppm_tracer_photon(j->s, j->a, j->sel, j->sel_pdf, j->g, j->px, &smp, &q,
                                  (int)j->shared);
      }
    }
  }
  __atomic_fetch_add(&j->q, q, __ATOMIC_RELAXED);
  __atomic_fetch_add(&j->deposits, dep, __ATOMIC_RELAXED);
  return NULL;
}

static int64_t sppm_phase(sppm_job *j, int n_threads) {
  j->next = 0;
  j->deposits = 0;
  j->shared = n_threads > 1;
  if (n_threads <= 1) {
    sppm_worker(j);
    return j->deposits;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, sppm_worker, j);
  for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
  free(th);
  return j->deposits;
}

int64_t orc_render_sppm(const orc_scene *s, const orc_camera *cam, const orc_sppm_args *a,
                        int32_t iterations, float *radiance, orc_sppm_pixel *state) {
  return orc_render_sppm_mt(s, cam, a, iterations, radiance, state, 1);
}

int64_t orc_render_sppm_mt(const orc_scene *s, const orc_camera *cam, const orc_sppm_args *a,
                           int32_t iterations, float *radiance, orc_sppm_pixel *state,
                           int n_threads) {
  const uint32_t W = (uint32_t)(int)cam->resolution[0], H = (uint32_t)(int)cam->resolution[1];
  if (W % 64 || H % 64) return -1;
  const uint32_t n = W * H;
  if (iterations <= 0 || iterations > a->iteration_count) iterations = a->iteration_count;
  sppm_px *px = (sppm_px *)calloc(n, sizeof(sppm_px));
  float init_radius = a->init_radius;
  if (init_radius < 0) init_radius = vlength(vsub(s->world.mx, s->world.mn)) / 1000.f;
  sppm_grid g;
  memset(&g, 0, sizeof(g));
  g.bound.mn = vsub(s->world.mn, V3(init_radius, init_radius, init_radius));
  g.bound.mx = vadd(s->world.mx, V3(init_radius, init_radius, init_radius));
  g.start = (uint32_t *)malloc(sizeof(uint32_t) * (SPPM_ENTRIES + 1));
  for (uint32_t i = 0; i < n; ++i) px[i].radius = init_radius;
  /* light selector: pdf_table of ones normalised by its sum (default_scene.cpp:50-55) */
  distrib sel;
  float sel_pdf = 0.f;
  const uint32_t nl = s->n_lights ? s->n_lights : 1;
  {
    float *tab = (float *)malloc(sizeof(float) * nl);
    float sum = 0.f;
    for (uint32_t i = 0; i < nl; ++i) tab[i] = 1.f;
    for (uint32_t i = 0; i < nl; ++i) sum += tab[i];
    const float ratio = 1.f / sum;
    for (uint32_t i = 0; i < nl; ++i) tab[i] *= ratio;
    sel_pdf = tab[0];
    distrib_init(&sel, tab, nl, 0);
    free(tab);
  }
  float max_radius = init_radius;
  int64_t deposits = 0;
  sppm_job job;
  memset(&job, 0, sizeof(job));
  job.s = s;
  job.cam = cam;
  job.a = a;
  job.sel = &sel;
  job.sel_pdf = sel_pdf;
  job.g = &g;
  job.px = px;
  job.W = W;
  for (int32_t it = 0; it < iterations; ++it) {
    g.side = 1.05f * max_radius;
    job.phase = 0;
    job.total = n;
    (void)sppm_phase(&job, n_threads);
    sppm_grid_build(&g, px, n);
    job.phase = 1;
    job.it = (uint32_t)it;
    job.total = (uint32_t)(a->photons_per_iteration > 0 ? a->photons_per_iteration : 0);
    deposits += sppm_phase(&job, n_threads);
    max_radius = 0.f;
    for (uint32_t i = 0; i < n; ++i)
      if (px[i].valid) max_radius = std_max(max_radius, px[i].radius);
    if (max_radius == 0.f) max_radius = init_radius;
    for (uint32_t i = 0; i < n; ++i)
      if (px[i].valid) sppm_update(a->update_alpha, &px[i]);
  }
  /* ComputePixelRadiance (sppm.cpp:169-178) with the full photon count */
  const uint64_t photon_n = (uint64_t)a->iteration_count * (uint64_t)a->photons_per_iteration;
  for (uint32_t i = 0; i < n; ++i) {
    const v3 direct = vdivs(px[i].direct, (float)a->iteration_count);
    const float dem = (float)photon_n * ORC_PI * px[i].radius * px[i].radius;
    const v3 L = vadd(direct, vdivs(px[i].tau, dem));
    if (radiance) vstore(radiance + 3 * (size_t)i, L);
    if (state) {
      orc_sppm_pixel *o = &state[i];
      vstore(o->pos, px[i].valid ? px[i].pos : V3(0, 0, 0));
      vstore(o->beta, px[i].valid ? px[i].beta_traced : V3(0, 0, 0));
      vstore(o->direct, px[i].direct);
      vstore(o->tau, px[i].tau);
      o->radius = px[i].radius;
      o->n = px[i].n;
      o->valid = (uint32_t)px[i].valid;
      o->pad = 0;
    }
  }
  distrib_free(&sel);
  free(g.start);
  free(g.list);
  free(px);
  return deposits;
}

/* ---- VCM (VCMRenderer, vcm.cpp:130-751) ----------------------------------
 * Restated as written, including: the reverse pdf of a scattering event is
 * PdfAll(-new_dir, new_dir) (the state's direction is overwritten first,
 * vcm.cpp:416-423); the camera path's first ray is offset twice (the
 * primary ray's origin already carries the Ray constructor's offset and
 * `Ray{origin, direction}` adds it again); a scattered path restarts from the
 * hit point without EpsOffset; every vertex draws a BSDF sample (Next3D) that
 * only its is_delta is read from; light tracing splats are summed over
 * iterations while camera-path colors are averaged (Film::WriteImage).
 * Streams: light path i of iteration k draws from PCG32 seeded with
 * mix(seed_base, i, 2k), the camera path of pixel p from mix(seed_base, p,
 * 2k+1); the reference draws every path from one sampler in sequence. */
typedef struct vcm_state {
  v3 origin, direction, throughput;
  uint32_t path_length, is_finite_light, specular_path;
  float dVCM, dVC, dVM;
} vcm_state;

typedef struct vcm_vertex {
  v3 hit_point, in_dir, throughput;
  uint32_t path_length;
  v3 normal;
  bsdf b;
  float dVCM, dVC, dVM;
} vcm_vertex;

typedef struct vcm_ctx {
  const orc_scene *s;
  const orc_camera *cam;
  const orc_vcm_args *a;
  float w2c[16], c2r[16], w2r[16];
  v3 cam_pos, cam_dir;
  int W, H;
  float light_count, mis_vm, mis_vc, vm_norm, radius, radius_sqr;
  float pix_k; /* pixels per unit image-plane area at the focal distance */
  distrib sel;
  float sel_pdf;
  float *splat;
  int shared_splat;    /* other threads splat too: atomic adds */
  float *camera_color; /* the camera phase's color sums (threaded phase) */
  /* light vertices of the iteration (path order) */
  vcm_vertex *lv;
  uint32_t n_lv, cap_lv;
  uint32_t *path_ends;
  /* HashGrid (hashgrid.h:30-175) */
  box3 gbox;
  float cell_size, inv_cell;
  uint32_t n_cells;
  int *cell_ends;
  int *indices;
  uint64_t q;
} vcm_ctx;

static void camera_derived(const orc_camera *cam, float *w2c, float *c2r, float *w2r) {
  double c2w[16], r2c[16], r2w[16], t[16];
  for (int i = 0; i < 16; ++i) c2w[i] = cam->camera2world[i], r2c[i] = cam->raster2camera[i];
  inv4d(c2w, t);
  for (int i = 0; i < 16; ++i) w2c[i] = (float)t[i];
  inv4d(r2c, t);
  for (int i = 0; i < 16; ++i) c2r[i] = (float)t[i];
  mul4d(c2w, r2c, r2w);
  inv4d(r2w, t);
  for (int i = 0; i < 16; ++i) w2r[i] = (float)t[i];
}

/* ThinLensCamera::ToRaster (thin_lens.cpp:85-119) */
static int vcm_to_raster(const vcm_ctx *c, v2 u, v3 pos, v3 *dir_to_camera, v2 *p_raster) {
  const orc_camera *cam = c->cam;
  const v3 cam_coord = xform_point(c->w2c, pos);
  v3 raster;
  if (cam->lens_radius > 0 && cam->focal_distance > 0) {
    v2 lens = disk_concentric(u);
    lens.x = cam->lens_radius * lens.x;
    lens.y = cam->lens_radius * lens.y;
    const v3 origin = V3(lens.x, lens.y, 0.f);
    const v3 direction = vnormalize(vsub(cam_coord, origin));
    const float ft = cam->focal_distance / direction.z;
    const v3 focus = vadd(origin, vscale(direction, ft));
    raster = xform_point(c->c2r, focus);
    if (!(raster.x < cam->resolution[0] && raster.x >= 0.f && raster.y < cam->resolution[1] &&
          raster.y >= 0.f))
      return 0;
    *dir_to_camera = vsub(xform_point(cam->camera2world, origin), pos);
  } else {
    raster = xform_point(c->w2r, pos);
    if (!(raster.x < cam->resolution[0] && raster.x >= 0.f && raster.y < cam->resolution[1] &&
          raster.y >= 0.f))
      return 0;
    *dir_to_camera = vsub(c->cam_pos, pos);
  }
  p_raster->x = raster.x;
  p_raster->y = raster.y;
  if (vdot(c->cam_dir, vneg(*dir_to_camera)) <= 0.f) return 0;
  return 1;
}

/* ImageAreaToSurfaceArea / ImageAreaToSolidAngle (thin_lens.cpp:121-147) */
static float vcm_image_to_surface(const vcm_ctx *c, v3 dir_to_camera, float cos_to_camera,
                                  float distance) {
  const float cos_at_camera = vdot(c->cam_dir, vneg(vnormalize(dir_to_camera)));
  const float d = c->cam->focal_distance / cos_at_camera;
  const float image_to_solid = d * d / cos_at_camera;
  return image_to_solid * fabsf(cos_to_camera) / (distance * distance);
}
static float vcm_image_to_solid(const vcm_ctx *c, v3 direction) {
  const float cos_at_camera = vdot(c->cam_dir, vnormalize(direction));
  const float d = c->cam->focal_distance / cos_at_camera;
  return d * d / cos_at_camera;
}

static int vcm_pick_light(const vcm_ctx *c, sampler *smp) {
  int idx;
  (void)distrib_sample_cont(&c->sel, next1(smp), NULL, &idx);
  return idx;
}

/* SampleScattering (vcm.cpp:409-451) */
static int vcm_scatter(const vcm_ctx *c, vcm_state *st, sampler *smp, const bsdf *b,
                       const inct *it) {
  v3 sam;
  sam.x = next1(smp);
  sam.y = next1(smp);
  sam.z = next1(smp);
  bsdf_sample bs = bsdf_sample_all(b, vneg(st->direction), sam);
  if (!bs.valid) bs.f = V3(0, 0, 0), bs.pdf = 0.f, bs.dir = V3(0, 0, 0);
  if (is_black(bs.f)) return 0;
  st->direction = bs.dir;
  float dir_pdf = bs.pdf;
  float rev_pdf = (c->a->variant & ORC_VCM_REVERSE_PDF) ? bsdf_pdf_all(b, it->wr, bs.dir)
                                                          : bsdf_pdf_all(b, vneg(st->direction), bs.dir);
  if (next1(smp) > c->a->cont_prob) return 0;
  dir_pdf *= c->a->cont_prob;
  rev_pdf *= c->a->cont_prob;
  const float cos_out = vdot(it->ng, bs.dir);
  st->dVC = (cos_out / dir_pdf) * (st->dVC * rev_pdf + st->dVCM + c->mis_vm);
  st->dVM = (cos_out / dir_pdf) * (st->dVM * rev_pdf + st->dVCM * c->mis_vc + 1.f);
  st->dVCM = 1.f / dir_pdf;
  st->specular_path &= 0;
  st->origin = it->pos;
  st->throughput = vmul(st->throughput, vscale(bs.f, cos_out / dir_pdf));
  return 1;
}

/* ConnectToCamera (vcm.cpp:544-600) */
static void vcm_connect_camera(vcm_ctx *c, const vcm_state *st, sampler *smp, const bsdf *b,
                               const inct *it) {
  v2 u;
  u.x = next1(smp);
  u.y = next1(smp);
  v3 dir;
  v2 img;
  if (!vcm_to_raster(c, u, it->pos, &dir, &img)) return;
  const float distance = vlength(dir);
  dir = vnormalize(dir);
  const v3 f = bsdf_eval_all(b, dir, it->wr);
  if (is_black(f)) return;
  float rev_pdf = bsdf_pdf_all(b, it->wr, dir);
  rev_pdf *= c->a->cont_prob;
  const float cos_to_cam = vdot(it->ng, dir);
  float i2s = vcm_image_to_surface(c, dir, cos_to_cam, distance);
  if (c->a->variant & ORC_VCM_PIXEL_UNITS) i2s *= c->pix_k;
  const float w_light = (i2s / c->light_count) * (c->mis_vm + st->dVCM + st->dVC * rev_pdf);
  const float mis = (c->a->variant & ORC_VCM_LT_ONLY) ? 1.f : 1.f / (w_light + 1.f);
  const float s2i = 1.f / i2s;
  const v3 contrib = vdivs(vmul(vscale(st->throughput, mis), f), c->light_count * s2i);
  if (is_black(contrib)) return;
  orc_ray sr = make_ray(it->pos, dir, ORC_EPS, distance * 0.9f);
  if (scene_occlude(c->s, &sr, &c->q)) return;
  const int ix = (int)img.x, iy = (int)img.y;
  float *px = c->splat + 3 * ((size_t)iy * c->W + ix);
  if (c->shared_splat) {
    atomic_addf(&px[0], contrib.x);
    atomic_addf(&px[1], contrib.y);
    atomic_addf(&px[2], contrib.z);
  } else {
    px[0] += contrib.x;
    px[1] += contrib.y;
    px[2] += contrib.z;
  }
}

/* Light::Sample as DirectIllumination reads it: position, normal, radiance,
 * pdf, emit_pdf (area_light.cpp:37-58, env_light.cpp:68-103,
 * directional_light.cpp:43-58). */
typedef struct vcm_lsample {
  v3 ref, pos, normal, radiance;
  float pdf, emit_pdf;
} vcm_lsample;
static vcm_lsample vcm_light_sample(const vcm_ctx *c, const light_rt *l, v3 ref, sampler *smp) {
  vcm_lsample r;
  memset(&r, 0, sizeof(r));
  if (l->type == ORC_LIGHT_AREA) {
    v3 sam;
    sam.x = next1(smp);
    sam.y = next1(smp);
    sam.z = next1(smp);
    v3 pos, ng;
    float pdf_area;
    area_geom_sample(l, sam, &pos, &ng, &pdf_area);
    if (vdot(ng, vsub(ref, pos)) <= 0) return r;
    const v3 i2r = vsub(ref, pos);
    const float dist2 = vdot(i2r, i2r);
    r.ref = ref;
    r.pos = pos;
    r.normal = ng;
    r.radiance = l->radiance;
    r.pdf = pdf_area * dist2 / fabsf(vdot(ng, i2r));
    r.emit_pdf = pdf_area * vdot(ng, vnormalize(i2r)) / ORC_PI;
  } else if (l->type == ORC_LIGHT_ENV) {
    const float first = next1(smp);
    const float second = next1(smp);
    v3 dir;
    light_sample ls = env_sample_uv(l, ref, second, first, &dir);
    if (ls.pdf == 0.f && is_black(ls.radiance)) return r;
    const float radius = vlength(vsub(l->aabb.mx, box_center(l->aabb)));
    r.ref = ref;
    r.pos = ls.pos;
    r.normal = vneg(dir);
    r.radiance = ls.radiance;
    r.pdf = ls.pdf;
    r.emit_pdf = ls.pdf / (ORC_PI * radius * radius);
  } else {
    v2 u;
    u.x = next1(smp);
    u.y = next1(smp);
    light_sample ls = dir_sample(l, ref, u);
    const float radius = vlength(vsub(l->aabb.mx, box_center(l->aabb)));
    r.ref = ref;
    r.pos = ls.pos;
    r.radiance = ls.radiance;
    r.pdf = ls.pdf;
    r.emit_pdf = ls.pdf / (ORC_PI * radius * radius);
  }
  return r;
}

/* DirectIllumination (vcm.cpp:604-668) */
static v3 vcm_direct(vcm_ctx *c, const vcm_state *st, const inct *it, const bsdf *b,
                     sampler *smp) {
  const v3 zero = V3(0, 0, 0);
  const int idx = vcm_pick_light(c, smp);
  const light_rt *l = &c->s->lights[idx];
  const float pick = c->sel_pdf;
  const vcm_lsample sr = vcm_light_sample(c, l, it->pos, smp);
  if (is_black(sr.radiance)) return zero;
  const v3 r2l = vnormalize(vsub(sr.pos, sr.ref));
  const v3 bsdf_eval = bsdf_eval_all(b, r2l, it->wr);
  if (is_black(bsdf_eval)) return zero;
  float bsdf_pdf_w = bsdf_pdf_all(b, r2l, it->wr);
  float bsdf_rev_pdf_w = bsdf_pdf_all(b, it->wr, r2l);
  bsdf_pdf_w *= l->type == ORC_LIGHT_DIRECTIONAL ? 0.f : c->a->cont_prob;
  bsdf_rev_pdf_w *= c->a->cont_prob;
  const float w_light = bsdf_pdf_w / (pick * sr.pdf);
  const float cos_to_light = vdot(it->ns, r2l);
  const float cos_at_light = vdot(sr.normal, vneg(r2l));
  const float w_camera = (sr.emit_pdf * cos_to_light / (sr.pdf * cos_at_light)) *
                         (c->mis_vm + st->dVCM + st->dVC * bsdf_rev_pdf_w);
  const float mis = 1.f / (w_light + 1.f + w_camera);
  const v3 contrib = vscale(vmul(sr.radiance, bsdf_eval), mis * cos_to_light / (pick * sr.pdf));
  if (is_black(contrib)) return zero;
  orc_ray ray = make_ray(it->pos, r2l, ORC_EPS, vlength(vsub(sr.pos, sr.ref)));
  if (scene_occlude(c->s, &ray, &c->q)) return zero;
  return contrib;
}

/* ConnectVertices (vcm.cpp:474-542) */
static v3 vcm_connect(vcm_ctx *c, const vcm_state *cs, const vcm_vertex *lv, const inct *it,
                      const bsdf *b) {
  const v3 zero = V3(0, 0, 0);
  v3 direction = vsub(lv->hit_point, it->pos);
  const float dist2 = vdot(direction, direction);
  const float distance = vlength(direction);
  direction = vnormalize(direction);
  const v3 bsdf_eval = bsdf_eval_all(b, direction, it->wr);
  if (is_black(bsdf_eval)) return zero;
  float cam_dir_pdf = bsdf_pdf_all(b, direction, it->wr);
  float cam_rev_pdf = bsdf_pdf_all(b, it->wr, direction);
  cam_dir_pdf *= c->a->cont_prob;
  cam_rev_pdf *= c->a->cont_prob;
  const v3 light_eval = bsdf_eval_all(&lv->b, vneg(direction), lv->in_dir);
  if (is_black(light_eval)) return zero;
  float light_dir_pdf = bsdf_pdf_all(&lv->b, vneg(direction), lv->in_dir);
  float light_rev_pdf = bsdf_pdf_all(&lv->b, lv->in_dir, vneg(direction));
  light_dir_pdf *= c->a->cont_prob;
  light_rev_pdf *= c->a->cont_prob;
  const float g = vdot(lv->normal, vneg(direction)) * vdot(it->ns, direction) / dist2;
  if (g < 0.f) return zero;
  const float cam_dir_pdf_a = cam_dir_pdf * fabsf(vdot(lv->normal, vneg(direction))) / dist2;
  const float light_dir_pdf_a = light_dir_pdf * fabsf(vdot(it->ns, direction)) / dist2;
  const float w_light = cam_dir_pdf_a * (c->mis_vm + lv->dVCM + lv->dVC * light_rev_pdf);
  const float w_camera = light_dir_pdf_a * (c->mis_vm + cs->dVCM + cs->dVC * cam_rev_pdf);
  const float mis = 1.f / (w_light + 1.f + w_camera);
  const v3 contrib = vmul(vscale(bsdf_eval, mis * g), light_eval);
  if (is_black(contrib)) return zero;
  orc_ray sr = make_ray(it->pos, direction, ORC_EPS, distance * 0.9f);
  if (scene_occlude(c->s, &sr, &c->q)) return zero;
  return contrib;
}

/* HashGrid::GetCellIndex (hashgrid.h:144-164) */
static uint32_t vcm_cell(const vcm_ctx *c, int x, int y, int z) {
  const uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
  return ((ux * 73856093u) ^ (uy * 19349663u) ^ (uz * 83492791u)) % c->n_cells;
}
static uint32_t vcm_cell_of(const vcm_ctx *c, v3 p) {
  const v3 d = vsub(p, c->gbox.mn);
  return vcm_cell(c, (int)floorf(c->inv_cell * d.x), (int)floorf(c->inv_cell * d.y),
                  (int)floorf(c->inv_cell * d.z));
}

/* HashGrid::Build (hashgrid.h:34-70) */
static void vcm_grid_build(vcm_ctx *c) {
  c->cell_size = c->radius * 2.f;
  c->inv_cell = 1.f / c->cell_size;
  c->gbox = box_empty();
  for (uint32_t i = 0; i < c->n_lv; ++i) {
    c->gbox.mn = vmin(c->gbox.mn, c->lv[i].hit_point);
    c->gbox.mx = vmax(c->gbox.mx, c->lv[i].hit_point);
  }
  c->indices = (int *)realloc(c->indices, sizeof(int) * (c->n_lv ? c->n_lv : 1));
  memset(c->cell_ends, 0, sizeof(int) * c->n_cells);
  for (uint32_t i = 0; i < c->n_lv; ++i) c->cell_ends[vcm_cell_of(c, c->lv[i].hit_point)]++;
  int sum = 0;
  for (uint32_t i = 0; i < c->n_cells; ++i) {
    const int t = c->cell_ends[i];
    c->cell_ends[i] = sum;
    sum += t;
  }
  for (uint32_t i = 0; i < c->n_lv; ++i) {
    const int at = c->cell_ends[vcm_cell_of(c, c->lv[i].hit_point)]++;
    c->indices[at] = (int)i;
  }
}

/* HashGrid::Process with RangeQuery::Process (hashgrid.h:72-136, vcm.cpp:88-126) */
static v3 vcm_merge(const vcm_ctx *c, const vcm_state *cs, const inct *it, const bsdf *b) {
  v3 contrib = V3(0, 0, 0);
  const v3 q = it->pos;
  const v3 dmin = vsub(q, c->gbox.mn), dmax = vsub(c->gbox.mx, q);
  if (vminc(dmin) < 0.f || vminc(dmax) < 0.f) return contrib;
  const v3 cp = vscale(dmin, c->inv_cell);
  const v3 cf = V3(floorf(cp.x), floorf(cp.y), floorf(cp.z));
  const int px = (int)cf.x, py = (int)cf.y, pz = (int)cf.z;
  const v3 fr = vsub(cp, cf);
  const int pxo = px + (fr.x < 0.5f ? -1 : +1);
  const int pyo = py + (fr.y < 0.5f ? -1 : +1);
  const int pzo = pz + (fr.z < 0.5f ? -1 : +1);
  const int cells[8][3] = {{px, py, pz},   {px, py, pzo},   {px, pyo, pz},   {px, pyo, pzo},
                           {pxo, py, pz},  {pxo, py, pzo},  {pxo, pyo, pz},  {pxo, pyo, pzo}};
  for (int j = 0; j < 8; ++j) {
    const uint32_t ci = vcm_cell(c, cells[j][0], cells[j][1], cells[j][2]);
    int lo = ci == 0 ? 0 : c->cell_ends[ci - 1], hi = c->cell_ends[ci];
    for (; lo < hi; ++lo) {
      const vcm_vertex *lv = &c->lv[c->indices[lo]];
      const v3 d = vsub(q, lv->hit_point);
      if (!(vdot(d, d) <= c->radius_sqr)) continue;
      const uint32_t len = lv->path_length + cs->path_length;
      if (len > (uint32_t)c->a->max_path_length || (int)len < c->a->min_path_length) continue;
      const v3 light_dir = lv->in_dir;
      const v3 bsdf_eval = bsdf_eval_all(b, light_dir, it->wr);
      if (is_black(bsdf_eval)) continue;
      float cam_dir_pdf = bsdf_pdf_all(b, light_dir, it->wr);
      float cam_rev_pdf = bsdf_pdf_all(b, it->wr, light_dir);
      if (cam_dir_pdf == 0.f || cam_rev_pdf == 0.f) continue;
      cam_dir_pdf *= c->a->cont_prob;
      cam_rev_pdf *= c->a->cont_prob;
      const float w_light = lv->dVCM * c->mis_vc + lv->dVM * cam_dir_pdf;
      const float w_camera = cs->dVCM * c->mis_vc + cs->dVM * cam_rev_pdf;
      const float mis = 1.f / (w_light + 1.f + w_camera);
      contrib = vadd(contrib, vmul(vscale(bsdf_eval, mis), lv->throughput));
    }
  }
  return contrib;
}

/* Geometry::Pdf of an area light (for PdfBdpt, area_light.cpp:87-98) */
static float vcm_area_pdf(const light_rt *l) {
  float apdf = l->kind == ORC_GEOM_CUBE ? l->cube_pdf0 : 1.f / l->area;
  if (l->two_sided) apdf = apdf * 0.5f;
  return apdf;
}

static void vcm_light_path(vcm_ctx *c, uint32_t path_idx, sampler *smp) {
  const orc_scene *s = c->s;
  const orc_vcm_args *a = c->a;
  vcm_state st;
  memset(&st, 0, sizeof(st));
  /* GenerateLightSample (vcm.cpp:670-701) */
  const int idx = vcm_pick_light(c, smp);
  const light_rt *l = &s->lights[idx];
  const light_emit em = light_sample_emit(l, smp);
  st.throughput = vscale(em.radiance, fabsf(vdot(em.normal, em.dir)));
  st.origin = em.pos;
  st.direction = em.dir;
  const float emission_pdf = em.pdf_pos * em.pdf_dir * c->sel_pdf;
  st.throughput = vdivs(st.throughput, emission_pdf);
  st.path_length = 1;
  st.is_finite_light = l->type == ORC_LIGHT_AREA;
  st.dVCM = 1.f / em.pdf_dir;
  if (l->type != ORC_LIGHT_DIRECTIONAL) {
    const float used_cos = st.is_finite_light ? fabsf(vdot(em.normal, em.dir)) : 1.f;
    st.dVC = used_cos / emission_pdf;
  } else {
    st.dVC = 0.f;
  }
  st.dVM = st.dVC * c->mis_vc;
  for (;; ++st.path_length) {
    orc_ray ray = make_ray(st.origin, st.direction, ORC_EPS, INFINITY);
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, &c->q)) break;
    inct it = post_intersect(s, &ray, &h);
    const bsdf b = make_bsdf(&s->mats[s->prims[s->tri_prim[h.id]].material], it.ng, it.ns);
    (void)next1(smp), (void)next1(smp), (void)next1(smp); /* SampleAll(Next3D): is_delta only */
    if (st.path_length > 1 || st.is_finite_light) st.dVCM *= h.t * h.t;
    const float cos_in = fabsf(vdot(it.ns, vneg(vload(ray.d))));
    st.dVCM /= fabsf(cos_in);
    st.dVC /= fabsf(cos_in);
    st.dVM /= fabsf(cos_in);
    if ((a->use_vc || a->use_vm) && !(a->variant & ORC_VCM_LT_ONLY)) {
      if (c->n_lv == c->cap_lv) {
        c->cap_lv = c->cap_lv ? 2 * c->cap_lv : 1024;
        c->lv = (vcm_vertex *)realloc(c->lv, sizeof(vcm_vertex) * c->cap_lv);
      }
      vcm_vertex *v = &c->lv[c->n_lv++];
      v->hit_point = it.pos;
      v->in_dir = it.wr;
      v->throughput = st.throughput;
      v->path_length = st.path_length;
      v->b = b;
      v->normal = it.ns;
      v->dVCM = st.dVCM;
      v->dVC = st.dVC;
      v->dVM = st.dVM;
    }
    if ((a->use_vc || (a->variant & ORC_VCM_LT_ONLY)) &&
        (int)st.path_length + 1 >= a->min_path_length)
      vcm_connect_camera(c, &st, smp, &b, &it);
    if ((int)st.path_length + 2 > a->max_path_length) break;
    if (!vcm_scatter(c, &st, smp, &b, &it)) break;
  }
  c->path_ends[path_idx] = c->n_lv;
}

static v3 vcm_camera_path(vcm_ctx *c, uint32_t path_idx, sampler *smp) {
  const orc_scene *s = c->s;
  const orc_vcm_args *a = c->a;
  vcm_state st;
  memset(&st, 0, sizeof(st));
  /* GenerateCameraSample (vcm.cpp:703-732) */
  const int x = (int)path_idx % c->W, y = (int)path_idx / c->W;
  const float u = (float)x + next1(smp);
  const float v = (float)y + next1(smp);
  float lens[2];
  lens[0] = next1(smp);
  lens[1] = next1(smp);
  orc_ray primary;
  orc_camera_ray(c->cam, u, v, lens, &primary);
  const float camera_pdf = vcm_image_to_solid(c, vload(primary.d));
  st.origin = vload(primary.o);
  st.direction = vload(primary.d);
  st.throughput = V3(1.f, 1.f, 1.f);
  st.path_length = 1;
  st.specular_path = 1;
  st.dVCM = (a->variant & ORC_VCM_PIXEL_UNITS) ? c->light_count / (camera_pdf * c->pix_k)
                                               : 1.f / camera_pdf;
  v3 color = V3(0, 0, 0);
  for (;; ++st.path_length) {
    orc_ray ray = make_ray(st.origin, st.direction, ORC_EPS, INFINITY);
    orc_hit h;
    if (!scene_intersect(s, &ray, &h, &c->q)) {
      if (s->env_index >= 0 && (int)st.path_length >= a->min_path_length)
        color = vadd(color, vmul(st.throughput, env_radiance(&s->lights[s->env_index],
                                                             vload(ray.d))));
      break;
    }
    inct it = post_intersect(s, &ray, &h);
    const bsdf b = make_bsdf(&s->mats[s->prims[s->tri_prim[h.id]].material], it.ng, it.ns);
    (void)next1(smp), (void)next1(smp), (void)next1(smp); /* SampleAll(Next3D): is_delta only */
    {
      const float cos_in = fabsf(vdot(it.ns, vneg(vload(ray.d))));
      st.dVCM *= h.t * h.t;
      st.dVCM /= fabsf(cos_in);
      st.dVC /= fabsf(cos_in);
      st.dVM /= fabsf(cos_in);
    }
    const int li = s->prim_light[s->tri_prim[h.id]];
    if (li >= 0) {
      if ((int)st.path_length >= a->min_path_length) {
        const light_rt *l = &s->lights[li];
        const v3 radiance = vdot(it.ng, it.wr) > 0 ? l->radiance : V3(0, 0, 0);
        if (!is_black(radiance)) {
          if (st.path_length == 1) {
            color = vadd(color, vmul(st.throughput, radiance));
          } else if (a->use_vm && !a->use_vc) {
            if (st.specular_path) color = vadd(color, vmul(st.throughput, radiance));
          } else {
            const float apdf = vcm_area_pdf(l);
            float direct_pdf_a = apdf;
            float emission_pdf_w = std_max(0.f, vdot(it.ng, it.wr)) / ORC_PI;
            emission_pdf_w *= apdf;
            direct_pdf_a *= a->cont_prob;
            emission_pdf_w *= a->cont_prob;
            const float w_camera = direct_pdf_a * st.dVCM + emission_pdf_w * st.dVC;
            const float mis = 1.f / (1.f + w_camera);
            color = vadd(color, vmul(vscale(st.throughput, mis), radiance));
          }
        }
      }
      break;
    }
    if (a->variant & ORC_VCM_LT_ONLY) break;
    if ((int)st.path_length >= a->max_path_length) break;
    if (a->use_vc && (int)st.path_length + 1 >= a->min_path_length)
      color = vadd(color, vmul(st.throughput, vcm_direct(c, &st, &it, &b, smp)));
    if (a->use_vc) {
      const uint32_t lo = path_idx == 0 ? 0 : c->path_ends[path_idx - 1];
      const uint32_t hi = c->path_ends[path_idx];
      for (uint32_t i = lo; i < hi; ++i) {
        const vcm_vertex *lv = &c->lv[i];
        const int len = (int)(lv->path_length + 1 + st.path_length);
        if (len < a->min_path_length || len > a->max_path_length) continue;
        color = vadd(color, vmul(vmul(st.throughput, lv->throughput),
                                 vcm_connect(c, &st, lv, &it, &b)));
      }
    }
    if (a->use_vm) {
      const v3 m = vcm_merge(c, &st, &it, &b);
      color = vadd(color, vmul(vscale(st.throughput, c->vm_norm), m));
    }
    if (!vcm_scatter(c, &st, smp, &b, &it)) break;
  }
  return color;
}

/* The light or camera sub-paths of one iteration on n_threads threads. Light
 * paths: thread t traces the contiguous range t of paths into its own
 * vertex buffer, and the buffers are joined in thread order (path order).
 * Camera paths write only their own pixel. */
typedef struct vcm_job {
  vcm_ctx *c;
  vcm_ctx tc;
  uint32_t lo, hi, iter, camera;
  uint32_t *next; /* camera phase cursor, guarded by the atomic builtins */
} vcm_job;

static void *vcm_worker(void *arg) {
  vcm_job *j = (vcm_job *)arg;
  vcm_ctx *c = &j->tc;
  const uint32_t n = (uint32_t)(c->W * c->H);
  if (!j->camera) {
    for (uint32_t i = j->lo; i < j->hi; ++i) {
      sampler smp;
      orc_pcg32_seed(&smp.rng, sample_seed(c->a->seed_base, i, 2u * j->iter));
      vcm_light_path(c, i, &smp);
    }
    return NULL;
  }
  float *color_sum = (float *)j->c->camera_color;
  for (;;) {
    const uint32_t i0 = __atomic_fetch_add(j->next, 64u, __ATOMIC_RELAXED);
    if (i0 >= n) break;
    const uint32_t i1 = i0 + 64 < n ? i0 + 64 : n;
    for (uint32_t i = i0; i < i1; ++i) {
      sampler smp;
      orc_pcg32_seed(&smp.rng, sample_seed(c->a->seed_base, i, 2u * j->iter + 1u));
      const v3 col = vcm_camera_path(c, i, &smp);
      color_sum[3 * (size_t)i] += col.x;
      color_sum[3 * (size_t)i + 1] += col.y;
      color_sum[3 * (size_t)i + 2] += col.z;
    }
  }
  return NULL;
}

static void vcm_light_phase_mt(vcm_ctx *c, uint32_t iter, int n_threads) {
  const uint32_t n = (uint32_t)(c->W * c->H);
  vcm_job *jobs = (vcm_job *)calloc((size_t)n_threads, sizeof(vcm_job));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].c = c;
    jobs[t].tc = *c;
    jobs[t].tc.lv = NULL;
    jobs[t].tc.n_lv = jobs[t].tc.cap_lv = 0;
    jobs[t].tc.q = 0;
    jobs[t].tc.shared_splat = 1;
    jobs[t].lo = (uint32_t)((uint64_t)n * (uint64_t)t / (uint64_t)n_threads);
    jobs[t].hi = (uint32_t)((uint64_t)n * (uint64_t)(t + 1) / (uint64_t)n_threads);
    jobs[t].iter = iter;
    pthread_create(&th[t], NULL, vcm_worker, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  uint32_t total = 0;
  for (int t = 0; t < n_threads; ++t) total += jobs[t].tc.n_lv;
  if (total > c->cap_lv) {
    c->cap_lv = total;
    c->lv = (vcm_vertex *)realloc(c->lv, sizeof(vcm_vertex) * (total ? total : 1));
  }
  uint32_t base = 0;
  for (int t = 0; t < n_threads; ++t) {
    vcm_ctx *tc = &jobs[t].tc;
    if (tc->n_lv) memcpy(c->lv + base, tc->lv, sizeof(vcm_vertex) * tc->n_lv);
    for (uint32_t i = jobs[t].lo; i < jobs[t].hi; ++i) c->path_ends[i] += base;
    base += tc->n_lv;
    c->q += tc->q;
    free(tc->lv);
  }
  c->n_lv = total;
  free(th);
  free(jobs);
}

static void vcm_camera_phase_mt(vcm_ctx *c, uint32_t iter, float *color_sum, int n_threads) {
  vcm_job *jobs = (vcm_job *)calloc((size_t)n_threads, sizeof(vcm_job));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  uint32_t next = 0;
  c->camera_color = color_sum;
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].c = c;
    jobs[t].tc = *c;
    jobs[t].tc.q = 0;
    jobs[t].iter = iter;
    jobs[t].camera = 1;
    jobs[t].next = &next;
    pthread_create(&th[t], NULL, vcm_worker, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  for (int t = 0; t < n_threads; ++t) c->q += jobs[t].tc.q;
  free(th);
  free(jobs);
}

int64_t orc_render_vcm(const orc_scene *s, const orc_camera *cam, const orc_vcm_args *a,
                       float *color_sum, float *splat) {
  return orc_render_vcm_mt(s, cam, a, color_sum, splat, 1);
}

int64_t orc_render_vcm_mt(const orc_scene *s, const orc_camera *cam, const orc_vcm_args *a,
                          float *color_sum, float *splat, int n_threads) {
  vcm_ctx c;
  memset(&c, 0, sizeof(c));
  c.s = s;
  c.cam = cam;
  c.a = a;
  c.W = (int)cam->resolution[0];
  c.H = (int)cam->resolution[1];
  camera_derived(cam, c.w2c, c.c2r, c.w2r);
  c.cam_pos = xform_point(cam->camera2world, V3(0.f, 0.f, 0.f));
  c.cam_dir = vnormalize(xform_vector(cam->camera2world, V3(0.f, 0.f, -1.f)));
  c.splat = splat;
  const uint32_t n = (uint32_t)(c.W * c.H);
  memset(color_sum, 0, sizeof(float) * 3 * n);
  memset(splat, 0, sizeof(float) * 3 * n);
  if (s->n_lights == 0) return 0;
  {
    const uint32_t nl = s->n_lights;
    float *tab = (float *)malloc(sizeof(float) * nl);
    float sum = 0.f;
    for (uint32_t i = 0; i < nl; ++i) tab[i] = 1.f;
    for (uint32_t i = 0; i < nl; ++i) sum += tab[i];
    const float ratio = 1.f / sum;
    for (uint32_t i = 0; i < nl; ++i) tab[i] *= ratio;
    c.sel_pdf = tab[0];
    distrib_init(&c.sel, tab, nl, 0);
    free(tab);
  }
  c.path_ends = (uint32_t *)calloc(n, sizeof(uint32_t));
  c.n_cells = n;
  c.cell_ends = (int *)calloc(n, sizeof(int));
  int64_t n_vertices = 0;
  {
    /* one pixel's area on the plane at the focal distance */
    const v3 q0 = xform_point(cam->raster2camera, V3(0.f, 0.f, 0.f));
    const v3 q1 = xform_point(cam->raster2camera, V3(1.f, 0.f, 0.f));
    const v3 q2 = xform_point(cam->raster2camera, V3(0.f, 1.f, 0.f));
    const float fd = cam->focal_distance;
    const v3 p0 = vscale(q0, fd / fabsf(q0.z)), p1 = vscale(q1, fd / fabsf(q1.z)),
             p2 = vscale(q2, fd / fabsf(q2.z));
    c.pix_k = 1.f / vlength(vcross(vsub(p1, p0), vsub(p2, p0)));
  }
  const v3 wsize = vsub(s->world.mx, s->world.mn);
  for (int iter = 0; iter < a->iteration; ++iter) {
    /* RenderIter (vcm.cpp:161-200) */
    c.light_count = (float)(c.W * c.H);
    float radius = a->base_radius * vminc(wsize);
    radius /= powf((float)(iter + 1), 0.5f * (1.f - a->radius_alpha));
    radius = std_max(radius, 1e-7f);
    c.radius = radius;
    c.radius_sqr = radius * radius;
    c.vm_norm = 1.f / (c.radius_sqr * ORC_PI * c.light_count);
    const float eta = (ORC_PI * c.radius_sqr) * c.light_count;
    c.mis_vm = a->use_vm ? eta : 0.f;
    c.mis_vc = a->use_vc ? 1.f / eta : 0.f;
    c.n_lv = 0;
    if (n_threads > 1) {
      vcm_light_phase_mt(&c, (uint32_t)iter, n_threads);
    } else {
      for (uint32_t i = 0; i < n; ++i) {
        sampler smp;
        orc_pcg32_seed(&smp.rng, sample_seed(a->seed_base, i, 2u * (uint32_t)iter));
        vcm_light_path(&c, i, &smp);
      }
    }
    n_vertices = c.n_lv;
    if (a->use_vm) vcm_grid_build(&c);
    if (n_threads > 1) {
      vcm_camera_phase_mt(&c, (uint32_t)iter, color_sum, n_threads);
    } else {
      for (uint32_t i = 0; i < n; ++i) {
        sampler smp;
        orc_pcg32_seed(&smp.rng, sample_seed(a->seed_base, i, 2u * (uint32_t)iter + 1u));
        const v3 col = vcm_camera_path(&c, i, &smp);
        color_sum[3 * (size_t)i] += col.x;
        color_sum[3 * (size_t)i + 1] += col.y;
        color_sum[3 * (size_t)i + 2] += col.z;
      }
    }
  }
  distrib_free(&c.sel);
  free(c.path_ends);
  free(c.cell_ends);
  free(c.indices);
  free(c.lv);
  return n_vertices;
}

/* Test hooks: ThinLensCamera::ToRaster and ImageAreaToSolidAngle with the
 * camera matrices VCM derives. */
int orc_vcm_to_raster(const orc_camera *cam, const float *u2, const float *pos, float *dir_out,
                      float *raster_out) {
  vcm_ctx c;
  memset(&c, 0, sizeof(c));
  c.cam = cam;
  camera_derived(cam, c.w2c, c.c2r, c.w2r);
  c.cam_pos = xform_point(cam->camera2world, V3(0.f, 0.f, 0.f));
  c.cam_dir = vnormalize(xform_vector(cam->camera2world, V3(0.f, 0.f, -1.f)));
  v2 u = {u2[0], u2[1]};
  v3 d;
  v2 r;
  if (!vcm_to_raster(&c, u, vload(pos), &d, &r)) return 0;
  vstore(dir_out, d);
  raster_out[0] = r.x;
  raster_out[1] = r.y;
  return 1;
}
float orc_vcm_image_to_solid(const orc_camera *cam, const float *dir) {
  vcm_ctx c;
  memset(&c, 0, sizeof(c));
  c.cam = cam;
  c.cam_dir = vnormalize(xform_vector(cam->camera2world, V3(0.f, 0.f, -1.f)));
  return vcm_image_to_solid(&c, vload(dir));
}
