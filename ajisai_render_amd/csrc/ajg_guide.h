// ajg_guide.h — learned path guiding: the network's layout, the input
// encoding and the von Mises-Fisher mixture that the network parameterises.
// (No reference counterpart: the reference samples continuation directions
// from the BSDF only, pt.cpp:215-222.)
//
// Network: x(112) -> 64 -> 64 -> 64 -> 48, ReLU on the hidden layers, BF16
// weights and activations, FP32 accumulation (MFMA 16x16x16 bf16).
// Input features x[16j + 4g + i], j = 0..6, g = 0..3, i = 0..3:
//   j < 6: component c = j / 2, harmonic h = 8 (j % 2) + 2g + i / 2,
//          x = sin(pi (h+1) n_c) for even i, cos(pi (h+1) n_c) for odd i;
//   j = 6: x[96 + t] = p_hat_t for t < 3, zero for features 99..111;
//   p_hat = 2 * (p - box_min) / (box_max - box_min) - 1.
// (Every 4 consecutive features are two whole harmonics of one component, so
// each lane of an MFMA operand fragment computes its own four in place.)
// Output o[0..39] (o[40..47] unused): lobe j uses o[5j..5j+4] =
//   (mu_raw xyz, kappa_raw, weight logit);
//   mu = mu_raw / |mu_raw|, kappa = exp(clamp(kappa_raw, -3, 7)), pi = softmax.
#pragma once

#include "ajg_common.h"

namespace ajg {

static constexpr int kGIn = 112;
static constexpr int kGUsedIn = 99;
static constexpr int kGHid = 64;
static constexpr int kGOut = 48;
static constexpr int kGUsedOut = 40;
static constexpr int kGLobes = 8;
static constexpr int kGFreq = 16;
// parameter offsets (floats) in the master FP32 vector: W[in][out] then b
static constexpr int kGW1 = 0;
static constexpr int kGB1 = kGW1 + kGIn * kGHid;
static constexpr int kGW2 = kGB1 + kGHid;
static constexpr int kGB2 = kGW2 + kGHid * kGHid;
static constexpr int kGW3 = kGB2 + kGHid;
static constexpr int kGB3 = kGW3 + kGHid * kGHid;
static constexpr int kGW4 = kGB3 + kGHid;
static constexpr int kGB4 = kGW4 + kGHid * kGOut;
static constexpr int kGParams = kGB4 + kGOut;  // 18672
static constexpr float kKappaLo = -3.f, kKappaHi = 7.f;

struct VmfLobe {
  f3 mu;
  float kappa, w;
};

// Decode network outputs into a normalised mixture.
__host__ __device__ __forceinline__ void guide_decode(const float *o, VmfLobe *lobes) {
  float mx = -3.4e38f;
  for (int j = 0; j < kGLobes; ++j) mx = fmaxf(mx, o[5 * j + 4]);
  float sum = 0.f;
  for (int j = 0; j < kGLobes; ++j) {
    const float mx0 = o[5 * j], my = o[5 * j + 1], mz = o[5 * j + 2];
    const float len = fmaxf(sqrtf(mx0 * mx0 + my * my + mz * mz), 1e-6f);
    lobes[j].mu = mk3(mx0 / len, my / len, mz / len);
    lobes[j].kappa = expf(fminf(fmaxf(o[5 * j + 3], kKappaLo), kKappaHi));
    lobes[j].w = expf(o[5 * j + 4] - mx);
    sum += lobes[j].w;
  }
  for (int j = 0; j < kGLobes; ++j) lobes[j].w /= sum;
}

// vMF density kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
__host__ __device__ __forceinline__ float vmf_pdf(const VmfLobe &l, f3 w) {
  const float k = l.kappa;
  const float norm = k / (6.28318531f * -expm1f(-2.f * k));
  return norm * expf(k * (l.mu.x * w.x + l.mu.y * w.y + l.mu.z * w.z - 1.f));
}

__host__ __device__ __forceinline__ float mixture_pdf(const VmfLobe *lobes, f3 w) {
  float p = 0.f;
  for (int j = 0; j < kGLobes; ++j) p += lobes[j].w * vmf_pdf(lobes[j], w);
  return p;
}

// Direction from one vMF lobe: w = 1 + log(u1 + (1 - u1) exp(-2 kappa)) / kappa
// (Jakob 2012, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2"), azimuth 2 pi u2 in a frame around mu (Duff et al.
// 2017, "Building an orthonormal basis, revisited").
__device__ __forceinline__ f3 vmf_sample_dir(const VmfLobe &l, float u1, float u2) {
  const float k = l.kappa;
  float wz = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k;
  wz = fminf(fmaxf(wz, -1.f), 1.f);
  const float s = sqrtf(fmaxf(0.f, 1.f - wz * wz));
  float sp, cp;
  sincosf(6.28318531f * u2, &sp, &cp);
  const f3 n = l.mu;
  const float sign = copysignf(1.f, n.z);
  const float a = -1.f / (sign + n.z);
  const float b = n.x * n.y * a;
  const f3 t1 = mk3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  const f3 t2 = mk3(b, sign + n.y * n.y * a, -n.y);
  return normalize(add(add(scale(t1, s * cp), scale(t2, s * sp)), scale(n, wz)));
}

// Sample the mixture: u0 picks the lobe (the first whose running weight sum
// exceeds u0, else the last), (u1, u2) the direction. The chosen lobe is
// copied out by a fully unrolled scan, so the lobes stay in registers (an
// indexed read would place them in scratch memory).
__device__ __forceinline__ f3 mixture_sample(const VmfLobe *lobes, float u0, float u1, float u2) {
  VmfLobe sel = lobes[kGLobes - 1];
  float c = 0.f;
  bool found = false;
#pragma unroll
  for (int j = 0; j < kGLobes - 1; ++j) {
    c += lobes[j].w;
    if (!found && u0 < c) {
      sel = lobes[j];
      found = true;
    }
  }
  return vmf_sample_dir(sel, u1, u2);
}

// The mixture as the renderer stores it per closest-hit queue entry: 40
// FP16 values, lobe j at m[5j..5j+4] = (mu xyz, kappa, weight). Sampling and
// density both decode the stored values the same way (mu renormalised,
// weights divided by their sum), so the density is exactly that of the
// sampler whatever the FP16 rounding did.
struct GMixLobes {
  VmfLobe l[kGLobes];
};

__device__ __forceinline__ GMixLobes gmix_load(const _Float16 *m) {
  GMixLobes g;
  float sw = 0.f;
#pragma unroll
  for (int j = 0; j < kGLobes; ++j) {
    g.l[j].mu = normalize(mk3((float)m[5 * j], (float)m[5 * j + 1], (float)m[5 * j + 2]));
    g.l[j].kappa = (float)m[5 * j + 3];
    g.l[j].w = (float)m[5 * j + 4];
    sw += g.l[j].w;
  }
  const float inv = 1.f / sw;
#pragma unroll
  for (int j = 0; j < kGLobes; ++j) g.l[j].w *= inv;
  return g;
}

__device__ __forceinline__ float gmix_pdf(const _Float16 *m, f3 w) {
  const GMixLobes g = gmix_load(m);
  return mixture_pdf(g.l, w);
}

__device__ __forceinline__ f3 gmix_sample(const _Float16 *m, float u0, float u1, float u2) {
  const GMixLobes g = gmix_load(m);
  return mixture_sample(g.l, u0, u1, u2);
}

}  // namespace ajg
