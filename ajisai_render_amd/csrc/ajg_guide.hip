// ajg_guide.hip — the path-guiding network on gfx950: BF16 MFMA forward and
// backward kernels with FP32 accumulation, the vMF-mixture loss, Adam.
//
// Every GEMM runs on v_mfma_f32_16x16x16bf16_1k tiles owned by one wave:
//   A (16 x 16, rows = samples or features) lane l holds row l%16, k = 4(l/16)..+3
//   B (16 x 16) lane l holds column l%16, k = 4(l/16)..+3
//   D (16 x 16) lane l holds column l%16, rows 4(l/16)..+3
// Weights are kept as FP32 masters W[in][out] plus two BF16 copies:
// out-major WT[out][in] (B operand of the forward pass) and in-major
// W[in][out] (B operand of the backward data pass).
// Training is layer-wise over the whole batch: forward (activations to HBM),
// loss gradient, backward data (dX), backward weights (per-block partial
// sums, reduced in a fixed order, so gradients are deterministic), Adam.
#include <cmath>
#include <cstring>
#include <vector>

#include "ajg_guide.h"
#include "ajg_internal.h"

struct ajg_guide {
  ajg_ctx *ctx = nullptr;
  float box[6] = {0, 0, 0, 1, 1, 1};
  float *w = nullptr;        // FP32 masters (kGParams)
  float *m = nullptr, *v = nullptr;  // Adam moments
  uint16_t *wt = nullptr;    // BF16 out-major copies of the 4 weight matrices
  uint16_t *wi = nullptr;    // BF16 in-major copies
  float *grad = nullptr;     // reduced gradient (kGParams)
  int step = 0;
  // training scratch, grown on demand
  uint32_t cap = 0;
  uint16_t *x0 = nullptr, *x1 = nullptr, *x2 = nullptr, *x3 = nullptr;
  uint16_t *dy4 = nullptr, *dx3 = nullptr, *dx2 = nullptr, *dx1 = nullptr;
  float *out = nullptr;
  float *partial = nullptr;  // per-block partial gradients
  float *scal = nullptr;     // [0] sum of weights, [1] loss
};

namespace ajg {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int kGThreads = 256;
static constexpr int kGWaves = kGThreads / 64;
static constexpr int kGradBlocks = 256;

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// four consecutive BF16 values (8-byte aligned in every caller)
__device__ __forceinline__ s16x4 ld4(const uint16_t *p) {
  return *reinterpret_cast<const s16x4 *>(p);
}
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// Y(16 x N) = X(16 x K, LDS row-major, stride ldx) * W, W given out-major
// (WT[n][k], LDS, stride K). acc[t] is N-tile t.
template <int K, int N>
__device__ __forceinline__ void tile_gemm(const uint16_t *X, int ldx, const uint16_t *WT, f32x4 *acc) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < N / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 16) {
    const s16x4 a = ld4(X + r * ldx + k0 + 4 * g);
#pragma unroll
    for (int t = 0; t < N / 16; ++t) {
      const s16x4 b = ld4(WT + (t * 16 + r) * K + k0 + 4 * g);
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
}

// Store D tiles (+bias, optional ReLU) as BF16 rows into LDS Y (stride ldy).
template <int N>
__device__ __forceinline__ void tile_store(const f32x4 *acc, const float *bias, bool relu,
                                           uint16_t *Y, int ldy) {
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < N / 16; ++t) {
    const float b = bias[t * 16 + c];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float y = acc[t][i] + b;
      if (relu) y = y > 0.f ? y : 0.f;
      Y[(4 * g + i) * ldy + t * 16 + c] = f2bf(y);
    }
  }
}

struct alignas(16) GuideLds {
  uint16_t wt1[kGHid * kGIn], wt2[kGHid * kGHid], wt3[kGHid * kGHid], wt4[kGOut * kGHid];
  float b[kGHid * 3 + kGOut];
  uint16_t xa[kGWaves][16 * kGIn];
  uint16_t xb[kGWaves][16 * kGHid];
};

__device__ void load_weights_lds(GuideLds &s, const uint16_t *wt, const float *w) {
  const uint16_t *src[4] = {wt, wt + kGIn * kGHid, wt + kGIn * kGHid + kGHid * kGHid,
                            wt + kGIn * kGHid + 2 * kGHid * kGHid};
  uint16_t *dst[4] = {s.wt1, s.wt2, s.wt3, s.wt4};
  const int cnt[4] = {kGHid * kGIn, kGHid * kGHid, kGHid * kGHid, kGOut * kGHid};
  for (int l = 0; l < 4; ++l)
    for (int i = threadIdx.x; i < cnt[l]; i += kGThreads) dst[l][i] = src[l][i];
  for (int i = threadIdx.x; i < kGHid; i += kGThreads) {
    s.b[i] = w[kGB1 + i];
    s.b[kGHid + i] = w[kGB2 + i];
    s.b[2 * kGHid + i] = w[kGB3 + i];
  }
  for (int i = threadIdx.x; i < kGOut; i += kGThreads) s.b[3 * kGHid + i] = w[kGB4 + i];
  __syncthreads();
}

// Encode 16 samples into xa: lane l writes features [28 (l/16), +28) of
// sample l%16 (one sincos per feature; no per-lane feature array).
__device__ __forceinline__ void encode_tile(uint16_t *xa, const float *pos, const float *nrm,
                                            uint32_t base, uint32_t n, uint32_t stride,
                                            const float *box) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const uint32_t s = base + r;
  f3 p = mk3(0.f, 0.f, 0.f), nn = mk3(0.f, 0.f, 1.f);
  const bool live = s < n;
  if (live) {
    p = ld3(pos + (size_t)s * stride);
    nn = ld3(nrm + (size_t)s * stride);
  }
  for (int i = 0; i < 28; ++i) {
    const int f = 28 * g + i;
    const float x = live ? guide_feature(p, nn, box, f) : 0.f;
    xa[r * kGIn + f] = f2bf(x);
  }
}

// Forward pass over a batch. pos/nrm: float rows with `stride` floats; the
// batch size is min(n, *n_dev) when n_dev is given.
// mode 0: inference, writes the decoded mixture (40 floats / sample) to
// `mix` and, if non-null, the raw outputs (48 / sample) to `raw`.
// mode 1: training, writes activations x0..x3 (BF16) and raw outputs.
__global__ void __launch_bounds__(kGThreads)
k_guide_fwd(const uint16_t *wt, const float *w, const float *pos, const float *nrm, uint32_t n,
            const uint32_t *n_dev, uint32_t stride, float bx0, float by0, float bz0, float bx1, float by1, float bz1,
            int mode, float *mix, float *raw, uint16_t *x0, uint16_t *x1, uint16_t *x2,
            uint16_t *x3) {
  __shared__ GuideLds s;
  if (n_dev) n = min(n, *n_dev);  // queue size written on the device
  if (n == 0) return;
  load_weights_lds(s, wt, w);
  const float box[6] = {bx0, by0, bz0, bx1, by1, bz1};
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint16_t *xa = s.xa[wv], *xb = s.xb[wv];
  const uint32_t waves = gridDim.x * kGWaves;
  for (uint32_t chunk = blockIdx.x * kGWaves + wv; chunk * 16 < n; chunk += waves) {
    const uint32_t base = chunk * 16;
    encode_tile(xa, pos, nrm, base, n, stride, box);
    wave_sync();
    f32x4 acc[4];
    tile_gemm<kGIn, kGHid>(xa, kGIn, s.wt1, acc);
    if (mode == 1)
      for (int i = lane; i < 16 * kGIn; i += 64) {
        const uint32_t sm = base + i / kGIn;
        if (sm < n) x0[(size_t)sm * kGIn + i % kGIn] = xa[i];
      }
    wave_sync();
    tile_store<kGHid>(acc, s.b, true, xb, kGHid);
    wave_sync();
    tile_gemm<kGHid, kGHid>(xb, kGHid, s.wt2, acc);
    if (mode == 1)
      for (int i = lane; i < 16 * kGHid; i += 64) {
        const uint32_t sm = base + i / kGHid;
        if (sm < n) x1[(size_t)sm * kGHid + i % kGHid] = xb[i];
      }
    wave_sync();
    tile_store<kGHid>(acc, s.b + kGHid, true, xa, kGHid);
    wave_sync();
    tile_gemm<kGHid, kGHid>(xa, kGHid, s.wt3, acc);
    if (mode == 1)
      for (int i = lane; i < 16 * kGHid; i += 64) {
        const uint32_t sm = base + i / kGHid;
        if (sm < n) x2[(size_t)sm * kGHid + i % kGHid] = xa[i];
      }
    wave_sync();
    tile_store<kGHid>(acc, s.b + 2 * kGHid, true, xb, kGHid);
    wave_sync();
    f32x4 o[3];
    tile_gemm<kGHid, kGOut>(xb, kGHid, s.wt4, o);
    if (mode == 1)
      for (int i = lane; i < 16 * kGHid; i += 64) {
        const uint32_t sm = base + i / kGHid;
        if (sm < n) x3[(size_t)sm * kGHid + i % kGHid] = xb[i];
      }
    // raw outputs: lane holds rows 4g..4g+3 of column c of tile t
    const int c = lane & 15, g = lane >> 4;
    float *stage = reinterpret_cast<float *>(xa);  // 16 x 48 floats fit in xa (16*112*2 bytes)
    wave_sync();
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) stage[(4 * g + i) * kGOut + t * 16 + c] = o[t][i] + s.b[3 * kGHid + t * 16 + c];
    wave_sync();
    if (raw)
      for (int i = lane; i < 16 * kGOut; i += 64) {
        const uint32_t sm = base + i / kGOut;
        if (sm < n) raw[(size_t)sm * kGOut + i % kGOut] = stage[i];
      }
    if (mode == 0 && mix && lane < 16 && base + lane < n) {
      VmfLobe lobes[kGLobes];
      guide_decode(stage + lane * kGOut, lobes);
      float *dst = mix + (size_t)(base + lane) * kGUsedOut;
      for (int j = 0; j < kGLobes; ++j) {
        dst[5 * j] = lobes[j].mu.x;
        dst[5 * j + 1] = lobes[j].mu.y;
        dst[5 * j + 2] = lobes[j].mu.z;
        dst[5 * j + 3] = lobes[j].kappa;
        dst[5 * j + 4] = lobes[j].w;
      }
    }
    wave_sync();
  }
}

// Sum of sample weights (training normalisation).
__global__ void __launch_bounds__(256)
k_guide_wsum(const float *samples, uint32_t n, float *scal) {
  __shared__ float red[256];
  float acc = 0.f;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc += samples[(size_t)i * 10 + 9];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(scal, red[0]);
}

// Loss -sum_k w_k log q(d_k) / sum_k w_k and its gradient w.r.t. the raw
// outputs (softmax logits, log-kappa with clamp, unnormalised mean
// directions).
__global__ void __launch_bounds__(256)
k_guide_loss(const float *raw, const float *samples, uint32_t n, const float *scal,
             uint16_t *dy4, float *loss) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  float lcontrib = 0.f;
  if (i < n) {
    const float *o = raw + (size_t)i * kGOut;
    const float *smp = samples + (size_t)i * 10;
    const f3 d = mk3(smp[6], smp[7], smp[8]);
    const float wgt = smp[9] / fmaxf(scal[0], 1e-30f);
    VmfLobe lobes[kGLobes];
    guide_decode(o, lobes);
    float qj[kGLobes], q = 0.f;
    for (int j = 0; j < kGLobes; ++j) {
      qj[j] = vmf_pdf(lobes[j], d);
      q += lobes[j].w * qj[j];
    }
    q = fmaxf(q, 1e-30f);
    lcontrib = -wgt * logf(q);
    float g[kGOut];
    for (int k = 0; k < kGOut; ++k) g[k] = 0.f;
    for (int j = 0; j < kGLobes; ++j) {
      const float r = lobes[j].w * qj[j] / q;  // responsibility
      // d log q / d logit_j = r_j - pi_j
      g[5 * j + 4] = -wgt * (r - lobes[j].w);
      // d log q_j / d kappa = 1/kappa - 2 e^{-2k}/(1 - e^{-2k}) + (mu.d - 1)
      const float k = lobes[j].kappa;
      const float md = dot(lobes[j].mu, d);
      const float dk = 1.f / k - 2.f / expm1f(2.f * k) + (md - 1.f);
      const float kr = o[5 * j + 3];
      const float dkr = (kr > kKappaLo && kr < kKappaHi) ? k * dk : 0.f;
      g[5 * j + 3] = -wgt * r * dkr;
      // d log q_j / d mu_raw = kappa (d - (mu.d) mu) / |mu_raw|
      const float len = fmaxf(sqrtf(o[5 * j] * o[5 * j] + o[5 * j + 1] * o[5 * j + 1] +
                                    o[5 * j + 2] * o[5 * j + 2]), 1e-6f);
      g[5 * j + 0] = -wgt * r * k * (d.x - md * lobes[j].mu.x) / len;
      g[5 * j + 1] = -wgt * r * k * (d.y - md * lobes[j].mu.y) / len;
      g[5 * j + 2] = -wgt * r * k * (d.z - md * lobes[j].mu.z) / len;
    }
    for (int k = 0; k < kGOut; ++k) dy4[(size_t)i * kGOut + k] = f2bf(g[k]);
  }
  // block reduction of the loss
  __shared__ float red[256];
  red[threadIdx.x] = lcontrib;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Backward data pass: dx3 = (dy4 W4^T) * [x3 > 0], dx2 = (dx3 W3^T) * [x2>0],
// dx1 = (dx2 W2^T) * [x1 > 0]. The B operand B[k = out][n = in] =
// W[in][out] reads four consecutive `out` of the in-major copy.
struct alignas(16) GuideBwdLds {
  uint16_t w2[kGHid * kGHid], w3[kGHid * kGHid], w4[kGHid * kGOut];
  uint16_t a[kGWaves][16 * kGHid];  // dY4 (16 x 48), then dX2 (16 x 64)
  uint16_t b[kGWaves][16 * kGHid];
};

template <int K, int N>
__device__ __forceinline__ void tile_gemm_bt(const uint16_t *X, int ldx, const uint16_t *Win,
                                             f32x4 *acc) {
  // Y(16 x N) = X(16 x K) * W^T where W is in-major [N][K] (N = in, K = out)
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < N / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 16) {
    const s16x4 a = ld4(X + r * ldx + k0 + 4 * g);
#pragma unroll
    for (int t = 0; t < N / 16; ++t) {
      const s16x4 b = ld4(Win + (t * 16 + r) * K + k0 + 4 * g);
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
}

template <int N>
__device__ __forceinline__ void tile_store_masked(const f32x4 *acc, const uint16_t *xg, uint32_t base,
                                                  uint32_t n, uint16_t *Y, uint16_t *dst) {
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < N / 16; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t sm = base + 4 * g + i;
      float v = acc[t][i];
      if (sm < n) {
        if (bf2f(xg[(size_t)sm * N + t * 16 + c]) <= 0.f) v = 0.f;
      } else {
        v = 0.f;
      }
      const uint16_t h = f2bf(v);
      Y[(4 * g + i) * N + t * 16 + c] = h;
      if (sm < n) dst[(size_t)sm * N + t * 16 + c] = h;
    }
}

__global__ void __launch_bounds__(kGThreads)
k_guide_bwd_data(const uint16_t *wi, const uint16_t *dy4, const uint16_t *x1, const uint16_t *x2,
                 const uint16_t *x3, uint32_t n, uint16_t *dx3, uint16_t *dx2, uint16_t *dx1) {
  __shared__ GuideBwdLds s;
  // in-major copies: W1 (112 x 64) first, then W2, W3 (64 x 64), W4 (64 x 48)
  const uint16_t *w2 = wi + kGIn * kGHid, *w3 = w2 + kGHid * kGHid, *w4 = w3 + kGHid * kGHid;
  for (int i = threadIdx.x; i < kGHid * kGHid; i += kGThreads) {
    s.w2[i] = w2[i];
    s.w3[i] = w3[i];
  }
  for (int i = threadIdx.x; i < kGHid * kGOut; i += kGThreads) s.w4[i] = w4[i];
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * kGWaves;
  for (uint32_t chunk = blockIdx.x * kGWaves + wv; chunk * 16 < n; chunk += waves) {
    const uint32_t base = chunk * 16;
    uint16_t *a = s.a[wv], *b = s.b[wv];
    for (int i = lane; i < 16 * kGOut; i += 64) {
      const uint32_t sm = base + i / kGOut;
      a[i] = sm < n ? dy4[(size_t)sm * kGOut + i % kGOut] : (uint16_t)0;
    }
    wave_sync();
    f32x4 acc[4];
    tile_gemm_bt<kGOut, kGHid>(a, kGOut, s.w4, acc);
    wave_sync();
    tile_store_masked<kGHid>(acc, x3, base, n, b, dx3);
    wave_sync();
    tile_gemm_bt<kGHid, kGHid>(b, kGHid, s.w3, acc);
    wave_sync();
    tile_store_masked<kGHid>(acc, x2, base, n, a, dx2);
    wave_sync();
    tile_gemm_bt<kGHid, kGHid>(a, kGHid, s.w2, acc);
    wave_sync();
    tile_store_masked<kGHid>(acc, x1, base, n, b, dx1);
    wave_sync();
  }
}

// Backward weight pass for one layer: dW[in][out] = sum_s X[s][in] dY[s][out],
// db[out] = sum_s dY[s][out]. MFMA with M = in, N = out, K = 16 samples.
// Each block writes its partial sums (deterministic reduction afterwards).
template <int KIN, int NOUT>
__global__ void __launch_bounds__(kGThreads)
k_guide_bwd_weight(const uint16_t *X, const uint16_t *DY, uint32_t n, float *partial,
                   int w_off, int b_off) {
  __shared__ uint16_t sx[kGWaves][16 * KIN];
  __shared__ uint16_t sd[kGWaves][16 * NOUT];
  __shared__ float red[KIN * NOUT + NOUT];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  constexpr int TM = KIN / 16, TN = NOUT / 16;
  f32x4 acc[TM * TN];
#pragma unroll
  for (int t = 0; t < TM * TN; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float dbias = 0.f;  // lane < NOUT: column sum
  const uint32_t waves = gridDim.x * kGWaves;
  for (uint32_t chunk = blockIdx.x * kGWaves + wv; chunk * 16 < n; chunk += waves) {
    const uint32_t base = chunk * 16;
    for (int i = lane; i < 16 * KIN; i += 64) {
      const uint32_t sm = base + i / KIN;
      sx[wv][i] = sm < n ? X[(size_t)sm * KIN + i % KIN] : (uint16_t)0;
    }
    for (int i = lane; i < 16 * NOUT; i += 64) {
      const uint32_t sm = base + i / NOUT;
      sd[wv][i] = sm < n ? DY[(size_t)sm * NOUT + i % NOUT] : (uint16_t)0;
    }
    wave_sync();
#pragma unroll
    for (int tm = 0; tm < TM; ++tm) {
      s16x4 a;  // A[m = in][k = sample] = X[sample][in]
      a.x = (short)sx[wv][(4 * g + 0) * KIN + tm * 16 + r];
      a.y = (short)sx[wv][(4 * g + 1) * KIN + tm * 16 + r];
      a.z = (short)sx[wv][(4 * g + 2) * KIN + tm * 16 + r];
      a.w = (short)sx[wv][(4 * g + 3) * KIN + tm * 16 + r];
#pragma unroll
      for (int tn = 0; tn < TN; ++tn) {
        s16x4 b;  // B[k = sample][n = out] = DY[sample][out]
        b.x = (short)sd[wv][(4 * g + 0) * NOUT + tn * 16 + r];
        b.y = (short)sd[wv][(4 * g + 1) * NOUT + tn * 16 + r];
        b.z = (short)sd[wv][(4 * g + 2) * NOUT + tn * 16 + r];
        b.w = (short)sd[wv][(4 * g + 3) * NOUT + tn * 16 + r];
        acc[tm * TN + tn] = mfma16(a, b, acc[tm * TN + tn]);
      }
    }
    if (lane < NOUT)
      for (int i = 0; i < 16; ++i) dbias += bf2f(sd[wv][i * NOUT + lane]);
    wave_sync();
  }
  // reduce the 4 waves in LDS in a fixed order
  for (int i = threadIdx.x; i < KIN * NOUT + NOUT; i += kGThreads) red[i] = 0.f;
  __syncthreads();
  for (int w = 0; w < kGWaves; ++w) {
    if (wv == w) {
#pragma unroll
      for (int tm = 0; tm < TM; ++tm)
#pragma unroll
        for (int tn = 0; tn < TN; ++tn)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            red[(tm * 16 + 4 * g + i) * NOUT + tn * 16 + r] += acc[tm * TN + tn][i];
      if (lane < NOUT) red[KIN * NOUT + lane] += dbias;
    }
    __syncthreads();
  }
  float *dst = partial + (size_t)blockIdx.x * kGParams;
  for (int i = threadIdx.x; i < KIN * NOUT; i += kGThreads) dst[w_off + i] = red[i];
  for (int i = threadIdx.x; i < NOUT; i += kGThreads) dst[b_off + i] = red[KIN * NOUT + i];
}

__global__ void __launch_bounds__(256)
k_guide_reduce(const float *partial, int blocks, float *grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kGParams) return;
  float s = 0.f;
  for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * kGParams + i];
  grad[i] = s;
}

// Adam on the FP32 masters, then refresh the BF16 copies.
__global__ void __launch_bounds__(256)
k_guide_adam(float *w, float *m, float *v, const float *g, float lr, float b1c, float b2c) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kGParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + eps);
}

__global__ void __launch_bounds__(256)
k_guide_cast(const float *w, uint16_t *wt, uint16_t *wi) {
  // layer l: W[in][out] at offset; wt = out-major, wi = in-major
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int offs[4] = {kGW1, kGW2, kGW3, kGW4};
  const int ins[4] = {kGIn, kGHid, kGHid, kGHid};
  const int outs[4] = {kGHid, kGHid, kGHid, kGOut};
  int base = 0;
  for (int l = 0; l < 4; ++l) {
    const int cnt = ins[l] * outs[l];
    if (i >= base && i < base + cnt) {
      const int k = i - base, in = k / outs[l], out = k % outs[l];
      const uint16_t h = f2bf(w[offs[l] + k]);
      wi[base + k] = h;
      wt[base + out * ins[l] + in] = h;
    }
    base += cnt;
  }
}

static constexpr int kGWeightCount = kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut;

static int ensure_train(ajg_guide *gd, uint32_t n) {
  if (gd->cap >= n) return AJG_OK;
  void *ps[] = {gd->x0, gd->x1, gd->x2, gd->x3, gd->dy4, gd->dx3, gd->dx2, gd->dx1, gd->out};
  for (void *p : ps) (void)hipFree(p);
  int st;
  if ((st = dev_alloc(gd->ctx, (void **)&gd->x0, 2ull * n * kGIn, "guide x0")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->x1, 2ull * n * kGHid, "guide x1")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->x2, 2ull * n * kGHid, "guide x2")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->x3, 2ull * n * kGHid, "guide x3")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->dy4, 2ull * n * kGOut, "guide dy4")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->dx3, 2ull * n * kGHid, "guide dx3")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->dx2, 2ull * n * kGHid, "guide dx2")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->dx1, 2ull * n * kGHid, "guide dx1")) ||
      (st = dev_alloc(gd->ctx, (void **)&gd->out, 4ull * n * kGOut, "guide out"))) {
    gd->cap = 0;
    return st;
  }
  gd->cap = n;
  return AJG_OK;
}

static void guide_free(ajg_guide *gd) {
  if (!gd) return;
  void *ps[] = {gd->w, gd->m, gd->v, gd->wt, gd->wi, gd->grad, gd->x0, gd->x1, gd->x2, gd->x3,
                gd->dy4, gd->dx3, gd->dx2, gd->dx1, gd->out, gd->partial, gd->scal};
  for (void *p : ps) (void)hipFree(p);
  delete gd;
}

int guide_forward_launch(ajg_guide *gd, const float *pos, const float *nrm, uint32_t n,
                         const uint32_t *n_dev, uint32_t stride, float *mix, float *raw,
                         hipStream_t s) {
  if (n == 0) return AJG_OK;
  const uint32_t blocks = std::min<uint32_t>(1024, (n + 16 * kGWaves - 1) / (16 * kGWaves));
  hipLaunchKernelGGL(k_guide_fwd, dim3(blocks), dim3(kGThreads), 0, s, gd->wt, gd->w, pos, nrm, n,
                     n_dev, stride, gd->box[0], gd->box[1], gd->box[2], gd->box[3], gd->box[4],
                     gd->box[5], 0, mix, raw, (uint16_t *)nullptr, (uint16_t *)nullptr,
                     (uint16_t *)nullptr, (uint16_t *)nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "k_guide_fwd");
}

// One training step enqueued on s (no host synchronisation).
int guide_train_launch(ajg_guide *gd, const float *samples, uint32_t n, float lr, hipStream_t s) {
  ajg_ctx *ctx = gd->ctx;
  if (int st = ensure_train(gd, n)) return st;
  AJG_HIP(ctx, hipMemsetAsync(gd->scal, 0, 16, s));
  hipLaunchKernelGGL(k_guide_wsum, dim3(256), dim3(256), 0, s, samples, n, gd->scal);
  const uint32_t fblocks = std::min<uint32_t>(1024, (n + 16 * kGWaves - 1) / (16 * kGWaves));
  // samples rows: pos(3) nrm(3) dir(3) weight(1)
  hipLaunchKernelGGL(k_guide_fwd, dim3(fblocks), dim3(kGThreads), 0, s, (const uint16_t *)gd->wt,
                     (const float *)gd->w, samples, samples + 3, n, (const uint32_t *)nullptr, 10u,
                     gd->box[0], gd->box[1], gd->box[2], gd->box[3], gd->box[4], gd->box[5], 1,
                     (float *)nullptr, gd->out, gd->x0, gd->x1, gd->x2, gd->x3);
  hipLaunchKernelGGL(k_guide_loss, dim3((n + 255) / 256), dim3(256), 0, s, (const float *)gd->out,
                     samples, n, (const float *)gd->scal, gd->dy4, gd->scal + 1);
  hipLaunchKernelGGL(k_guide_bwd_data, dim3(fblocks), dim3(kGThreads), 0, s, (const uint16_t *)gd->wi,
                     (const uint16_t *)gd->dy4, (const uint16_t *)gd->x1, (const uint16_t *)gd->x2,
                     (const uint16_t *)gd->x3, n, gd->dx3, gd->dx2, gd->dx1);
  const uint32_t wb = kGradBlocks;
  hipLaunchKernelGGL((k_guide_bwd_weight<kGHid, kGOut>), dim3(wb), dim3(kGThreads), 0, s,
                     (const uint16_t *)gd->x3, (const uint16_t *)gd->dy4, n, gd->partial, kGW4, kGB4);
  hipLaunchKernelGGL((k_guide_bwd_weight<kGHid, kGHid>), dim3(wb), dim3(kGThreads), 0, s,
                     (const uint16_t *)gd->x2, (const uint16_t *)gd->dx3, n, gd->partial, kGW3, kGB3);
  hipLaunchKernelGGL((k_guide_bwd_weight<kGHid, kGHid>), dim3(wb), dim3(kGThreads), 0, s,
                     (const uint16_t *)gd->x1, (const uint16_t *)gd->dx2, n, gd->partial, kGW2, kGB2);
  hipLaunchKernelGGL((k_guide_bwd_weight<kGIn, kGHid>), dim3(wb), dim3(kGThreads), 0, s,
                     (const uint16_t *)gd->x0, (const uint16_t *)gd->dx1, n, gd->partial, kGW1, kGB1);
  hipLaunchKernelGGL(k_guide_reduce, dim3((kGParams + 255) / 256), dim3(256), 0, s,
                     (const float *)gd->partial, (int)wb, gd->grad);
  if (lr > 0.f) {
    gd->step += 1;
    const float b1c = 1.f - std::pow(0.9f, (float)gd->step);
    const float b2c = 1.f - std::pow(0.999f, (float)gd->step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, gd->w, gd->m,
                       gd->v, (const float *)gd->grad, lr, b1c, b2c);
    hipLaunchKernelGGL(k_guide_cast, dim3((kGWeightCount + 255) / 256), dim3(256), 0, s,
                       (const float *)gd->w, gd->wt, gd->wi);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(ctx, e, "guide train step");
}

int guide_steps(const ajg_guide *gd) { return gd->step; }
const float *guide_loss_dev(const ajg_guide *gd) { return gd->scal + 1; }
const float *guide_box(const ajg_guide *gd) { return gd->box; }

}  // namespace ajg

using namespace ajg;

extern "C" {

int ajg_guide_create(ajg_ctx *ctx, uint32_t seed, const float *scene_box, ajg_guide **out) {
  if (!ctx || !out || !scene_box) return AJG_ERR_INVALID;
  *out = nullptr;
  if (int st = injected(ctx)) return st;
  ajg_guide *gd = new ajg_guide();
  gd->ctx = ctx;
  for (int i = 0; i < 6; ++i) gd->box[i] = scene_box[i];
  for (int i = 0; i < 3; ++i)
    if (!(gd->box[3 + i] > gd->box[i])) gd->box[3 + i] = gd->box[i] + 1.f;
  int st;
  if ((st = dev_alloc(ctx, (void **)&gd->w, 4 * kGParams, "guide w")) ||
      (st = dev_alloc(ctx, (void **)&gd->m, 4 * kGParams, "guide m")) ||
      (st = dev_alloc(ctx, (void **)&gd->v, 4 * kGParams, "guide v")) ||
      (st = dev_alloc(ctx, (void **)&gd->grad, 4 * kGParams, "guide grad")) ||
      (st = dev_alloc(ctx, (void **)&gd->wt, 2 * kGWeightCount, "guide wt")) ||
      (st = dev_alloc(ctx, (void **)&gd->wi, 2 * kGWeightCount, "guide wi")) ||
      (st = dev_alloc(ctx, (void **)&gd->partial, 4ull * kGradBlocks * kGParams, "guide partial")) ||
      (st = dev_alloc(ctx, (void **)&gd->scal, 16, "guide scalars"))) {
    guide_free(gd);
    return st;
  }
  // He-uniform initialisation from a splitmix64 stream; zero biases; small
  // output layer so the initial mixture is close to uniform lobes.
  std::vector<float> h(kGParams, 0.f);
  uint64_t st64 = seed ? seed : 1;
  auto rnd = [&]() {
    st64 += 0x9E3779B97F4A7C15ull;
    uint64_t z = st64;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)((z >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f;
  };
  const int offs[4] = {kGW1, kGW2, kGW3, kGW4};
  const int ins[4] = {kGUsedIn, kGHid, kGHid, kGHid};
  const int rows[4] = {kGIn, kGHid, kGHid, kGHid};
  const int outs[4] = {kGHid, kGHid, kGHid, kGOut};
  for (int l = 0; l < 4; ++l) {
    const float sc = std::sqrt(6.f / (float)ins[l]) * (l == 3 ? 0.1f : 1.f);
    for (int r = 0; r < rows[l]; ++r)
      for (int c = 0; c < outs[l]; ++c) {
        const float val = rnd() * sc;
        if (r < ins[l] && !(l == 3 && c >= kGUsedOut)) h[offs[l] + r * outs[l] + c] = val;
      }
  }
  hipError_t e = hipMemcpy(gd->w, h.data(), 4 * kGParams, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(gd->m, 0, 4 * kGParams);
  if (e == hipSuccess) e = hipMemset(gd->v, 0, 4 * kGParams);
  if (e != hipSuccess) {
    guide_free(gd);
    return hip_fail(ctx, e, "guide init");
  }
  hipLaunchKernelGGL(k_guide_cast, dim3((kGWeightCount + 255) / 256), dim3(256), 0, nullptr,
                     (const float *)gd->w, gd->wt, gd->wi);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    guide_free(gd);
    return hip_fail(ctx, e, "guide cast");
  }
  *out = gd;
  return AJG_OK;
}

int ajg_guide_destroy(ajg_guide *gd) {
  guide_free(gd);
  return AJG_OK;
}

int ajg_guide_get_params(ajg_guide *gd, float *params) {
  if (!gd || !params) return AJG_ERR_INVALID;
  hipError_t e = hipMemcpy(params, gd->w, 4 * kGParams, hipMemcpyDeviceToHost);
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "guide get params");
}

int ajg_guide_set_params(ajg_guide *gd, const float *params) {
  if (!gd || !params) return AJG_ERR_INVALID;
  hipError_t e = hipMemcpy(gd->w, params, 4 * kGParams, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(gd->ctx, e, "guide set params");
  hipLaunchKernelGGL(k_guide_cast, dim3((kGWeightCount + 255) / 256), dim3(256), 0, nullptr,
                     (const float *)gd->w, gd->wt, gd->wi);
  e = hipDeviceSynchronize();
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "guide cast");
}

int ajg_guide_param_count(void) { return kGParams; }

int ajg_guide_forward(ajg_guide *gd, const float *pos, const float *nrm, uint32_t n,
                      float *mixture, float *raw, void *stream) {
  if (!gd || (n && (!pos || !nrm))) return AJG_ERR_INVALID;
  if (int st = injected(gd->ctx)) return st;
  return guide_forward_launch(gd, pos, nrm, n, nullptr, 3, mixture, raw, (hipStream_t)stream);
}

int ajg_guide_train_step(ajg_guide *gd, const float *samples, uint32_t n, float lr,
                         float *loss_out, float *grad_out, void *stream) {
  if (!gd || !samples || n == 0 || !(lr >= 0.f)) return AJG_ERR_INVALID;
  ajg_ctx *ctx = gd->ctx;
  if (int st = injected(ctx)) return st;
  hipStream_t s = (hipStream_t)stream;
  if (int st = guide_train_launch(gd, samples, n, lr, s)) return st;
  if (loss_out) {
    float sc[2];
    AJG_HIP(ctx, hipMemcpyAsync(sc, gd->scal, 8, hipMemcpyDeviceToHost, s));
    AJG_HIP(ctx, hipStreamSynchronize(s));
    *loss_out = sc[1];
  }
  if (grad_out) {
    AJG_HIP(ctx, hipMemcpyAsync(grad_out, gd->grad, 4 * kGParams, hipMemcpyDeviceToHost, s));
    AJG_HIP(ctx, hipStreamSynchronize(s));
  }
  return AJG_OK;
}

}  // extern "C"
