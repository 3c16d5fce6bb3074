// ajg_guide.hip — the path-guiding network on gfx950: BF16 MFMA forward and
// backward kernels with FP32 accumulation, the vMF-mixture loss, Adam.
//
// Every GEMM runs on v_mfma_f32_16x16x16bf16_1k tiles owned by one wave:
//   A (16 x 16, rows m, depth k) lane l holds A[l%16][4(l/16) .. +3]
//   B (16 x 16, depth k, cols n) lane l holds B[4(l/16) .. +3][l%16]
//   D (16 x 16)                  lane l holds D[4(l/16) .. +3][l%16]
// so a D tile of a layer (rows = 16 samples, cols = 16 features) gives each
// lane four consecutive samples of one feature.
//
// Layouts. Weights: FP32 masters W[in][out] plus two BF16 copies, out-major
// WT[out][in] (B operand of the forward pass) and in-major W[in][out] (B
// operand of the backward-data pass), both staged in LDS. Training
// activations and their gradients are stored feature-major, Xt[f][sample]
// (sample count padded to 16): the D-tile of a layer lands as one 8-byte
// store per lane, and the weight-gradient GEMM dW = X^T dY reads both of its
// operands straight from HBM/L2 as 8-byte loads (k = 4 consecutive samples).
//
// A training step is three launches:
//   k_guide_fwd<1>   encode + forward + vMF loss/gradient + backward data,
//                    one 16-sample tile per wave, everything in LDS/registers
//                    except the feature-major activations it writes;
//   k_guide_wgrad    dW and db of all four layers (blockIdx.y = layer), one
//                    partial sum per block;
//   k_guide_update   fixed-order reduction of the partials, normalisation by
//                    the batch's total sample weight, Adam on the FP32
//                    masters, refresh of both BF16 copies.
// All reductions run in a fixed order, so a step is deterministic.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ajg_guide.h"
#include "ajg_internal.h"

namespace ajg {
static constexpr int kGWeightCount = kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut;  // 18432
static constexpr int kFwdMaxBlocks = 1024;
static constexpr int kWgBlocks = 128;   // weight-gradient blocks per layer
static constexpr int kWgWaves = 4;
}  // namespace ajg

struct ajg_guide {
  ajg_ctx *ctx = nullptr;
  float box[6] = {0, 0, 0, 1, 1, 1};
  float *w = nullptr;                // FP32 masters (kGParams)
  float *m = nullptr, *v = nullptr;  // Adam moments
  uint16_t *wt = nullptr;            // BF16 out-major copies of the 4 weight matrices
  uint16_t *wi = nullptr;            // BF16 in-major copies
  float *grad = nullptr;             // normalised gradient of the last step
  float *partial = nullptr;          // kWgBlocks x kGParams
  float *scal = nullptr;             // 2 x kFwdMaxBlocks: per forward block (sum w, sum -w log q)
  float *lossbuf = nullptr;          // [0] sum of weights, [1] loss of the last step
  int step = 0;
  // training-record pool filled by guided renders (kept across renders):
  // rows of 10 floats; rec_ctr[0] = rows held, rec_ctr[2..3] = u64 total
  float *rec = nullptr;
  uint32_t rec_cap = 0;
  uint32_t *rec_ctr = nullptr;
  // feature-major training activations, npad = capacity (multiple of 16)
  uint32_t cap = 0;
  uint16_t *x0t = nullptr, *x1t = nullptr, *x2t = nullptr, *x3t = nullptr;
  uint16_t *dy4t = nullptr, *dx3t = nullptr, *dx2t = nullptr, *dx1t = nullptr;
};

namespace ajg {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// Two k-steps of 16, accumulated in order. (One v_mfma_f32_16x16x32_bf16
// over both steps gave results that changed from run to run in the
// training kernel as this toolchain schedules it: identical steps on
// identical inputs differed by up to 12% of the largest gradient entry.
// With two 16x16x16 MFMAs a step is bitwise reproducible.)
__device__ __forceinline__ f32x4 mfma32(s16x4 a0, s16x4 a1, s16x4 b0, s16x4 b1, f32x4 c) {
  return mfma16(a1, b1, mfma16(a0, b0, c));
}
// four consecutive BF16 values; every caller's address is 8-byte aligned
__device__ __forceinline__ s16x4 ld4(const uint16_t *p) {
  return *reinterpret_cast<const s16x4 *>(p);
}
__device__ __forceinline__ void st4(uint16_t *p, s16x4 v) { *reinterpret_cast<s16x4 *>(p) = v; }
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// ---- output-layer column order. The output layer's D tile gives lane
// (c, g) the outputs 16 mt + 4 g + i (mt = 0..2, i = 0..3) of sample c;
// the BF16 copies of W4 store the columns permuted so that these 12 slots
// q = 4 mt + i hold lobes 2g (q 0..4) and 2g+1 (q 5..9) of the logical
// output vector, plus two unused outputs. Masters stay in logical order.
__host__ __device__ __forceinline__ int out_logical(int p) {
  const int mt = p >> 4, g = (p & 15) >> 2, i = p & 3, q = 4 * mt + i;
  return q < 10 ? 10 * g + q : kGUsedOut + 2 * g + (q - 10);
}
__host__ __device__ __forceinline__ int out_physical(int L) {
  int g, q;
  if (L < kGUsedOut) {
    g = L / 10;
    q = L % 10;
  } else {
    g = (L - kGUsedOut) >> 1;
    q = 10 + ((L - kGUsedOut) & 1);
  }
  return 16 * (q >> 2) + 4 * g + (q & 3);
}

// ---- LDS images of the weights, rows padded by 4 BF16 so that the 16 rows
// one MFMA fragment reads fall on distinct bank pairs
static constexpr int kLd1 = kGIn + 4;   // 116
static constexpr int kLdH = kGHid + 4;  // 68
static constexpr int kLdO = kGOut + 4;  // 52

template <int MODE>
struct NetCfg {
  static constexpr int kWaves = 4;
  static constexpr int kNT = MODE ? 1 : 4;  // 16-sample tiles per wave iteration
};

template <int MODE>
struct alignas(16) NetLds {
  // A operands of the forward pass: out-major rows
  uint16_t wt1[kGHid * kLd1], wt2[kGHid * kLdH], wt3[kGHid * kLdH], wt4[kGOut * kLdH];
  // A operands of the backward-data pass: in-major rows
  uint16_t wi2[MODE ? kGHid * kLdH : 8], wi3[MODE ? kGHid * kLdH : 8],
      wi4[MODE ? kGHid * kLdO : 8];
  float b1[kGHid], b2[kGHid], b3[kGHid], b4[kGOut];  // b4 in physical order
  float red[2 * NetCfg<MODE>::kWaves];
};

struct FwdArgs {
  const uint16_t *wt, *wi;
  const float *w;
  const float *pos, *nrm;  // rows of `stride` floats
  uint32_t stride, n;
  const uint32_t *n_dev;   // optional device-side count (inference)
  float box[6];
  float *mix, *raw;        // inference outputs (optional)
  _Float16 *mixh;          // inference: FP16 mixture for the renderer (optional)
  const float *samples;    // training rows: pos nrm dir weight
  uint32_t npad;
  uint16_t *x0t, *x1t, *x2t, *x3t, *dy4t, *dx3t, *dx2t, *dx1t;
  float *scal;
};

// copy rows x K BF16 (K contiguous) into LDS rows of stride ld
template <int T>
__device__ __forceinline__ void copy_rows(uint16_t *dst, int ld, const uint16_t *src, int rows, int K) {
  const int q = K / 4;
  for (int idx = threadIdx.x; idx < rows * q; idx += T) {
    const int row = idx / q, c4 = idx - row * q;
    st4(dst + row * ld + 4 * c4, ld4(src + row * K + 4 * c4));
  }
}

template <int MODE>
__device__ void load_net(NetLds<MODE> &s, const FwdArgs &a) {
  constexpr int T = 64 * NetCfg<MODE>::kWaves;
  const int o2 = kGIn * kGHid, o3 = o2 + kGHid * kGHid, o4 = o3 + kGHid * kGHid;
  copy_rows<T>(s.wt1, kLd1, a.wt, kGHid, kGIn);
  copy_rows<T>(s.wt2, kLdH, a.wt + o2, kGHid, kGHid);
  copy_rows<T>(s.wt3, kLdH, a.wt + o3, kGHid, kGHid);
  copy_rows<T>(s.wt4, kLdH, a.wt + o4, kGOut, kGHid);
  if (MODE) {
    copy_rows<T>(s.wi2, kLdH, a.wi + o2, kGHid, kGHid);
    copy_rows<T>(s.wi3, kLdH, a.wi + o3, kGHid, kGHid);
    copy_rows<T>(s.wi4, kLdO, a.wi + o4, kGHid, kGOut);
  }
  for (int i = threadIdx.x; i < kGHid; i += T) {
    s.b1[i] = a.w[kGB1 + i];
    s.b2[i] = a.w[kGB2 + i];
    s.b3[i] = a.w[kGB3 + i];
  }
  for (int i = threadIdx.x; i < kGOut; i += T) s.b4[i] = a.w[kGB4 + out_logical(i)];
  __syncthreads();
}

// B fragment of input k-step j for one sample (the lane's features
// 16j + 4g .. +3; see the encoding in ajg_guide.h). sin/cos(pi m n) come
// from the hardware sine/cosine, whose argument is in revolutions
// (m n / 2, |.| <= 8 here).
__device__ __forceinline__ s16x4 encode_frag(int j, int g, f3 ph, f3 nn, bool live) {
  s16x4 r = {0, 0, 0, 0};
  if (!live) return r;
  if (j < 6) {
    const int c = j >> 1;
    const float nc = c == 0 ? nn.x : (c == 1 ? nn.y : nn.z);
    const int h0 = 8 * (j & 1) + 2 * g;
    const float r0 = 0.5f * (float)(h0 + 1) * nc, r1 = 0.5f * (float)(h0 + 2) * nc;
    const float s0 = __builtin_amdgcn_sinf(r0), c0 = __builtin_amdgcn_cosf(r0);
    const float s1 = __builtin_amdgcn_sinf(r1), c1 = __builtin_amdgcn_cosf(r1);
    r[0] = (short)f2bf(s0);
    r[1] = (short)f2bf(c0);
    r[2] = (short)f2bf(s1);
    r[3] = (short)f2bf(c1);
  } else if (g == 0) {
    r[0] = (short)f2bf(ph.x);
    r[1] = (short)f2bf(ph.y);
    r[2] = (short)f2bf(ph.z);
  }
  return r;
}

// D[m][n] = sum_k A[m][k] B[k][n] over KT 16-wide k-steps; A = weight rows
// from LDS (row stride LD), B = fragments in registers; MT row tiles, NT
// column tiles; one 16x16x16 MFMA per k-step (pairs through mfma32).
template <int KT, int MT, int NT, int LD>
__device__ __forceinline__ void net_layer(const uint16_t *Wl, const s16x4 (&bf)[KT][NT],
                                          f32x4 (&acc)[MT][NT]) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt + 1 < KT; kt += 2)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const uint16_t *row = Wl + (mt * 16 + r) * LD + kt * 16 + 4 * g;
      const s16x4 a0 = ld4(row), a1 = ld4(row + 16);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = mfma32(a0, a1, bf[kt][nt], bf[kt + 1][nt], acc[mt][nt]);
    }
  if (KT & 1)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const s16x4 av = ld4(Wl + (mt * 16 + r) * LD + (KT - 1) * 16 + 4 * g);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma16(av, bf[KT - 1][nt], acc[mt][nt]);
    }
}

// relu(D + bias) as BF16 B fragments of the next layer
template <int MT, int NT>
__device__ __forceinline__ void activate(const f32x4 (&acc)[MT][NT], const float *bias,
                                         s16x4 (&out)[MT][NT]) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float y = acc[mt][nt][i] + bias[mt * 16 + 4 * g + i];
        out[mt][nt][i] = (short)f2bf(y > 0.f ? y : 0.f);
      }
}

// feature-major store of one sample's fragments: Xt[16 kt + 4 g + i][smp]
template <int KT>
__device__ __forceinline__ void store_t(uint16_t *Xt, uint32_t npad, uint32_t smp,
                                        const s16x4 (&f)[KT][1]) {
  const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int i = 0; i < 4; ++i) Xt[(size_t)(kt * 16 + 4 * g + i) * npad + smp] = (uint16_t)f[kt][0][i];
}

// dX = D masked by [X > 0] as BF16 fragments
template <int MT>
__device__ __forceinline__ void mask_grad(const f32x4 (&acc)[MT][1], const s16x4 (&x)[MT][1],
                                          s16x4 (&out)[MT][1]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      out[mt][0][i] = (short)f2bf(bf2f((uint16_t)x[mt][0][i]) > 0.f ? acc[mt][0][i] : 0.f);
}

// Forward pass, everything in registers: the weights are the A operand, the
// activations the B operand, so the D tile of one layer (4 consecutive
// features of one sample per lane) is the B fragment of the next.
// MODE 0: inference (decoded mixture, optional raw outputs), 4 tiles of 16
// samples per wave iteration. MODE 1: training (forward, vMF loss gradient,
// backward data), one tile per iteration, feature-major activations out.
// Inference runs two waves per SIMD (256 registers, a few spilled): 1M
// points in 0.125 ms against 0.158 ms at the compiler's own budget (one
// wave, 234 + 72 registers); fewer tiles per wave or three waves spill far
// more and are 2x slower. The training pass keeps one wave (two: 13%
// slower).
template <int MODE>
__global__ void __launch_bounds__(64 * NetCfg<MODE>::kWaves)
__attribute__((amdgpu_waves_per_eu(MODE ? 1 : 2)))
k_guide_fwd(FwdArgs a) {
  constexpr int W = NetCfg<MODE>::kWaves, NT = NetCfg<MODE>::kNT;
  __shared__ NetLds<MODE> s;
  uint32_t n = a.n;
  if (!MODE && a.n_dev) n = min(n, *a.n_dev);  // queue size written on the device
  const uint32_t limit = MODE ? a.npad : n;
  if (limit == 0) return;
  load_net<MODE>(s, a);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  float sum_w = 0.f, sum_l = 0.f;
  for (uint32_t tile = blockIdx.x * W + wv; tile * 16 * NT < limit; tile += gridDim.x * W) {
    const uint32_t base = tile * 16 * NT;
    // ---- encoding
    s16x4 x0[7][NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const uint32_t smp = base + 16 * nt + c;
      const bool live = smp < n;
      f3 ph = mk3(0.f, 0.f, 0.f), nn = mk3(0.f, 0.f, 1.f);
      if (live) {
        const float *pp = a.pos + (size_t)smp * a.stride;
        const float *np = a.nrm + (size_t)smp * a.stride;
        ph = mk3(2.f * (pp[0] - a.box[0]) / (a.box[3] - a.box[0]) - 1.f,
                 2.f * (pp[1] - a.box[1]) / (a.box[4] - a.box[1]) - 1.f,
                 2.f * (pp[2] - a.box[2]) / (a.box[5] - a.box[2]) - 1.f);
        nn = mk3(np[0], np[1], np[2]);
      }
#pragma unroll
      for (int j = 0; j < 7; ++j) x0[j][nt] = encode_frag(j, g, ph, nn, live);
    }
    // ---- forward
    f32x4 acc[4][NT];
    s16x4 x1[4][NT], x2[4][NT], x3[4][NT];
    net_layer<7, 4, NT, kLd1>(s.wt1, x0, acc);
    activate<4, NT>(acc, s.b1, x1);
    net_layer<4, 4, NT, kLdH>(s.wt2, x1, acc);
    activate<4, NT>(acc, s.b2, x2);
    net_layer<4, 4, NT, kLdH>(s.wt3, x2, acc);
    activate<4, NT>(acc, s.b3, x3);
    f32x4 o4[3][NT];
    net_layer<4, 3, NT, kLdH>(s.wt4, x3, o4);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const uint32_t smp = base + 16 * nt + c;
      const bool live = smp < n;
      // the lane's 12 outputs: slots q = 4 mt + i (lobes 2g, 2g+1, 2 unused)
      float ov[12];
#pragma unroll
      for (int mt = 0; mt < 3; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) ov[4 * mt + i] = o4[mt][nt][i] + s.b4[16 * mt + 4 * g + i];
      // ---- decode: lanes c, c+16, c+32, c+48 hold the 8 lobes of sample c
      f3 mu[2];
      float kap[2], len[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float vx = ov[5 * q], vy = ov[5 * q + 1], vz = ov[5 * q + 2];
        len[q] = fmaxf(sqrtf(vx * vx + vy * vy + vz * vz), 1e-6f);
        const float il = __builtin_amdgcn_rcpf(len[q]);
        mu[q] = mk3(vx * il, vy * il, vz * il);
        kap[q] = expf(fminf(fmaxf(ov[5 * q + 3], kKappaLo), kKappaHi));
      }
      float mx = fmaxf(ov[4], ov[9]);
      mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
      mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
      const float e0 = expf(ov[4] - mx), e1 = expf(ov[9] - mx);
      float se = e0 + e1;
      se += __shfl_xor(se, 16, 64);
      se += __shfl_xor(se, 32, 64);
      const float ise = __builtin_amdgcn_rcpf(se);
      const float pi[2] = {e0 * ise, e1 * ise};
      if (!MODE) {
        if (a.raw && live)
#pragma unroll
          for (int q = 0; q < 12; ++q)
            a.raw[(size_t)smp * kGOut + (q < 10 ? 10 * g + q : kGUsedOut + 2 * g + q - 10)] = ov[q];
        if (a.mixh && live) {
          _Float16 *dst = a.mixh + (size_t)smp * kGUsedOut + 10 * g;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            dst[5 * q] = (_Float16)mu[q].x;
            dst[5 * q + 1] = (_Float16)mu[q].y;
            dst[5 * q + 2] = (_Float16)mu[q].z;
            dst[5 * q + 3] = (_Float16)kap[q];
            dst[5 * q + 4] = (_Float16)pi[q];
          }
        }
        if (a.mix && live) {
          float *dst = a.mix + (size_t)smp * kGUsedOut + 10 * g;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            dst[5 * q] = mu[q].x;
            dst[5 * q + 1] = mu[q].y;
            dst[5 * q + 2] = mu[q].z;
            dst[5 * q + 3] = kap[q];
            dst[5 * q + 4] = pi[q];
          }
        }
        continue;
      }
      // ---- training: loss -sum w log q(d) and its gradient w.r.t. the outputs
      f3 d = mk3(0.f, 0.f, 1.f);
      float wgt = 0.f;
      if (live) {
        const float *sr = a.samples + (size_t)smp * 10;
        d = mk3(sr[6], sr[7], sr[8]);
        wgt = sr[9];
      }
      float pq[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        VmfLobe l;
        l.mu = mu[q];
        l.kappa = kap[q];
        l.w = 1.f;
        pq[q] = pi[q] * vmf_pdf(l, d);
      }
      float qs = pq[0] + pq[1];
      qs += __shfl_xor(qs, 16, 64);
      qs += __shfl_xor(qs, 32, 64);
      qs = fmaxf(qs, 1e-30f);
      if (g == 0) {
        sum_w += wgt;
        sum_l -= wgt * logf(qs);
      }
      float gr[12];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float rj = pq[q] * __builtin_amdgcn_rcpf(qs);  // responsibility
        const float k = kap[q];
        const float md = dot(mu[q], d);
        // d log q / d logit_j = r_j - pi_j
        gr[5 * q + 4] = -wgt * (rj - pi[q]);
        // d log q_j / d kappa = 1/kappa - 2/(e^{2 kappa} - 1) + (mu.d - 1); d kappa / d raw = kappa
        const float dk =
            __builtin_amdgcn_rcpf(k) - 2.f * __builtin_amdgcn_rcpf(expm1f(2.f * k)) + (md - 1.f);
        const float kr = ov[5 * q + 3];
        gr[5 * q + 3] = (kr > kKappaLo && kr < kKappaHi) ? -wgt * rj * k * dk : 0.f;
        // d log q_j / d mu_raw = kappa (d - (mu.d) mu) / |mu_raw|
        const float sc = -wgt * rj * k * __builtin_amdgcn_rcpf(len[q]);
        gr[5 * q + 0] = sc * (d.x - md * mu[q].x);
        gr[5 * q + 1] = sc * (d.y - md * mu[q].y);
        gr[5 * q + 2] = sc * (d.z - md * mu[q].z);
      }
      gr[10] = 0.f;
      gr[11] = 0.f;
      s16x4 dy[3][1];
#pragma unroll
      for (int mt = 0; mt < 3; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) dy[mt][0][i] = (short)f2bf(gr[4 * mt + i]);
      // ---- feature-major copies for the weight gradient
      {
        s16x4 t0[7][1], t1[4][1], t2[4][1], t3[4][1];
#pragma unroll
        for (int j = 0; j < 7; ++j) t0[j][0] = x0[j][0];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          t1[j][0] = x1[j][0];
          t2[j][0] = x2[j][0];
          t3[j][0] = x3[j][0];
        }
        store_t<7>(a.x0t, a.npad, smp, t0);
        store_t<4>(a.x1t, a.npad, smp, t1);
        store_t<4>(a.x2t, a.npad, smp, t2);
        store_t<4>(a.x3t, a.npad, smp, t3);
        store_t<3>(a.dy4t, a.npad, smp, dy);
      }
      // ---- backward data: dX3 = (W4 dY4) * [X3 > 0], dX2, dX1
      f32x4 bacc[4][1];
      s16x4 d3[4][1], d2[4][1], d1[4][1], xm[4][1];
      net_layer<3, 4, 1, kLdO>(s.wi4, dy, bacc);
#pragma unroll
      for (int j = 0; j < 4; ++j) xm[j][0] = x3[j][0];
      mask_grad<4>(bacc, xm, d3);
      store_t<4>(a.dx3t, a.npad, smp, d3);
      net_layer<4, 4, 1, kLdH>(s.wi3, d3, bacc);
#pragma unroll
      for (int j = 0; j < 4; ++j) xm[j][0] = x2[j][0];
      mask_grad<4>(bacc, xm, d2);
      store_t<4>(a.dx2t, a.npad, smp, d2);
      net_layer<4, 4, 1, kLdH>(s.wi2, d2, bacc);
#pragma unroll
      for (int j = 0; j < 4; ++j) xm[j][0] = x1[j][0];
      mask_grad<4>(bacc, xm, d1);
      store_t<4>(a.dx1t, a.npad, smp, d1);
    }
  }
  if (MODE) {
    // per-block (sum w, sum -w log q), waves combined in a fixed order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sum_w += __shfl_xor(sum_w, off, 64);
      sum_l += __shfl_xor(sum_l, off, 64);
    }
    if (lane == 0) {
      s.red[2 * wv] = sum_w;
      s.red[2 * wv + 1] = sum_l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float a0 = 0.f, a1 = 0.f;
      for (int i = 0; i < W; ++i) {
        a0 += s.red[2 * i];
        a1 += s.red[2 * i + 1];
      }
      a.scal[2 * blockIdx.x] = a0;
      a.scal[2 * blockIdx.x + 1] = a1;
    }
  }
}

// dW[in][out] = sum_s Xt[in][s] dYt[out][s] and db = sum_s dYt[out][s] over
// this block's sample range; A = X^T and B = dY load straight from memory.
template <int KIN, int NOUT, bool PERM>
__device__ __forceinline__ void wgrad_layer(const uint16_t *Xt, const uint16_t *Dt, uint32_t npad,
                                            uint32_t kbeg, uint32_t kend, float *red, float *dst,
                                            int w_off, int b_off) {
  constexpr int TM = KIN / 16, TN = NOUT / 16;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  f32x4 acc[TM * TN];
#pragma unroll
  for (int t = 0; t < TM * TN; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bs[TN];
#pragma unroll
  for (int t = 0; t < TN; ++t) bs[t] = 0.f;
  for (uint32_t s0 = kbeg + 16 * wv; s0 < kend; s0 += 16 * kWgWaves) {
    s16x4 av[TM], bv[TN];
#pragma unroll
    for (int tm = 0; tm < TM; ++tm) av[tm] = ld4(Xt + (size_t)(tm * 16 + r) * npad + s0 + 4 * g);
#pragma unroll
    for (int tn = 0; tn < TN; ++tn) {
      bv[tn] = ld4(Dt + (size_t)(tn * 16 + r) * npad + s0 + 4 * g);
      bs[tn] += (bf2f((uint16_t)bv[tn][0]) + bf2f((uint16_t)bv[tn][1])) +
                (bf2f((uint16_t)bv[tn][2]) + bf2f((uint16_t)bv[tn][3]));
    }
#pragma unroll
    for (int tm = 0; tm < TM; ++tm)
#pragma unroll
      for (int tn = 0; tn < TN; ++tn) acc[tm * TN + tn] = mfma16(av[tm], bv[tn], acc[tm * TN + tn]);
  }
#pragma unroll
  for (int tn = 0; tn < TN; ++tn) {
    bs[tn] += __shfl_xor(bs[tn], 16, 64);
    bs[tn] += __shfl_xor(bs[tn], 32, 64);
  }
  for (int i = threadIdx.x; i < KIN * NOUT + NOUT; i += 64 * kWgWaves) red[i] = 0.f;
  __syncthreads();
  for (int w = 0; w < kWgWaves; ++w) {
    if (wv == w) {
#pragma unroll
      for (int tm = 0; tm < TM; ++tm)
#pragma unroll
        for (int tn = 0; tn < TN; ++tn)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            red[(tm * 16 + 4 * g + i) * NOUT + tn * 16 + r] += acc[tm * TN + tn][i];
      if (g == 0)
#pragma unroll
        for (int tn = 0; tn < TN; ++tn) red[KIN * NOUT + tn * 16 + r] += bs[tn];
    }
    __syncthreads();
  }
  // PERM: the output layer's rows of dYt are in physical column order
  for (int i = threadIdx.x; i < KIN * NOUT; i += 64 * kWgWaves) {
    const int in = i / NOUT, o = i - in * NOUT;
    dst[w_off + in * NOUT + (PERM ? out_logical(o) : o)] = red[i];
  }
  for (int i = threadIdx.x; i < NOUT; i += 64 * kWgWaves)
    dst[b_off + (PERM ? out_logical(i) : i)] = red[KIN * NOUT + i];
}

struct WgArgs {
  const uint16_t *x0t, *x1t, *x2t, *x3t, *dy4t, *dx3t, *dx2t, *dx1t;
  uint32_t npad, chunk;
  float *partial;
};

__global__ void __launch_bounds__(64 * kWgWaves)
k_guide_wgrad(WgArgs a) {
  __shared__ float red[kGIn * kGHid + kGHid];
  const uint32_t kbeg = blockIdx.x * a.chunk;
  const uint32_t kend = min(a.npad, kbeg + a.chunk);
  float *dst = a.partial + (size_t)blockIdx.x * kGParams;
  switch (blockIdx.y) {
    case 0:
      wgrad_layer<kGIn, kGHid, false>(a.x0t, a.dx1t, a.npad, kbeg, kend, red, dst, kGW1, kGB1);
      break;
    case 1:
      wgrad_layer<kGHid, kGHid, false>(a.x1t, a.dx2t, a.npad, kbeg, kend, red, dst, kGW2, kGB2);
      break;
    case 2:
      wgrad_layer<kGHid, kGHid, false>(a.x2t, a.dx3t, a.npad, kbeg, kend, red, dst, kGW3, kGB3);
      break;
    default:
      wgrad_layer<kGHid, kGOut, true>(a.x3t, a.dy4t, a.npad, kbeg, kend, red, dst, kGW4, kGB4);
      break;
  }
}

// Weight-copy indices of master parameter i (false for a bias): in-major
// and out-major index within the BF16 copies (W4 columns in physical order).
__device__ __forceinline__ bool weight_copy_index(int i, int *wi_idx, int *wt_idx) {
  const int offs[4] = {kGW1, kGW2, kGW3, kGW4};
  const int ins[4] = {kGIn, kGHid, kGHid, kGHid};
  const int outs[4] = {kGHid, kGHid, kGHid, kGOut};
  int base = 0;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int cnt = ins[l] * outs[l];
    if (i >= offs[l] && i < offs[l] + cnt) {
      const int k = i - offs[l], in = k / outs[l], lo = k % outs[l];
      const int out = l == 3 ? out_physical(lo) : lo;
      *wi_idx = base + in * outs[l] + out;
      *wt_idx = base + out * ins[l] + in;
      return true;
    }
    base += cnt;
  }
  return false;
}

// Reduce the per-block partials in block order, normalise by the batch's
// total weight, Adam (beta 0.9 / 0.999, eps 1e-8) on the masters, refresh
// the BF16 copies.
__global__ void __launch_bounds__(256)
k_guide_update(const float *partial, int nb, const float *scal, int nfb, float *grad, float *w,
               float *m, float *v, uint16_t *wt, uint16_t *wi, float lr, float b1c, float b2c,
               float *lossbuf) {
  __shared__ float rw[256], rl[256];
  float aw = 0.f, al = 0.f;
  for (int b = threadIdx.x; b < nfb; b += 256) {
    aw += scal[2 * b];
    al += scal[2 * b + 1];
  }
  rw[threadIdx.x] = aw;
  rl[threadIdx.x] = al;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      rw[threadIdx.x] += rw[threadIdx.x + st];
      rl[threadIdx.x] += rl[threadIdx.x + st];
    }
    __syncthreads();
  }
  const float S = rw[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    lossbuf[0] = S;
    lossbuf[1] = S > 0.f ? rl[0] / S : 0.f;
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kGParams) return;
  // eight independent sums (partials b = 8 k + j go to sum j), combined in
  // a fixed order: eight loads in flight per thread instead of one chain
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int pb = 0;
  for (; pb + 8 <= nb; pb += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += partial[(size_t)(pb + j) * kGParams + i];
  for (; pb < nb; ++pb) acc[pb & 7] += partial[(size_t)pb * kGParams + i];
  const float gs = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  const float gi = S > 0.f ? gs / S : 0.f;
  grad[i] = gi;
  if (!(lr > 0.f) || !(S > 0.f)) return;
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wn = w[i] - lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
  w[i] = wn;
  int a, b;
  if (weight_copy_index(i, &a, &b)) {
    const uint16_t h = f2bf(wn);
    wi[a] = h;
    wt[b] = h;
  }
}

// BF16 copies of all weights from the masters (creation, set_params).
__global__ void __launch_bounds__(256)
k_guide_cast(const float *w, uint16_t *wt, uint16_t *wi) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kGParams) return;
  int a, b;
  if (weight_copy_index(i, &a, &b)) {
    const uint16_t h = f2bf(w[i]);
    wi[a] = h;
    wt[b] = h;
  }
}

static int ensure_train(ajg_guide *gd, uint32_t npad) {
  if (gd->cap >= npad) return AJG_OK;
  uint16_t **ps[] = {&gd->x0t, &gd->x1t, &gd->x2t, &gd->x3t, &gd->dy4t, &gd->dx3t, &gd->dx2t, &gd->dx1t};
  const int rows[] = {kGIn, kGHid, kGHid, kGHid, kGOut, kGHid, kGHid, kGHid};
  for (uint16_t **p : ps) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  gd->cap = 0;
  for (int k = 0; k < 8; ++k)
    if (int st = dev_alloc(gd->ctx, (void **)ps[k], 2ull * rows[k] * npad, "guide activations"))
      return st;
  gd->cap = npad;
  return AJG_OK;
}

static void guide_free(ajg_guide *gd) {
  if (!gd) return;
  void *ps[] = {gd->w, gd->m, gd->v, gd->wt, gd->wi, gd->grad, gd->partial, gd->scal, gd->lossbuf,
                gd->rec, gd->rec_ctr,
                gd->x0t, gd->x1t, gd->x2t, gd->x3t, gd->dy4t, gd->dx3t, gd->dx2t, gd->dx1t};
  for (void *p : ps) (void)hipFree(p);
  delete gd;
}

static FwdArgs fwd_args(const ajg_guide *gd) {
  FwdArgs a;
  memset(&a, 0, sizeof(a));
  a.wt = gd->wt;
  a.wi = gd->wi;
  a.w = gd->w;
  for (int i = 0; i < 6; ++i) a.box[i] = gd->box[i];
  return a;
}

int guide_forward_launch(ajg_guide *gd, const float *pos, const float *nrm, uint32_t n,
                         const uint32_t *n_dev, uint32_t stride, float *mix, _Float16 *mixh,
                         float *raw, hipStream_t s) {
  if (n == 0) return AJG_OK;
  FwdArgs a = fwd_args(gd);
  a.pos = pos;
  a.nrm = nrm;
  a.stride = stride;
  a.n = n;
  a.n_dev = n_dev;
  a.mix = mix;
  a.mixh = mixh;
  a.raw = raw;
  constexpr int W = NetCfg<0>::kWaves, NT = NetCfg<0>::kNT;
  const uint32_t tiles = (n + 16 * NT - 1) / (16 * NT);
  const uint32_t blocks = std::min<uint32_t>(kFwdMaxBlocks, (tiles + W - 1) / W);
  hipLaunchKernelGGL(k_guide_fwd<0>, dim3(blocks), dim3(64 * W), 0, s, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "k_guide_fwd");
}

// Per-step work buffers of a training step: the feature-major activations
// and their gradients (rows x npad BF16 each, columns = records padded to
// 16), the weight-gradient partials of every block and the forward blocks'
// weight and loss sums. Each step writes every element it later reads, so
// none needs clearing between steps.
struct TrainBufs {
  uint16_t *x0t, *x1t, *x2t, *x3t, *dy4t, *dx3t, *dx2t, *dx1t;
  float *partial, *scal;
};

// Feature rows of the eight activation arrays, in TrainBufs order.
static constexpr int kActRows[8] = {kGIn, kGHid, kGHid, kGHid, kGOut, kGHid, kGHid, kGHid};
static constexpr size_t kActRowsTotal = kGIn + 6 * kGHid + kGOut;  // 544

// One training step enqueued on s (no host synchronisation).
static int train_launch(ajg_guide *gd, const TrainBufs &tb, const float *samples, uint32_t n,
                        uint32_t npad, float lr, hipStream_t s) {
  FwdArgs a = fwd_args(gd);
  a.samples = samples;
  a.pos = samples;
  a.nrm = samples + 3;
  a.stride = 10;
  a.n = n;
  a.npad = npad;
  a.x0t = tb.x0t;
  a.x1t = tb.x1t;
  a.x2t = tb.x2t;
  a.x3t = tb.x3t;
  a.dy4t = tb.dy4t;
  a.dx3t = tb.dx3t;
  a.dx2t = tb.dx2t;
  a.dx1t = tb.dx1t;
  a.scal = tb.scal;
  constexpr int W = NetCfg<1>::kWaves;
  const uint32_t tiles = npad / 16;
  const uint32_t fblocks = std::min<uint32_t>(kFwdMaxBlocks, (tiles + W - 1) / W);
  hipLaunchKernelGGL(k_guide_fwd<1>, dim3(fblocks), dim3(64 * W), 0, s, a);
  WgArgs wa;
  wa.x0t = tb.x0t;
  wa.x1t = tb.x1t;
  wa.x2t = tb.x2t;
  wa.x3t = tb.x3t;
  wa.dy4t = tb.dy4t;
  wa.dx3t = tb.dx3t;
  wa.dx2t = tb.dx2t;
  wa.dx1t = tb.dx1t;
  wa.npad = npad;
  wa.chunk = (tiles + kWgBlocks - 1) / kWgBlocks * 16;
  wa.partial = tb.partial;
  hipLaunchKernelGGL(k_guide_wgrad, dim3(kWgBlocks, 4), dim3(64 * kWgWaves), 0, s, wa);
  float b1c = 1.f, b2c = 1.f;
  if (lr > 0.f) {
    gd->step += 1;
    b1c = 1.f - std::pow(0.9f, (float)gd->step);
    b2c = 1.f - std::pow(0.999f, (float)gd->step);
  }
  hipLaunchKernelGGL(k_guide_update, dim3((kGParams + 255) / 256), dim3(256), 0, s,
                     (const float *)tb.partial, kWgBlocks, (const float *)tb.scal, (int)fblocks,
                     gd->grad, gd->w, gd->m, gd->v, gd->wt, gd->wi, lr, b1c, b2c, gd->lossbuf);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "guide train step");
}

// The guide's own buffers (activations grown on demand: may allocate).
int guide_train_launch(ajg_guide *gd, const float *samples, uint32_t n, float lr, hipStream_t s) {
  const uint32_t npad = (n + 15) / 16 * 16;
  if (int st = ensure_train(gd, npad)) return st;
  const TrainBufs tb = {gd->x0t, gd->x1t, gd->x2t, gd->x3t, gd->dy4t, gd->dx3t, gd->dx2t, gd->dx1t,
                        gd->partial, gd->scal};
  return train_launch(gd, tb, samples, n, npad, lr, s);
}

// ---- allocation-free training step
// Its work buffers live in `scratch`, one copy per device for the process's
// lifetime, sized for kNoAllocMaxRecords records: gradient partials
// (9,560,064 B), forward block sums (8,192 B), activations (1,088 B per
// record). Host code never passes `scratch` itself (in host code the name is
// a stand-in, not a device address): guide_scratch_setup records each
// device's address of it from hipGetSymbolAddress.
static constexpr uint32_t kNoAllocMaxRecords = 1u << 15;
static constexpr size_t kScratchPartialBytes = 4ull * kWgBlocks * kGParams;
static constexpr size_t kScratchScalBytes = 8ull * kFwdMaxBlocks;
static constexpr size_t kScratchBytes =
    kScratchPartialBytes + kScratchScalBytes + 2ull * kActRowsTotal * kNoAllocMaxRecords;
static_assert(kScratchPartialBytes % 256 == 0 && kScratchScalBytes % 256 == 0,
              "scratch sub-buffers start on 256-byte boundaries");
static_assert(kScratchBytes == 45219840ull, "scratch size as documented");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxScratchDevices];

// Called by ajg_ctx_create with the context's device current.
int guide_scratch_setup(ajg_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
    return fail(ctx, AJG_ERR_INVALID, "device index above the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return AJG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[ctx->device] = (unsigned char *)p;
  return AJG_OK;
}

// Carves the step's buffers for npad records out of the device's scratch.
static TrainBufs scratch_bufs(unsigned char *base, uint32_t npad) {
  TrainBufs tb;
  tb.partial = (float *)base;
  tb.scal = (float *)(base + kScratchPartialBytes);
  unsigned char *act = base + kScratchPartialBytes + kScratchScalBytes;
  uint16_t **dst[8] = {&tb.x0t, &tb.x1t, &tb.x2t, &tb.x3t, &tb.dy4t, &tb.dx3t, &tb.dx2t, &tb.dx1t};
  for (int k = 0; k < 8; ++k) {
    *dst[k] = (uint16_t *)act;
    act += 2ull * kActRows[k] * npad;  // a multiple of 256 bytes (npad % 16 == 0)
  }
  return tb;
}

int guide_train_launch_noalloc(ajg_guide *gd, const float *samples, uint32_t n, float lr,
                               hipStream_t s) {
  if (n > kNoAllocMaxRecords)
    return fail(gd->ctx, AJG_ERR_INVALID,
                "guide train step (no allocation): at most 32768 records per step");
  const int dev = gd->ctx->device;
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev >= 0 && dev < kMaxScratchDevices) base = g_scratch_ptr[dev];
  }
  if (!base) return fail(gd->ctx, AJG_ERR_INTERNAL, "guide scratch not set up on this device");
  const uint32_t npad = (n + 15) / 16 * 16;
  return train_launch(gd, scratch_bufs(base, npad), samples, n, npad, lr, s);
}

int guide_records(ajg_guide *gd, uint32_t cap, float **rec, uint32_t **ctr, uint32_t *cap_out) {
  if (gd->rec_cap < cap) {
    (void)hipFree(gd->rec);
    gd->rec = nullptr;
    gd->rec_cap = 0;
    if (int st = dev_alloc(gd->ctx, (void **)&gd->rec, 40ull * cap, "guide training records"))
      return st;
    gd->rec_cap = cap;
    hipError_t e = hipMemset(gd->rec_ctr, 0, 4);  // the pool restarts empty
    if (e != hipSuccess) return hip_fail(gd->ctx, e, "guide record counter");
  }
  *rec = gd->rec;
  *ctr = gd->rec_ctr;
  *cap_out = gd->rec_cap;
  return AJG_OK;
}

int guide_steps(const ajg_guide *gd) { return gd->step; }
const float *guide_loss_dev(const ajg_guide *gd) { return gd->lossbuf + 1; }
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
      (st = dev_alloc(ctx, (void **)&gd->partial, 4ull * kWgBlocks * kGParams, "guide partial")) ||
      (st = dev_alloc(ctx, (void **)&gd->scal, 8ull * kFwdMaxBlocks, "guide block sums")) ||
      (st = dev_alloc(ctx, (void **)&gd->lossbuf, 16, "guide loss")) ||
      (st = dev_alloc(ctx, (void **)&gd->rec_ctr, 16, "guide record counter"))) {
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
  if (e == hipSuccess) e = hipMemset(gd->rec_ctr, 0, 16);
  if (e != hipSuccess) {
    guide_free(gd);
    return hip_fail(ctx, e, "guide init");
  }
  hipLaunchKernelGGL(k_guide_cast, dim3((kGParams + 255) / 256), dim3(256), 0, nullptr,
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
  hipLaunchKernelGGL(k_guide_cast, dim3((kGParams + 255) / 256), dim3(256), 0, nullptr,
                     (const float *)gd->w, gd->wt, gd->wi);
  e = hipDeviceSynchronize();
  return e == hipSuccess ? AJG_OK : hip_fail(gd->ctx, e, "guide cast");
}

int ajg_guide_param_count(void) { return kGParams; }

int ajg_guide_forward(ajg_guide *gd, const float *pos, const float *nrm, uint32_t n,
                      float *mixture, float *raw, void *stream) {
  if (!gd || (n && (!pos || !nrm))) return AJG_ERR_INVALID;
  if (int st = injected(gd->ctx)) return st;
  return guide_forward_launch(gd, pos, nrm, n, nullptr, 3, mixture, nullptr, raw,
                              (hipStream_t)stream);
}

int ajg_guide_train_step_noalloc(ajg_guide *gd, const float *samples, uint32_t n, float lr,
                                 void *stream) {
  if (!gd || !samples || n == 0 || !(lr >= 0.f)) return AJG_ERR_INVALID;
  if (int st = injected(gd->ctx)) return st;
  return guide_train_launch_noalloc(gd, samples, n, lr, (hipStream_t)stream);
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
    AJG_HIP(ctx, hipMemcpyAsync(sc, gd->lossbuf, 8, hipMemcpyDeviceToHost, s));
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
