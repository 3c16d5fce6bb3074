// ajg_render.hip — DefaultScene on the device and the wavefront MIS path
// tracer (PathTracing::TraceMIS, ajisai/core/renderer/pt.cpp:98-227, with
// the helpers of renderer/helper.h and TiledIntegrator::Render, pt.cpp:40-73).
//
// Stages per wavefront iteration:
//   trace_closest  extension rays + MIS-BSDF rays   (exact-order traversal)
//   trace_any      shadow rays of the NEE samples
//   shade          1) finish the previous bounce's direct lighting
//                     (direct = sum_lights beta*NEE + beta*MIS-BSDF, added to
//                     the path's radiance in the reference's order),
//                  2) shade the extension hit: PostIntersect, BSDF, light
//                     samples -> shadow rays, MIS-BSDF sample -> ray,
//                     continuation sample, Russian roulette,
//                  3) when a sample ends, accumulate it (Kahan, FP32) and
//                     start the slot's next sample (ray generation).
// Every path slot keeps its own PCG32 stream, so the random numbers each
// sample consumes are those of the reference's per-sample loop.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ajg_guide.h"
#include "ajg_internal.h"
#include "ajg_shade.h"
#include "ajg_traverse.h"
#include "ajg_scene.h"


namespace ajg {

int injected(ajg_ctx *ctx);
int flatten_views(ajg_ctx *ctx, const ajg_mesh_view *views, uint32_t n_views,
                  std::vector<float> *tri, float *box);
int upload_and_build(ajg_ctx *ctx, const std::vector<float> &tri, const float *box,
                     uint32_t r, hipStream_t s, ajg_accel **out);

// Closest-queue entries per path slot: the extension ray and, in scenes with
// area lights, the MIS-BSDF ray (which must find the emitter it hits).
static constexpr size_t kCqPerSlot = 2;

// ------------------------------------------------------------ env tables
// EnvLight constructor weights (env_light.cpp:33-45): luminance of the
// bilinear texture sample at texel centres times sin(theta), luminance in
// double (spectrum.h:80-82).
__global__ void __launch_bounds__(256)
k_env_weights(DevEnv e, float *wts) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= e.w * e.h) return;
  const uint32_t x = i % e.w, y = i / e.w;
  const float v = ((float)y + 0.5f) / (float)e.h;
  const float sin_t = sinf(AJG_PI * ((float)y + 0.5f) / (float)e.h);
  const float u = ((float)x + 0.5f) / (float)e.w;
  wts[i] = luminance(tex_sample(e, u, v)) * sin_t;
}

// Distribution::SetFunction (distribution.h:40-58) for one row (or the
// marginal): prefix sums in FP64 (rounded to float) or FP32 as the
// reference, then normalisation by funcInt.
__device__ void distrib_build(const float *f, uint32_t n, float *func, float *cdf,
                              float *func_int, int fp64) {
  const float inv = 1.f / (float)n;
  cdf[0] = 0.f;
  if (fp64) {
    double acc = 0.0;
    for (uint32_t i = 1; i <= n; ++i) {
      func[i - 1] = f[i - 1];
      acc += (double)f[i - 1] * (double)inv;
      cdf[i] = (float)acc;
    }
  } else {
    for (uint32_t i = 1; i <= n; ++i) {
      func[i - 1] = f[i - 1];
      cdf[i] = cdf[i - 1] + f[i - 1] * inv;
    }
  }
  const float fi = cdf[n];
  *func_int = fi;
  if (fi > 0.f) {
    for (uint32_t i = 0; i <= n; ++i) cdf[i] /= fi;
  } else {
    for (uint32_t i = 0; i <= n; ++i) cdf[i] = (float)i / (float)n;
  }
}

__global__ void __launch_bounds__(64)
k_env_rows(const float *wts, uint32_t w, uint32_t h, float *cond_func, float *cond_cdf,
           float *cond_int, int fp64) {
  const uint32_t y = blockIdx.x * 64 + threadIdx.x;
  if (y >= h) return;
  distrib_build(wts + (size_t)y * w, w, cond_func + (size_t)y * w, cond_cdf + (size_t)y * (w + 1),
                cond_int + y, fp64);
}

__global__ void k_env_marginal(const float *cond_int, uint32_t h, float *marg_func,
                               float *marg_cdf, float *marg_int, int fp64) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    distrib_build(cond_int, h, marg_func, marg_cdf, marg_int, fp64);
}

// --------------------------------------------------------------- render

struct RenderParams {
  DevCamera cam;
  uint32_t W, H, Wp, n_pix;
  uint32_t spp, chunk, n_chunks, total_work;
  int min_bounces, max_bounces;
  float cont_prob;
  int seed_mode;
  uint32_t seed_base;
  uint32_t slots;
  uint32_t tile, tiles_x;  // work-list pixel order: T x T tiles (0: scanline)
  uint32_t flags;          // AJG_RENDER_* bits
  uint32_t use_mis;        // 1: TraceMIS (pt.cpp:98-227), 0: Trace (pt.cpp:229-311)
};

// Pixel of the i-th entry of a work list's pixel order: scanline, or T x T
// tiles in scanline order of tiles (T divides the traced area).
__device__ __forceinline__ uint32_t pixel_of(uint32_t i, uint32_t tile, uint32_t tiles_x,
                                             uint32_t Wp) {
  if (tile == 0) return i;
  const uint32_t tt = tile * tile;
  const uint32_t t = i / tt, r = i - t * tt;
  const uint32_t ty = t / tiles_x, tx = t - ty * tiles_x;
  const uint32_t ry = r / tile, rx = r - ry * tile;
  return (ty * tile + ry) * Wp + tx * tile + rx;
}

// pend_flags
static constexpr uint32_t kPendActive = 1u;
static constexpr uint32_t kPendEnded = 2u;
static constexpr uint32_t kPendLast = 4u;
static constexpr uint32_t kPendMis = 8u;
static constexpr uint32_t kPendHist = 16u;    // guided: the vertex has a history entry
static constexpr uint32_t kPendLight0 = 32u;  // + k
static constexpr uint32_t kPendMisClosest = 1u << 12;  // the MIS-BSDF ray is in the closest queue
// guided: the env light's MIS weights of this vertex use the mixture density
static constexpr uint32_t kPendMisMix = 1u << 13;

struct PathState {
  uint32_t *work;       // current work item (kNone = none)
  uint32_t *sample;     // sample index within the work item's range (absolute)
  uint32_t *sample_end;
  uint32_t *bounce;
  uint32_t *bank;       // accumulator bank of the current work item
  uint64_t *rng;
  float *L, *beta;      // 3 per slot
  // pending direct lighting of the last shaded bounce
  uint32_t *pend_flags, *pend_mis, *pend_shadow;  // shadow: kMaxLights x slots (pl_idx)
  uint32_t n_slots;     // slots the per-light arrays are laid out for
  uint32_t *pend_work, *pend_sample, *pend_bank;
  float *pend_L, *pend_beta, *pend_c, *pend_f, *pend_pdf, *pend_ns;
  float *pend_pb;       // the BSDF's own pdf of the MIS-BSDF direction
  // accumulators: 2 banks x {sum3, comp3, sq3, compsq3}
  float *acc;
  uint32_t *acc_n;      // 2 banks
  // batched renders: job of the current work item and of the pending record
  uint32_t *job, *pend_job;
};

// A striped ray queue as its consumers see it (the striping is described
// with q_append below).
struct QIn {
  const uint32_t *ctr;  // kQStripes counter words, kQStride apart
  uint32_t R;           // entries per stripe
};

struct Queues {
  DevRay *c_rays_in, *c_rays_out;
  float4 *c_hits_in;   // t, leaf slot (int bits), u, v
  const uint32_t *c_inst_in;  // instanced scenes: instance of each hit
  uint32_t *c_count_out;   // kQStripes counter words of the striped queue
  uint32_t c_len;          // entries per stripe
  QIn c_in;                // the traced closest queue's stripes (guided: mixture index)
  DevRay *a_rays_in, *a_rays_out;
  uint8_t *a_occ_in;
  uint32_t *a_count_out;
  uint32_t a_len;
  uint32_t *live;
  uint32_t *work_counter;  // work_stripes counter words, kQStride apart
  uint32_t work_stripes;   // 1 or kWStripes
  uint32_t first_static;  // a slot's first work item is its own index (next_sample)
  uint32_t work_static;   // first_static: items [0, work_static) are the slots' first
  // Live-slot lists: the shade stage runs over the slots the previous
  // shade left live (dense index -> stripe entry -> slot), and appends the
  // slots it leaves live to the next list, so slots with nothing left to
  // do take no lanes (see k_shade).
  const uint32_t *lv_in;
  const uint32_t *lv_ei_in;  // per entry: the slot's extension ray (closest queue index, kNone)
  QIn lv_q;                  // the input list's stripes
  const uint32_t *lv_total;  // its length (written by k_iter_begin)
  uint32_t *lv_out, *lv_ei_out;
  uint32_t *lv_count_out;    // kQStripes counter words of the output list
  uint32_t lv_len;           // entries per stripe
  float *partial;      // total_work x 6 (sum3, sq3) or the film when n_chunks == 1
  uint32_t *partial_n;
  float *per_sample;   // optional
  unsigned long long *counters;  // closest, any, paths
};

// Guided rendering state. The decoded guide mixture of closest-queue entry i
// is mix[40 i ..] (FP16). Training slots (every train_stride-th slot, so the
// records cover the whole image) record training data: the NEE and
// MIS-BSDF directions of each vertex with their MIS-weighted direct
// radiance, and (through a per-slot vertex history) each continuation
// direction with the radiance the rest of the path brought back along it.
// Training record: pos(3) normal(3) direction(3) weight(1), weight =
// luminance(L_in) |cos| / pdf.
struct GuideParams {
  const _Float16 *mix;
  float alpha;          // probability of sampling the guide
  uint32_t train_stride;  // every train_stride-th slot records (0: none)
  uint32_t hist_depth;
  float *hist;          // (slots / train_stride) x hist_depth x 16: pos nrm dir cw L_upto beta
  uint32_t *hist_n;     // indexed by slot / train_stride
  float *rec;
  uint32_t *rec_count;
  uint32_t rec_cap;
  float *pend_pos;      // 3 per slot
  float *pend_tw;       // kMaxLights per slot: NEE training weight before visibility
};

static constexpr int kGRec = 10;
static constexpr int kGHist = 16;
static constexpr int kGMix = 40;

// Batched renders: many (scene, camera, args, film) jobs share one
// wavefront. Global work item w belongs to the job whose [work_begin,
// work_begin + P.total_work) holds it; every queued ray carries its job so
// the traversal kernels trace it in that job's BVH.
struct BatchJob {
  DevScene sc;
  RenderParams P;
  float *partial;
  uint32_t *partial_n;
  float *per_sample;
};

struct BatchParams {
  const BatchJob *jobs;
  const uint32_t *job_begin;        // n_jobs + 1 prefix of work items
  uint32_t n_jobs, slots, total_work;
  uint32_t *remaining;              // per job: work items not finished
  unsigned long long *done_clock;   // per job: device wall clock at completion
  uint16_t *c_job_out, *a_job_out;  // job of each ray appended this iteration
  const uint16_t *c_job_in, *a_job_in;
};

// Where a finished sample goes.
struct AccTarget {
  float *partial;
  uint32_t *partial_n;
  float *per_sample;
  uint32_t n_pix, spp, Wp, tile, tiles_x;
  uint32_t *remaining;              // batched: job completion counter
  unsigned long long *done;
};

__device__ __forceinline__ uint32_t find_job(const BatchParams &B, uint32_t w) {
  uint32_t lo = 0, hi = B.n_jobs;  // job_begin[lo] <= w < job_begin[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (B.job_begin[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void kahan_add(float &sum, float &c, float x) {
  const float y = x - c;
  const float t = sum + y;
  c = (t - sum) - y;
  sum = t;
}

// Wave-aggregated append: returns this lane's slot in the queue.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return kNone;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  return want ? base + (uint32_t)__popcll(mask & lt) : kNone;
}

// Per-render clears in one launch: each entry sets n 32-bit words (or, for
// a byte array, n bytes) to v. Separate hipMemsetAsync calls cost a launch
// each (about 10 us apart), and a fill kernel may be captured in a graph.
struct FillEntry {
  void *p;
  uint64_t n;
  uint32_t v;
  uint32_t bytes;  // 1: p is a byte array
};
static constexpr int kFillEntries = 24;
struct FillTable {
  FillEntry e[kFillEntries];
  int count;
};
__global__ void __launch_bounds__(256) k_fill(FillTable t) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < t.count; ++k) {
    const FillEntry f = t.e[k];
    if (f.bytes) {
      uint8_t *p = (uint8_t *)f.p;
      for (uint64_t i = i0; i < f.n; i += stride) p[i] = (uint8_t)f.v;
    } else {
      uint32_t *p = (uint32_t *)f.p;
      for (uint64_t i = i0; i < f.n; i += stride) p[i] = f.v;
    }
  }
}
struct Fills {
  FillTable t;
  uint64_t max_n = 0;
  Fills() { t.count = 0; }
  void words(void *p, uint64_t n, uint32_t v = 0u) { add(p, n, v, 0u); }
  void bytes(void *p, uint64_t n, uint32_t v = 0u) { add(p, n, v, 1u); }
  void add(void *p, uint64_t n, uint32_t v, uint32_t b) {
    if (!p || n == 0) return;
    t.e[t.count++] = FillEntry{p, n, v, b};
    max_n = std::max(max_n, n);
  }
  bool full() const { return t.count + 4 > kFillEntries; }
  hipError_t launch(hipStream_t s) {
    if (t.count == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, (max_n + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, s, t);
    t.count = 0;
    max_n = 0;
    return hipGetLastError();
  }
};

// Striped ray queues. Atomics on one counter word are served one at a time,
// and every wave of a shade launch appends at the same few points of the
// kernel, so a single counter per queue holds thousands of waves in line
// (measured: one more append per wave on the closest-queue counter cost the
// shade stage 0.85 ms per launch of 64K waves). Each queue is split into
// kQStripes stripes: wave w appends to stripe w % kQStripes, whose entries
// occupy [s R, s R + R) of the queue (R covers the stripe's worst case) and
// whose counter word sits on a 256-byte line of its own. Consumers walk a
// dense index v in [0, total) and map it through the stripes' prefix sums
// to the entry; the entry index is what the shade stage records, and what
// hits and occlusion results are stored at.
static constexpr uint32_t kQStripes = 32;
static constexpr uint32_t kQStride = 64;  // counter words between stripes
static constexpr uint32_t kQWords = kQStripes * kQStride;
// Fetch-cursor stripes of the closest-hit queue kernels (lane refill); the
// cursor words follow the four queue counter blocks, kCursorStride apart.
static constexpr uint32_t kCursorStripes = 16;
// Work-item counter stripes follow the cursors, kQStride apart (see
// next_sample); they are cleared once per render, not per iteration.
static constexpr uint32_t kWStripes = 32;
static constexpr uint32_t kWorkOff = 4 * kQWords + 2 * kCursorStripes * kCursorStride;
// The live-slot lists (two generations, striped like the ray queues) take
// the last two counter blocks.
static constexpr uint32_t kListOff = kWorkOff + kWStripes * kQStride;
static constexpr uint32_t kQctrWords = kListOff + 2 * kQWords;
// A stripe's items are taken only by its own waves (wave w takes from
// stripe w % stripes), so every stripe needs a wave: at most one stripe
// per wave of slots.
static inline uint32_t work_stripes_for(uint32_t slots) {
  return std::max(1u, std::min(kWStripes, (slots + 63) / 64));
}

// Entries one stripe holds: every wave of the stripe appending per_slot
// entries for each of its 64 lanes.
static inline uint32_t stripe_len(size_t slots, size_t per_slot) {
  const size_t waves = (slots + 63) / 64;
  return (uint32_t)((waves + kQStripes - 1) / kQStripes * 64 * per_slot);
}

__device__ __forceinline__ uint32_t q_append(bool want, uint32_t *ctr, uint32_t R, uint32_t stripe) {
  const uint32_t i = wave_append(want, ctr + stripe * kQStride);
  return i == kNone ? kNone : stripe * R + i;
}

// Prefix sums of a queue's stripe counts into s_pre[0 .. kQStripes]
// (block-wide: every thread of the block calls it). Returns the total.
__device__ __forceinline__ uint32_t q_prefix(const QIn &q, uint32_t *s_pre) {
  if (threadIdx.x < 64) {
    const uint32_t lane = threadIdx.x;
    uint32_t v = lane < kQStripes ? q.ctr[lane * kQStride] : 0u;
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    if (lane < kQStripes) s_pre[lane + 1] = v;
    if (lane == 0) s_pre[0] = 0;
  }
  __syncthreads();
  return s_pre[kQStripes];
}

// Entry of dense index v < total: the last stripe whose prefix is <= v.
__device__ __forceinline__ uint32_t q_entry(const uint32_t *s_pre, uint32_t R, uint32_t v) {
  uint32_t k = 0;
  for (uint32_t step = kQStripes / 2; step > 0; step >>= 1)
    if (s_pre[k + step] <= v) k += step;
  return k * R + (v - s_pre[k]);
}

// Queue capacities (entries) for S slots: a closest-queue entry per slot for
// the extension ray and one for the MIS-BSDF ray of scenes with area lights;
// a shadow-queue entry per light and one for the MIS-BSDF ray.
static inline size_t cq_capacity(size_t S) { return (size_t)kQStripes * stripe_len(S, kCqPerSlot); }
static inline size_t aq_capacity(size_t S) {
  return (size_t)kQStripes * stripe_len(S, kMaxLights + 1);
}

// BSDF value as the sampling routine reports it: LambertianReflection's
// Sample_f multiplies in the shading-normal correction that its f() leaves
// out (the plain path's continuation and MIS-BSDF ray use the sampled
// value), Disney's sample returns f() itself.
__device__ __forceinline__ f3 bsdf_eval_sampled(const Bsdf &b, f3 wi, f3 wo) {
  if (b.m.type == AJG_MAT_DISNEY) return disney_eval(b, wi, wo);
  return scale(lambert_eval(b, wi, wo), correct_shading_normal(b.ng, b.ns, wi));
}

struct GSample {
  f3 dir, f;
  float pdf;  // density of the mixture the direction was drawn from
  float pb;   // the BSDF's own pdf of the direction
  bool ok;
};

// One-sample MIS of the guide mixture and the BSDF: with probability alpha
// the direction comes from the mixture, else from the BSDF sampler; the
// estimator divides by the mixture density alpha q + (1 - alpha) p_bsdf.
__device__ __forceinline__ GSample guided_sample(const Bsdf &bs, f3 wr, const _Float16 *gm,
                                                 float alpha, Pcg32 &rng) {
  const float usel = pcg_next_float(rng);
  f3 sam;
  sam.x = pcg_next_float(rng);
  sam.y = pcg_next_float(rng);
  sam.z = pcg_next_float(rng);
  GSample g;
  if (usel < alpha) {
    g.dir = gmix_sample(gm, sam.x, sam.y, sam.z);
    g.ok = true;
  } else {
    const BsdfSample s = bsdf_sample_all(bs, wr, sam);
    g.ok = s.valid;
    g.dir = normalize(s.dir);
  }
  if (!g.ok) return g;
  g.f = bsdf_eval_sampled(bs, g.dir, wr);
  const float pb = bsdf_pdf_all(bs, g.dir, wr);
  g.pb = pb;
  g.pdf = alpha > 0.f ? alpha * gmix_pdf(gm, g.dir) + (1.f - alpha) * pb : pb;
  g.ok = g.pdf > 0.f && isfinite(g.pdf) && !is_black(g.f);
  return g;
}

__device__ __forceinline__ void store_ray(DevRay *q, uint32_t i, const DevRay &r) {
  float4 *p = reinterpret_cast<float4 *>(q + i);
  p[0] = make_float4(r.o.x, r.o.y, r.o.z, r.d.x);
  p[1] = make_float4(r.d.y, r.d.z, r.tmin, r.tmax);
}
__device__ __forceinline__ DevRay load_ray(const DevRay *q, uint32_t i) {
  const float4 *p = reinterpret_cast<const float4 *>(q + i);
  const float4 a = p[0], b = p[1];
  DevRay r;
  r.o = mk3(a.x, a.y, a.z);
  r.d = mk3(a.w, b.x, b.y);
  r.tmin = b.z;
  r.tmax = b.w;
  return r;
}

// Sum of a per-lane count over the wave, added once per wave.
__device__ __forceinline__ void wave_count_add(uint32_t v, unsigned long long *dst) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// Persistent traversal of a queue whose size the shade stage wrote on the
// device: a fixed grid of waves runs the lane-refill traversal of
// ajg_traverse.h (closest-hit or any-hit; one atomic on one of the striped
// fetch cursors per refill); an empty queue costs one short launch.
// 6 blocks per CU: with lane refill from striped cursors measured 1-2% faster
// than 4, 8 or 12 on the heavy mesh (the two queue kernels run concurrently).
static constexpr int kTravBlocksPerCU = 6;
// Every fetch-cursor stripe needs a wave that drains it: wave w of a
// persistent traversal launch takes from stripe w mod kCursorStripes, so a
// grid of fewer than kCursorStripes waves would leave rays untraced.
static constexpr uint32_t kMinTravGrid = (kCursorStripes * 64 + kTravThreads - 1) / kTravThreads;
static_assert(kMinTravGrid * (kTravThreads / 64) >= kCursorStripes, "traversal grid covers every stripe");
static inline uint32_t trav_grid_size(uint32_t blocks) { return std::max(blocks, kMinTravGrid); }
static constexpr int kItersPerCheck = 4;
// Work items per path slot aimed at by a plain seed_mode-1 render's
// automatic sample chunking (setup_params).
static constexpr uint64_t kItemsPerSlot = 8;

template <int SPILL, bool COUNT>
__global__ void __launch_bounds__(kTravThreads)
k_pq_closest(DevBVH bvh, const DevRay *rays, float4 *hits, QIn qin, uint32_t *cursor,
             unsigned long long *counters) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  uint32_t nn = 0, nt = 0;
  persistent_closest<SPILL, COUNT, false, kLdsStack, kCursorStripes>(
      bvh, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v, DevBVH &) {
        const DevRay ray = load_ray(rays, q_entry(s_pre, qin.R, v));
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, const THit &h) {
        hits[q_entry(s_pre, qin.R, v)] = make_float4(h.t, __int_as_float(h.slot), h.u, h.v);
      },
      &nn, &nt);
  if (COUNT) {
    wave_count_add(nn, counters);
    wave_count_add(nt, counters + 1);
  }
}

template <int SPILL, bool COUNT>
__global__ void __launch_bounds__(kTravThreads)
k_pq_any(DevBVH bvh, const DevRay *rays, uint8_t *occ, QIn qin, uint32_t *cursor,
         unsigned long long *counters) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  uint32_t nn = 0, nt = 0;
  persistent_any<SPILL, COUNT, false, kLdsStack, kCursorStripes>(
      bvh, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v, DevBVH &) {
        const DevRay ray = load_ray(rays, q_entry(s_pre, qin.R, v));
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, bool hit) { occ[q_entry(s_pre, qin.R, v)] = hit ? 1 : 0; }, &nn, &nt);
  if (COUNT) {
    wave_count_add(nn, counters);
    wave_count_add(nt, counters + 1);
  }
}

// Batched renders: the BVH of each ray is its job's.
template <int SPILL, int LDS>
__global__ void __launch_bounds__(kTravThreads)
k_pq_closest_b(const BatchJob *jobs, const uint16_t *job_of, const DevRay *rays, float4 *hits,
               QIn qin, uint32_t *cursor) {
  __shared__ int s_node[LDS * kTravThreads];
  __shared__ float s_entry[LDS * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  const DevBVH first = jobs[0].sc.bvh;
  uint32_t nn = 0, nt = 0;
  persistent_closest<SPILL, false, true, LDS, kCursorStripes>(
      first, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v, DevBVH &b) {
        const uint32_t i = q_entry(s_pre, qin.R, v);
        b = jobs[job_of[i]].sc.bvh;
        const DevRay ray = load_ray(rays, i);
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, const THit &h) {
        hits[q_entry(s_pre, qin.R, v)] = make_float4(h.t, __int_as_float(h.slot), h.u, h.v);
      },
      &nn, &nt);
}

template <int SPILL, int LDS>
__global__ void __launch_bounds__(kTravThreads)
k_pq_any_b(const BatchJob *jobs, const uint16_t *job_of, const DevRay *rays, uint8_t *occ,
           QIn qin, uint32_t *cursor) {
  __shared__ int s_node[LDS * kTravThreads];
  __shared__ float s_entry[LDS * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  const DevBVH first = jobs[0].sc.bvh;
  uint32_t nn = 0, nt = 0;
  persistent_any<SPILL, false, true, LDS, kCursorStripes>(
      first, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v, DevBVH &b) {
        const uint32_t i = q_entry(s_pre, qin.R, v);
        b = jobs[job_of[i]].sc.bvh;
        const DevRay ray = load_ray(rays, i);
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, bool hit) { occ[q_entry(s_pre, qin.R, v)] = hit ? 1 : 0; }, &nn, &nt);
}

// Instanced scenes: two-level traversal (top-level tree over instances).
// (6 waves per SIMD: the register budget that measured fastest; 8 spills)
template <int SPILL, bool COUNT>
__global__ void __launch_bounds__(kTravThreads, 6)
k_pq_closest2(DevBVH tlas, const DevInst *insts, const DevBlas *blases, const DevRay *rays,
              float4 *hits, uint32_t *hit_inst, QIn qin, uint32_t *cursor,
              unsigned long long *counters) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  uint32_t nn = 0, nt = 0;
  persistent_closest_2l<SPILL, kLdsStack, kCursorStripes, COUNT>(
      tlas, insts, blases, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v) {
        const DevRay ray = load_ray(rays, q_entry(s_pre, qin.R, v));
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, const THit &h, uint32_t inst) {
        const uint32_t i = q_entry(s_pre, qin.R, v);
        hits[i] = make_float4(h.t, __int_as_float(h.slot), h.u, h.v);
        hit_inst[i] = inst;
      },
      &nn, &nt);
  if (COUNT) {
    wave_count_add(nn, counters);
    wave_count_add(nt, counters + 1);
  }
}

// (8 waves per SIMD measured faster for this kernel than the compiler's
// default register budget)
template <int SPILL, bool COUNT>
__global__ void __launch_bounds__(kTravThreads, 8)
k_pq_any2(DevBVH tlas, const DevInst *insts, const DevBlas *blases, const DevRay *rays,
          uint8_t *occ, QIn qin, uint32_t *cursor, unsigned long long *counters) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (n == 0) return;
  uint32_t nn = 0, nt = 0;
  persistent_any_2l<SPILL, kLdsStack, kCursorStripes, COUNT>(
      tlas, insts, blases, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t v) {
        const DevRay ray = load_ray(rays, q_entry(s_pre, qin.R, v));
        return make_tray(ray.o, ray.d, ray.tmin, ray.tmax);
      },
      [&](uint32_t v, bool hit) { occ[q_entry(s_pre, qin.R, v)] = hit ? 1 : 0; }, &nn, &nt);
  if (COUNT) {
    wave_count_add(nn, counters);
    wave_count_add(nt, counters + 1);
  }
}

// C-ABI scene traces of instanced scenes (rays in the ajg_ray layout).
template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_scene_closest2(DevBVH tlas, const DevInst *insts, const DevBlas *blases, const ajg_ray *rays,
                 uint32_t n, ajg_hit *hits, uint32_t *hit_inst, uint32_t *cursor) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  persistent_closest_2l<SPILL, kLdsStack>(
      tlas, insts, blases, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t i) {
        const ajg_ray ray = rays[i];
        return make_tray(ld3(ray.o), ld3(ray.d), ray.tmin, ray.tmax);
      },
      [&](uint32_t i, const THit &h, uint32_t inst) {
        ajg_hit o;
        o.t = h.t;
        o.tri_id = h.slot >= 0 ? blases[insts[inst].info.x].perm[h.slot] : kNone;
        o.u = h.u;
        o.v = h.v;
        hits[i] = o;
        if (hit_inst) hit_inst[i] = inst;
      });
}

template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_scene_any2(DevBVH tlas, const DevInst *insts, const DevBlas *blases, const ajg_ray *rays,
             uint32_t n, uint8_t *occ) {
  __shared__ int s_node[kLdsStack * kTravThreads];
  __shared__ float s_entry[kLdsStack * kTravThreads];
  const uint32_t i = blockIdx.x * kTravThreads + threadIdx.x;
  if (i >= n) return;
  const ajg_ray ray = rays[i];
  const TRay r = make_tray(ld3(ray.o), ld3(ray.d), ray.tmin, ray.tmax);
  occ[i] = trace_any_2l<SPILL, kLdsStack>(tlas, insts, blases, r, s_node + threadIdx.x,
                                         s_entry + threadIdx.x) ? 1 : 0;
}

__global__ void k_clear_word(uint32_t *w) {
  if (threadIdx.x == 0) *w = 0u;
}

__global__ void __launch_bounds__(256)
k_one_level_instance(const ajg_hit *hits, uint32_t n, uint32_t *inst) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) inst[i] = hits[i].tri_id == kNone ? kNone : 0u;
}

template <bool COUNT>
static void launch_pq2(int stack, const DevScene &ds, bool closest, const DevRay *rays, void *out,
                       uint32_t *hit_inst, QIn count, uint32_t *cursor,
                       unsigned long long *ctr, uint32_t grid, hipStream_t s) {
  const dim3 g(grid), b(kTravThreads);
  if (closest) {
    if (stack == kSpillS)
      hipLaunchKernelGGL((k_pq_closest2<kSpillS, COUNT>), g, b, 0, s, ds.bvh, ds.inst, ds.blas,
                         rays, (float4 *)out, hit_inst, count, cursor, ctr);
    else
      hipLaunchKernelGGL((k_pq_closest2<kSpillL, COUNT>), g, b, 0, s, ds.bvh, ds.inst, ds.blas,
                         rays, (float4 *)out, hit_inst, count, cursor, ctr);
  } else {
    if (stack == kSpillS)
      hipLaunchKernelGGL((k_pq_any2<kSpillS, COUNT>), g, b, 0, s, ds.bvh, ds.inst, ds.blas, rays,
                         (uint8_t *)out, count, cursor, ctr);
    else
      hipLaunchKernelGGL((k_pq_any2<kSpillL, COUNT>), g, b, 0, s, ds.bvh, ds.inst, ds.blas, rays,
                         (uint8_t *)out, count, cursor, ctr);
  }
}

template <bool COUNT>
static void launch_pq(int stack, const DevBVH &bvh, bool closest, const DevRay *rays, void *out,
                      QIn count, uint32_t *cursor, unsigned long long *ctr,
                      uint32_t grid, hipStream_t s) {
  const dim3 g(grid), b(kTravThreads);
  if (closest) {
    if (stack == kSpillS)
      hipLaunchKernelGGL((k_pq_closest<kSpillS, COUNT>), g, b, 0, s, bvh, rays, (float4 *)out, count,
                         cursor, ctr);
    else
      hipLaunchKernelGGL((k_pq_closest<kSpillL, COUNT>), g, b, 0, s, bvh, rays, (float4 *)out, count,
                         cursor, ctr);
  } else {
    if (stack == kSpillS)
      hipLaunchKernelGGL((k_pq_any<kSpillS, COUNT>), g, b, 0, s, bvh, rays, (uint8_t *)out, count,
                         cursor, ctr);
    else
      hipLaunchKernelGGL((k_pq_any<kSpillL, COUNT>), g, b, 0, s, bvh, rays, (uint8_t *)out, count,
                         cursor, ctr);
  }
}

template <int SPILL>
static void launch_pq_b(int lds, bool closest, const BatchJob *jobs, const uint16_t *job_of,
                        const DevRay *rays, void *out, QIn count, uint32_t *cursor,
                        uint32_t grid, hipStream_t s) {
  const dim3 g(grid), b(kTravThreads);
#define AJG_PQB(L)                                                                            \
  if (closest)                                                                                \
    hipLaunchKernelGGL((k_pq_closest_b<SPILL, L>), g, b, 0, s, jobs, job_of, rays, (float4 *)out, \
                       count, cursor);                                                        \
  else                                                                                        \
    hipLaunchKernelGGL((k_pq_any_b<SPILL, L>), g, b, 0, s, jobs, job_of, rays, (uint8_t *)out, count, \
                       cursor);
  if (lds == 4) {
    AJG_PQB(4)
  } else if (lds == 16) {
    AJG_PQB(16)
  } else {
    AJG_PQB(8)
  }
#undef AJG_PQB
}

// Event pairs around launches (AJG_RENDER_TIMING).
struct StageTimer {
  bool on = false;
  std::vector<hipEvent_t> evs[5];
  size_t used[5] = {0, 0, 0, 0, 0};
  uint32_t launches[5] = {0, 0, 0, 0, 0};
  hipEvent_t get(int k) {
    if (used[k] == evs[k].size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      evs[k].push_back(e);
    }
    return evs[k][used[k]++];
  }
  void begin(int k, hipStream_t s) {
    ++launches[k];
    if (on) (void)hipEventRecord(get(k), s);
  }
  void end(int k, hipStream_t s) {
    if (on) (void)hipEventRecord(get(k), s);
  }
  float total_ms(int k) {
    float t = 0.f;
    for (size_t i = 0; i + 1 < used[k]; i += 2) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, evs[k][i], evs[k][i + 1]);
      t += ms;
    }
    return t;
  }
  ~StageTimer() {
    for (auto &v : evs)
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
  }
};

// Start of a wavefront iteration: add the sizes of the queues traced in the
// previous iteration to the totals, clear the queues the shade stage is
// about to fill, the live flag and the traversal fetch cursors, and write
// the length of the live-slot list the shade stage reads to ctr[5]. qctr
// holds the striped counters of closest queues 0 and 1, then shadow queues
// 0 and 1, kQWords apart, then the fetch-cursor stripes, the work-item
// counters and the live-slot lists 0 and 1. One wave.
__global__ void k_iter_begin(uint32_t *ctr, uint32_t *qctr, int cur, int nxt,
                             unsigned long long *totals) {
  const uint32_t lane = threadIdx.x;
  uint32_t c = 0, a = 0, l = 0;
  if (lane < kQStripes) {
    c = qctr[cur * kQWords + lane * kQStride];
    a = qctr[(2 + cur) * kQWords + lane * kQStride];
    l = qctr[kListOff + cur * kQWords + lane * kQStride];
    qctr[nxt * kQWords + lane * kQStride] = 0;
    qctr[(2 + nxt) * kQWords + lane * kQStride] = 0;
    qctr[kListOff + nxt * kQWords + lane * kQStride] = 0;
  }
  if (lane < 2 * kCursorStripes) qctr[4 * kQWords + lane * kCursorStride] = 0;
  for (uint32_t o = 32; o > 0; o >>= 1) {
    c += __shfl_down(c, o, 64);
    a += __shfl_down(a, o, 64);
    l += __shfl_down(l, o, 64);
  }
  if (lane == 0) {
    totals[0] += c;
    totals[1] += a;
    ctr[4] = 0;
    ctr[5] = l;
  }
}

// The first live-slot list: every slot, in the order a shade launch over
// the slots would append them (wave w to stripe w % kQStripes), with the
// stripe counts. qctr_list: the list's kQStripes counter words.
__global__ void __launch_bounds__(256)
k_lv_init(uint32_t *lv, uint32_t *lv_ei, uint32_t *qctr_list, uint32_t S, uint32_t R) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < S) {
    const uint32_t w = i >> 6;
    const uint32_t e = (w % kQStripes) * R + (w / kQStripes) * 64 + (i & 63);
    lv[e] = i;
    lv_ei[e] = kNone;
  }
  if (blockIdx.x == 0 && threadIdx.x < kQStripes) {
    const uint32_t st = threadIdx.x, full = S >> 6, rem = S & 63;
    const uint32_t n = full > st ? (full - 1 - st) / kQStripes + 1 : 0u;
    qctr_list[st * kQStride] = 64 * n + ((rem && full % kQStripes == st) ? rem : 0u);
  }
}

// Device wall clock at the start of a batched render (job completion times
// are measured against it).
__global__ void k_clock(unsigned long long *dst) {
  if (threadIdx.x == 0) *dst = wall_clock64();
}

__device__ __forceinline__ f3 ld_f3(const float *p, uint32_t i) {
  return mk3(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]);
}
__device__ __forceinline__ void st_f3(float *p, uint32_t i, f3 v) {
  p[3 * (size_t)i] = v.x;
  p[3 * (size_t)i + 1] = v.y;
  p[3 * (size_t)i + 2] = v.z;
}

// Per-light pending entries (shadow-ray index, contribution) are light
// major: light k of every slot, then light k + 1, so the step-1 kernel's
// reads of one light fill whole cache lines (heavy mesh shading 45.8 ->
// 44.8 ms per render against slot-major, where a one-light scene used a
// quarter of each line).
__device__ __forceinline__ uint32_t pl_idx(const PathState &ps, uint32_t slot, uint32_t k) {
  return k * ps.n_slots + slot;
}

__device__ __forceinline__ void emit_record(const GuideParams &G, f3 pos, f3 nrm, f3 dir, float w) {
  const bool want = w > 0.f && isfinite(w);
  const uint32_t i = wave_append(want, G.rec_count);
  if (want && i < G.rec_cap) {
    float *r = G.rec + (size_t)i * kGRec;
    st3(r, pos);
    st3(r + 3, nrm);
    st3(r + 6, dir);
    r[9] = w;
  }
}

// Path end: one record per recorded continuation k with
// L_in = (L_final - L_upto_k) / beta_{k+1} (what the rest of the path
// brought back along the direction, per unit throughput).
__device__ void emit_history(const GuideParams &G, uint32_t hslot, f3 Lf) {
  const uint32_t n = G.hist_n[hslot];
  for (uint32_t k = 0; k < n; ++k) {
    const float *h = G.hist + ((size_t)hslot * G.hist_depth + k) * kGHist;
    const float lx = h[13] > 0.f ? fmaxf(Lf.x - h[10], 0.f) / h[13] : 0.f;
    const float ly = h[14] > 0.f ? fmaxf(Lf.y - h[11], 0.f) / h[14] : 0.f;
    const float lz = h[15] > 0.f ? fmaxf(Lf.z - h[12], 0.f) / h[15] : 0.f;
    emit_record(G, ld3(h), ld3(h + 3), ld3(h + 6), lum_f(mk3(lx, ly, lz)) * h[9]);
  }
  G.hist_n[hslot] = 0;
}

// Accumulate a finished sample into its work item's bank; flush the bank
// when it was the item's last sample (and, in a batch, count the job down).
__device__ void accumulate(const PathState &ps, const AccTarget &t, uint32_t slot, uint32_t bank,
                           uint32_t work, uint32_t sample, f3 L, bool last) {
  float *a = ps.acc + ((size_t)slot * 2 + bank) * 12;
  kahan_add(a[0], a[3], L.x);
  kahan_add(a[1], a[4], L.y);
  kahan_add(a[2], a[5], L.z);
  kahan_add(a[6], a[9], L.x * L.x);
  kahan_add(a[7], a[10], L.y * L.y);
  kahan_add(a[8], a[11], L.z * L.z);
  uint32_t &n = ps.acc_n[(size_t)slot * 2 + bank];
  n += 1;
  const uint32_t chunk = work / t.n_pix;
  const uint32_t p = pixel_of(work - chunk * t.n_pix, t.tile, t.tiles_x, t.Wp);
  if (t.per_sample) {
    float *o = t.per_sample + 3 * ((size_t)p * t.spp + sample);
    o[0] = L.x;
    o[1] = L.y;
    o[2] = L.z;
  }
  if (last) {
    const size_t cell = (size_t)chunk * t.n_pix + p;  // chunk-major, pixel-minor
    float *dst = t.partial + cell * 6;
    dst[0] = a[0] - a[3];
    dst[1] = a[1] - a[4];
    dst[2] = a[2] - a[5];
    dst[3] = a[6] - a[9];
    dst[4] = a[7] - a[10];
    dst[5] = a[8] - a[11];
    t.partial_n[cell] = n;
    for (int k = 0; k < 12; ++k) a[k] = 0.f;
    n = 0;
    if (t.remaining) {
      __threadfence();
      if (atomicSub(t.remaining, 1u) == 1u) *t.done = wall_clock64();
    }
  }
}

__device__ __forceinline__ AccTarget acc_single(const Queues &q, const RenderParams &P) {
  AccTarget t;
  t.partial = q.partial;
  t.partial_n = q.partial_n;
  t.per_sample = q.per_sample;
  t.n_pix = P.n_pix;
  t.spp = P.spp;
  t.Wp = P.Wp;
  t.tile = P.tile;
  t.tiles_x = P.tiles_x;
  t.remaining = nullptr;
  t.done = nullptr;
  return t;
}

__device__ __forceinline__ AccTarget acc_job(const BatchParams &B, uint32_t j) {
  const BatchJob &J = B.jobs[j];
  AccTarget t;
  t.partial = J.partial;
  t.partial_n = J.partial_n;
  t.per_sample = J.per_sample;
  t.n_pix = J.P.n_pix;
  t.spp = J.P.spp;
  t.Wp = J.P.Wp;
  t.tile = J.P.tile;
  t.tiles_x = J.P.tiles_x;
  t.remaining = B.remaining + j;
  t.done = B.done_clock + j;
  return t;
}

// Start the next sample of the slot (fetching a new work item when the
// current one is exhausted) and emit its camera ray. Wave-uniform call.
// P0 is the render's parameters (single renders); a batch takes each job's.
template <bool BATCH>
__device__ void next_sample(const PathState &ps, const Queues &q, const RenderParams &P0,
                            const BatchParams &B, uint32_t slot, bool need, Pcg32 &rng,
                            bool &live, uint32_t stripe, uint32_t &ei_out) {
  uint32_t work = need ? ps.work[slot] : kNone;
  uint32_t job = BATCH && need ? ps.job[slot] : 0u;
  uint32_t sample = 0;
  bool fresh = false;
  if (need && work != kNone) {
    sample = ps.sample[slot] + 1;
    if (sample >= ps.sample_end[slot]) work = kNone;
  }
  const bool fetch = need && work == kNone;
  // A slot's first work item is its own index: every slot fetches in the
  // render's first shade (its sample_end is still 0), and one atomic per
  // wave would hold S/64 waves in line; items [0, work_static) are taken
  // that way. Later fetches take the next items from a counter. With
  // work_stripes > 1 the counter is striped like the fetch cursors of the
  // traversal kernels: wave w takes from stripe s = w % work_stripes,
  // whose value j stands for item work_static + (s + stripes (j / 64)) 64
  // + j % 64, so each stripe hands out every stripes-th 64-item block in
  // order (atomics on one word are served one at a time, and with several
  // work items per slot every wave fetches many times). A stripe whose
  // items are used up stops taking atomics (a plain read first: the
  // counter only grows, so a stale value only costs the atomic). A lane
  // whose fetch finds its stripe used up stays idle for the rest of the
  // render; the stripe's other waves finish its items. The stripe belongs
  // to the slot (its wave of slots), and a shade wave runs the slots of a
  // live-slot list, whose lanes may come from several waves of slots: one
  // append per distinct stripe among the fetching lanes.
  const bool first = q.first_static && fetch && ps.sample_end[slot] == 0u;
  const uint32_t limit = BATCH ? B.total_work : P0.total_work;
  const uint32_t ws = q.work_stripes;
  const uint32_t wst = ws > 1 ? (slot >> 6) % ws : 0u;
  uint32_t *const wctr = q.work_counter + wst * kQStride;
  auto item_at = [&](uint32_t j) -> uint64_t {
    const uint64_t base = q.first_static ? q.work_static : 0u;
    return ws > 1 ? base + (((uint64_t)wst + (uint64_t)ws * (j >> 6)) << 6) + (j & 63u)
                  : base + j;
  };
  const bool want = fetch && !first && __ballot(fetch && !first) != 0ull && item_at(*wctr) < limit;
  uint32_t jw = kNone;
  for (uint64_t left = __ballot(want); left != 0ull;) {
    const uint32_t ls = __shfl(wst, __ffsll((unsigned long long)left) - 1, 64);
    const bool mine = want && wst == ls;
    const uint32_t j = wave_append(mine, q.work_counter + ls * kQStride);
    if (mine) jw = j;
    left &= ~__ballot(mine);
  }
  uint64_t w = jw == kNone ? (uint64_t)limit : item_at(jw);
  if (first) w = slot;
  if (fetch) {
    if (w < (uint64_t)limit) {
      uint32_t local = (uint32_t)w;
      if (BATCH) {
        job = find_job(B, (uint32_t)w);
        local = (uint32_t)w - B.job_begin[job];
      }
      const RenderParams &P = BATCH ? B.jobs[job].P : P0;
      work = local;
      fresh = true;
      const uint32_t c = local / P.n_pix;
      sample = c * P.chunk;
      ps.sample_end[slot] = min(P.spp, sample + P.chunk);
      ps.bank[slot] ^= 1u;
    } else {
      work = kNone;
    }
  }
  if (!need) return;
  ps.work[slot] = work;
  if (work == kNone) {
    ei_out = kNone;
    return;
  }
  if (BATCH) ps.job[slot] = job;
  const RenderParams &P = BATCH ? B.jobs[job].P : P0;
  ps.sample[slot] = sample;
  const uint32_t p = pixel_of(work % P.n_pix, P.tile, P.tiles_x, P.Wp);
  const uint32_t x = p % P.Wp, y = p / P.Wp;
  if (P.seed_mode == 0) {
    if (fresh) pcg_seed(rng, (uint64_t)(int64_t)(int)(x + y * P.W));
  } else {
    pcg_seed(rng, sample_seed(P.seed_base, p, sample));
  }
  // TiledIntegrator::Render per-sample draws (pt.cpp:58-61)
  const float u = (float)x + pcg_next_float(rng);
  const float v = (float)y + pcg_next_float(rng);
  const float l0 = pcg_next_float(rng);
  const float l1 = pcg_next_float(rng);
  const DevRay r = camera_ray(P.cam, u, v, l0, l1);
  st_f3(ps.L, slot, mk3(0.f, 0.f, 0.f));
  st_f3(ps.beta, slot, mk3(1.f, 1.f, 1.f));
  ps.bounce[slot] = 0;
  live = true;
  // Russian roulette for bounce 0 only if min_bounces < 0 (pt.cpp:107-110)
  bool alive = true;
  if (0 > P.min_bounces) {
    if (pcg_next_float(rng) > P.cont_prob) {
      alive = false;
    } else {
      st_f3(ps.beta, slot, divs(mk3(1.f, 1.f, 1.f), P.cont_prob));
    }
  }
  if (P.max_bounces <= 0) alive = false;
  // A path that dies before its first intersection contributes zero.
  const uint32_t idx = q_append(alive, q.c_count_out, q.c_len, stripe);
  if (alive) {
    store_ray(q.c_rays_out, idx, r);
    if (BATCH) B.c_job_out[idx] = (uint16_t)job;
    ei_out = idx;
  } else {
    ei_out = kNone;
  }
  if (!alive) {
    // zero-radiance sample: accumulate immediately, then the next sample
    // starts in the following iteration (slot stays live).
    const bool last = sample + 1 >= ps.sample_end[slot];
    accumulate(ps, BATCH ? acc_job(B, job) : acc_single(q, P), slot, ps.bank[slot], work, sample,
               mk3(0.f, 0.f, 0.f), last);
    ps.pend_flags[slot] = 0;
  }
}

// Shading step 1 for one live slot (k_shade_pend): finish the previous
// bounce's direct lighting (shadow-ray results, the MIS-BSDF ray's light)
// and accumulate a finished sample.
template <bool GUIDED, bool BATCH>
__device__ __forceinline__ void shade_pend(const DevScene &sc0, const RenderParams &P0,
                                           const PathState &ps, const Queues &q,
                                           const GuideParams &G, const BatchParams &B,
                                           uint32_t slot, bool train, uint32_t hslot) {
  // ---- 1. finish the previous bounce's direct lighting
  const uint32_t pf = ps.pend_flags[slot];
  if (pf & kPendActive) {
    const uint32_t pj = BATCH ? ps.pend_job[slot] : 0u;
    const DevScene &sc = BATCH ? B.jobs[pj].sc : sc0;
    f3 tpos, tnrm;
    if (train) {
      tpos = ld_f3(G.pend_pos, slot);
      tnrm = normalize(ld_f3(ps.pend_ns, slot));
    }
    f3 direct = mk3(0.f, 0.f, 0.f);
    for (uint32_t k = 0; k < sc.n_lights; ++k) {
      f3 c = mk3(0.f, 0.f, 0.f);
      if (pf & (kPendLight0 << k)) {
        const uint32_t si = ps.pend_shadow[pl_idx(ps, slot, k)];
        if (!q.a_occ_in[si]) {
          c = ld_f3(ps.pend_c, pl_idx(ps, slot, k));
          if (train)
            emit_record(G, tpos, tnrm, load_ray(q.a_rays_in, si).d,
                        G.pend_tw[(size_t)slot * kMaxLights + k]);
        }
      }
      direct = add(direct, c);
    }
    f3 mis = mk3(0.f, 0.f, 0.f);
    if (pf & kPendMis) {
      // MISSampleBsdf (helper.h:208-262). Without area lights the ray
      // contributes only when it escapes to the environment, so it is
      // traced as an any-hit query (it escapes iff the reference's
      // closest-hit query misses); with area lights it is a closest-hit
      // query, and a hit on the emitting side of an emissive primitive
      // adds that light.
      const uint32_t mi = ps.pend_mis[slot];
      const bool closest = (pf & kPendMisClosest) != 0;
      const DevRay mr = load_ray(closest ? q.c_rays_in : q.a_rays_in, mi);
      float4 mh = make_float4(0.f, 0.f, 0.f, 0.f);
      if (closest) mh = q.c_hits_in[mi];
      const bool escaped = closest ? __float_as_int(mh.y) < 0 : !q.a_occ_in[mi];
      if (!escaped && closest) {
        const Inct hi = post_intersect(sc, mr, mh.x, __float_as_int(mh.y), mh.z, mh.w,
                                       sc.inst ? q.c_inst_in[mi] : 0u);
        if (hi.light && dot(hi.ng, hi.wr) > 0) {  // AreaLight::Radiance
          const DevLight &l = sc.lights[hi.light - 1];
          if (!is_black(l.radiance)) {
            const f3 f = scale(mul(l.radiance, ld_f3(ps.pend_f, slot)),
                               fabsf(dot(ld_f3(ps.pend_ns, slot), mr.d)));
            const float pdf = ps.pend_pdf[slot];
            const float lpdf = area_pdf(l, mr.o, hi.pos, hi.ng);
            // guided: the weight keeps the BSDF's own pdf (the area
            // light's pdf is not a density of the same measure, so the
            // two weights do not sum to one; keeping the plain path's
            // weight functions keeps the value the estimator converges to)
            const float wb = power_heuristic(GUIDED ? ps.pend_pb[slot] : pdf, lpdf);
            mis = scale(divs(f, pdf), wb);
            if (train)
              emit_record(G, tpos, tnrm, mr.d, lum_f(l.radiance) * fabsf(dot(tnrm, mr.d)) / pdf * wb);
          }
        }
      }
      if (escaped && sc.env_index >= 0) {
        const DevEnv &e = sc.lights[sc.env_index].env;
        const f3 le = env_radiance(e, mr.d);
        if (!is_black(le)) {
          const f3 f = scale(mul(le, ld_f3(ps.pend_f, slot)),
                             fabsf(dot(ld_f3(ps.pend_ns, slot), mr.d)));
          const float pdf = ps.pend_pdf[slot];
          const float lpdf = env_pdf(e, mr.d);
          // guided: the mixture density is the MIS partner only where the
          // light-side and BSDF-side integrands are the same function
          // (flag set in the shade that queued this ray); elsewhere the
          // plain path's weight functions, still dividing by the density
          // the direction was drawn from
          const float wb = power_heuristic(
              (!GUIDED || (pf & kPendMisMix)) ? pdf : ps.pend_pb[slot], lpdf);
          mis = scale(divs(f, pdf), wb);
          if (train) emit_record(G, tpos, tnrm, mr.d, lum_f(le) * fabsf(dot(tnrm, mr.d)) / pdf * wb);
        }
      }
    }
    direct = add(direct, mul(ld_f3(ps.pend_beta, slot), mis));
    const f3 Lf = add(ld_f3(ps.pend_L, slot), scale(direct, 1.f));
    if (train) {
      if (pf & kPendEnded) {
        emit_history(G, hslot, Lf);
      } else if (pf & kPendHist) {
        const uint32_t hn = G.hist_n[hslot];
        st3(G.hist + ((size_t)hslot * G.hist_depth + hn - 1) * kGHist + 10, Lf);
      }
    }
    if (pf & kPendEnded) {
      accumulate(ps, BATCH ? acc_job(B, pj) : acc_single(q, P0), slot, ps.pend_bank[slot],
                 ps.pend_work[slot], ps.pend_sample[slot], Lf, (pf & kPendLast) != 0);
    } else {
      st_f3(ps.L, slot, Lf);
    }
    ps.pend_flags[slot] = 0;
  }

}

// Shading step 2 for one live slot (k_shade): shade the extension hit of
// the current sample (NEE, MIS-BSDF and continuation rays). (As a function
// of its own the plain kernel compiles to fewer spills. A second,
// wave-uniform-job copy for batches, taken when a ballot shows that all
// lanes shade slots of one job, made the batched kernel spill 34 registers
// and the batch 3% slower, and was dropped.)
template <bool GUIDED, bool BATCH>
__device__ __forceinline__ void shade_slot(const DevScene &sc0, const RenderParams &P0,
                                           const PathState &ps, const Queues &q,
                                           const GuideParams &G, const BatchParams &B,
                                           uint32_t slot, uint32_t stripe, bool train,
                                           uint32_t hslot, const uint32_t *s_pre, Pcg32 &rng,
                                           uint32_t ei, uint32_t &ei_next, bool &live,
                                           bool &need_new) {
  // ---- 2. shade the extension hit of the current sample
  if (ei == kNone) {
    if (ps.work[slot] != kNone) {
      // a sample ended before its first intersection: start the next one
      need_new = true;
      live = true;
    } else if (ps.sample_end[slot] == 0) {
      need_new = true;  // never had work: try to fetch some
    }
  }
  if (ei != kNone) {
    live = true;
    const uint32_t j = BATCH ? ps.job[slot] : 0u;
    const DevScene &sc = BATCH ? B.jobs[j].sc : sc0;
    const RenderParams &P = BATCH ? B.jobs[j].P : P0;
    const DevRay ray = load_ray(q.c_rays_in, ei);
    const float4 h = q.c_hits_in[ei];
    const int ls = __float_as_int(h.y);
    const uint32_t b = ps.bounce[slot];
    f3 L = ld_f3(ps.L, slot);
    f3 beta = ld_f3(ps.beta, slot);
    const uint32_t work = ps.work[slot], sample = ps.sample[slot];
    const bool last = sample + 1 >= ps.sample_end[slot];
    if (ls < 0) {
      // TraceMIS adds the environment at the first bounce only (pt.cpp:
      // 114-121; later escapes are its MIS-BSDF term); Trace adds it at
      // every bounce, unweighted by the throughput (pt.cpp:245-249)
      if ((b == 0 || !P.use_mis) && sc.env_index >= 0)
        L = add(L, env_radiance(sc.lights[sc.env_index].env, ray.d));
      if (train) emit_history(G, hslot, L);
      accumulate(ps, BATCH ? acc_job(B, j) : acc_single(q, P), slot, ps.bank[slot], work, sample,
                 L, last);
      need_new = true;
    } else {
      const Inct it = post_intersect(sc, ray, h.x, ls, h.z, h.w, sc.inst ? q.c_inst_in[ei] : 0u);
      const Bsdf bs = make_bsdf(sc.mats[it.material], it.ng, it.ns);
      const _Float16 *gm = nullptr;
      if (GUIDED) {
        const uint32_t k = ei / q.c_in.R;
        gm = G.mix + (size_t)(s_pre[k] + (ei - k * q.c_in.R)) * kGMix;
      }
      // The guide is mixed in only where the BSDF's pdf is the density of
      // its own sampler (Lambert). The Disney sampler's selection between
      // its specular and diffuse branches reuses the half-vector's random
      // numbers (material/disney.cpp:113-130), so its Pdf (:186-220) is not
      // that density
      // and the plain estimator converges to E[f cos / Pdf] rather than to
      // the integral of f cos; mixing another strategy in would change the
      // value the render converges to. Disney vertices therefore sample
      // the BSDF alone (alpha 0), exactly as the plain path does.
      const float alpha = (GUIDED && bs.m.type != AJG_MAT_DISNEY) ? G.alpha : 0.f;
      // The env light's two estimates measure different integrands where
      // the shading normal differs from the geometric one: the light side
      // takes Eval and |wi . normalize(ns)| (helper.h:120-131), the BSDF
      // side the sample's f, which carries |ns . wi| / |ng . wi|
      // (aggregate_bsdf.cpp:82-89), times |ns . wi| (helper.h:228-231).
      // The plain render converges to the sum of the two weighted by the
      // BSDF-pdf power heuristic, so guided weights may differ from plain
      // only where the two integrands coincide: ns equal to ng. Their
      // ratio is |ns|^2 |ns^ . wi| / |ng . wi|, within 1 + 2e-5 + theta tan(beta)
      // of 1 for ns at angle theta from ng; the test below admits
      // theta <= 1e-5 and |ns|^2 within 2e-5 of 1 (quads and cubes, whose
      // ns is ng, and flat-shaded meshes up to rounding): there the two
      // integrands differ by at most 2e-5 + 1e-5 tan(beta) of their value
      // at angle beta from ng, so the weights' choice moves the converged
      // value by no more than that.
      bool mix_w = false;
      if (GUIDED && alpha > 0.f) {
        const float n2 = dot(it.ns, it.ns);
        const f3 cx = cross(it.ns, it.ng);
        mix_w = dot(cx, cx) <= 1e-10f * n2 && fabsf(n2 - 1.f) <= 2e-5f;
      }
      // emitted radiance at the first hit (pt.cpp:188-194); Trace adds
      // it at every hit (pt.cpp:289-293)
      if ((b == 0 || !P.use_mis) && it.light && dot(it.ng, it.wr) > 0)
        L = add(L, mul(beta, sc.lights[it.light - 1].radiance));
      uint32_t pf2 = kPendActive;
      st_f3(ps.pend_L, slot, L);
      st_f3(ps.pend_beta, slot, beta);
      ps.pend_work[slot] = work;
      ps.pend_sample[slot] = sample;
      ps.pend_bank[slot] = ps.bank[slot];
      if (BATCH) ps.pend_job[slot] = j;
      if (train) {
        st_f3(G.pend_pos, slot, it.pos);
        st_f3(ps.pend_ns, slot, it.ns);
      }
      // NEE for each light (MISSampleLight, helper.h:100-133); Trace
      // samples no light (its direct-light loop is empty, pt.cpp:296-301)
      const uint32_t n_nee = P.use_mis ? sc.n_lights : 0u;
      for (uint32_t k = 0; k < n_nee; ++k) {
        const DevLight &l = sc.lights[k];
        LightSample lsmp;
        if (l.type == AJG_LIGHT_ENV) {
          const float first = pcg_next_float(rng);
          const float second = pcg_next_float(rng);
          lsmp = env_sample(l.env, it.pos, second, first, sc.radius);
        } else if (l.type == kLightArea) {
          f3 sam;
          sam.x = pcg_next_float(rng);
          sam.y = pcg_next_float(rng);
          sam.z = pcg_next_float(rng);
          lsmp = area_sample(l, it.pos, sam);
        } else {
          const float u0 = pcg_next_float(rng);
          const float u1 = pcg_next_float(rng);
          lsmp = dir_sample(l, it.pos, u0, u1, sc.radius);
        }
        bool emit = false;
        DevRay sr;
        if (!(is_black(lsmp.radiance) || lsmp.pdf <= 0.f)) {
          const float len = length(sub(lsmp.pos, it.pos)) - AJG_EPS;
          if (!(len < AJG_EPS)) {
            const f3 to_light = normalize(sub(lsmp.pos, it.pos));
            const f3 bf = bsdf_eval_all(bs, to_light, it.wr);
            if (!is_black(bf)) {
              // MISSampleAreaLight takes the shading normal as it is,
              // the other lights normalise it (helper.h:64, 92, 125)
              const f3 nrm = normalize(it.ns);
              const f3 f = scale(mul(lsmp.radiance, bf),
                                 fabsf(dot(to_light, l.type == kLightArea ? it.ns : nrm)));
              float bpdf = bsdf_pdf_all(bs, to_light, it.wr);
              // guided: the env light's MIS partner is the mixture density
              // where ns equals ng (see mix_w); a delta light is never
              // reached by directional sampling, so its weight keeps the
              // BSDF pdf, as in the plain path
              if (mix_w && l.type == AJG_LIGHT_ENV)
                bpdf = alpha * gmix_pdf(gm, to_light) + (1.f - alpha) * bpdf;
              const float wl = power_heuristic(lsmp.pdf, bpdf);
              const f3 c = mul(beta, scale(divs(f, lsmp.pdf), wl));
              st_f3(ps.pend_c, pl_idx(ps, slot, k), c);
              if (train)
                G.pend_tw[(size_t)slot * kMaxLights + k] =
                    l.type != AJG_LIGHT_DIRECTIONAL
                        ? lum_f(lsmp.radiance) * fabsf(dot(to_light, nrm)) / lsmp.pdf * wl
                        : 0.f;
              sr = make_ray(it.pos, to_light, AJG_EPS, 0.99f * len);
              emit = true;
            }
          }
        }
        const uint32_t si = q_append(emit, q.a_count_out, q.a_len, stripe);
        if (emit) {
          store_ray(q.a_rays_out, si, sr);
          if (BATCH) B.a_job_out[si] = (uint16_t)j;
          ps.pend_shadow[pl_idx(ps, slot, k)] = si;
          pf2 |= kPendLight0 << k;
        }
      }
      // Two independent draws of the BSDF sampler (of the guide mixture
      // when guided): the MIS-BSDF sample (MISSampleBsdf, helper.h:209-246),
      // then the continuation (pt.cpp:215-223). One loop over the two keeps
      // a single inlined copy of the sampler in the kernel.
      // Where ns differs from ng the MIS-BSDF estimate carries the factor
      // |ns . wi| / |ng . wi| (aggregate_bsdf.cpp:82-89), unbounded for
      // directions grazing ng; the plain sampler draws those rarely, a
      // guide may not, so there the MIS-BSDF direction comes from the BSDF
      // alone (alpha 0, with the plain weights) and only the continuation,
      // whose weight stays bounded, is guided.
      GSample g_mis, g_cont;
      g_mis.ok = false;
#pragma unroll 1
      for (int k = P.use_mis ? 0 : 1; k < 2; ++k) {
        GSample g;
        if (GUIDED) {
          g = guided_sample(bs, it.wr, gm, (k == 0 && !mix_w) ? 0.f : alpha, rng);
        } else {
          f3 sam;
          sam.x = pcg_next_float(rng);
          sam.y = pcg_next_float(rng);
          sam.z = pcg_next_float(rng);
          const BsdfSample s = bsdf_sample_all(bs, it.wr, sam);
          g.dir = s.dir;
          g.f = s.f;
          g.pdf = s.pdf;
          g.pb = s.pdf;
          g.ok = s.valid;
        }
        if (k == 0)
          g_mis = g;
        else
          g_cont = g;
      }
      // MIS-BSDF ray
      {
        bool emit = false;
        DevRay mr;
        const bool lit = sc.env_index >= 0 || sc.n_area > 0;
        if (GUIDED) {
          if (g_mis.ok && lit) {
            mr = make_ray(eps_offset(it, g_mis.dir), g_mis.dir, AJG_EPS, __int_as_float(0x7f800000));
            st_f3(ps.pend_f, slot, g_mis.f);
            ps.pend_pdf[slot] = g_mis.pdf;
            ps.pend_pb[slot] = g_mis.pb;
            st_f3(ps.pend_ns, slot, it.ns);
            emit = true;
          }
        } else {
          if (g_mis.ok && !(is_black(g_mis.f) || g_mis.pdf == 0.f) && lit) {
            const f3 dn = normalize(g_mis.dir);
            mr = make_ray(eps_offset(it, dn), dn, AJG_EPS, __int_as_float(0x7f800000));
            st_f3(ps.pend_f, slot, g_mis.f);
            ps.pend_pdf[slot] = g_mis.pdf;
            st_f3(ps.pend_ns, slot, it.ns);
            emit = true;
          }
        }
        // with area lights: a closest-hit query (the emitter it hits)
        const bool to_c = emit && sc.n_area > 0;
        const uint32_t mc = q_append(to_c, q.c_count_out, q.c_len, stripe);
        const uint32_t ma = q_append(emit && !to_c, q.a_count_out, q.a_len, stripe);
        if (to_c) {
          store_ray(q.c_rays_out, mc, mr);
          if (BATCH) B.c_job_out[mc] = (uint16_t)j;
          ps.pend_mis[slot] = mc;
          pf2 |= kPendMis | kPendMisClosest;
        } else if (emit) {
          store_ray(q.a_rays_out, ma, mr);
          if (BATCH) B.a_job_out[ma] = (uint16_t)j;
          ps.pend_mis[slot] = ma;
          pf2 |= kPendMis;
        }
        if (mix_w) pf2 |= kPendMisMix;
      }
      // continuation (pt.cpp:215-223) and Russian roulette of the next
      // bounce (pt.cpp:107-110)
      bool cont;
      DevRay nr;
      float cw = 0.f;
      {
        const f3 dir = g_cont.dir, fv = g_cont.f;
        const float pdf = g_cont.pdf;
        cont = g_cont.ok;
        if (!GUIDED) {
          cont = cont && !(is_black(fv) || pdf < 0.f);
          if ((P.flags & AJG_RENDER_FINITE_PDF) && pdf == 0.f) cont = false;
        }
        if (cont) {
          beta = mul(beta, divs(scale(fv, fabsf(dot(it.ng, dir))), pdf));
          nr = make_ray(eps_offset(it, dir), dir, AJG_EPS, __int_as_float(0x7f800000));
          if (train) cw = fabsf(dot(normalize(it.ns), dir)) / pdf;
          const int nb = (int)b + 1;
          if (nb >= P.max_bounces) {
            cont = false;
          } else if (nb > P.min_bounces) {
            if (pcg_next_float(rng) > P.cont_prob)
              cont = false;
            else
              beta = divs(beta, P.cont_prob);
          }
        }
      }
      const uint32_t ni = q_append(cont, q.c_count_out, q.c_len, stripe);
      if (cont) {
        store_ray(q.c_rays_out, ni, nr);
        if (BATCH) B.c_job_out[ni] = (uint16_t)j;
        ei_next = ni;
        ps.bounce[slot] = b + 1;
        st_f3(ps.beta, slot, beta);
        if (train) {
          const uint32_t hn = G.hist_n[hslot];
          if (hn < G.hist_depth) {
            float *hr = G.hist + ((size_t)hslot * G.hist_depth + hn) * kGHist;
            st3(hr, it.pos);
            st3(hr + 3, normalize(it.ns));
            st3(hr + 6, nr.d);
            hr[9] = cw;
            st3(hr + 13, beta);
            G.hist_n[hslot] = hn + 1;
            pf2 |= kPendHist;
          }
        }
      } else {
        pf2 |= kPendEnded;
        if (last) pf2 |= kPendLast;
        need_new = true;
      }
      ps.pend_flags[slot] = pf2;
    }
  }
  if (ps.pend_flags[slot] & kPendActive) live = true;
}

// Plain and batched shading run three waves per SIMD (at most 168
// registers): measured faster than the compiler's own budget (two waves)
// and than four (128 registers, with spills). The guided variant keeps the
// compiler's budget (at three waves it would spill more than twice as much).
template <bool GUIDED, bool BATCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GUIDED ? 1 : 3)))
k_shade(DevScene sc0, RenderParams P0, PathState ps, Queues q, GuideParams G, BatchParams B) {
  // The launch covers the largest possible list (every slot); blocks past
  // the list's end leave at once. Thread v of the launch shades the slot at
  // dense index v of the live-slot list. Slots without work, pending
  // light or rays are not in the list: late in a render most slots are
  // idle, and shading by slot would keep their lanes in every wave.
  const uint32_t total = *q.lv_total;
  if (blockIdx.x * 256 >= total) return;
  // guided: the guide mixtures are stored by dense index of the closest queue
  __shared__ uint32_t s_pre[GUIDED ? kQStripes + 1 : 1];
  __shared__ uint32_t s_lv[kQStripes + 1];
  if (GUIDED) q_prefix(q.c_in, s_pre);
  q_prefix(q.lv_q, s_lv);
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  const bool valid = v < total;
  const uint32_t le = valid ? q_entry(s_lv, q.lv_q.R, v) : 0u;
  const uint32_t slot = valid ? q.lv_in[le] : 0u;
  const uint32_t ei = valid ? q.lv_ei_in[le] : kNone;
  const uint32_t stripe = (v >> 6) % kQStripes;  // queue stripe of this wave
  const bool train = GUIDED && valid && G.train_stride && slot % G.train_stride == 0;
  const uint32_t hslot = train ? slot / G.train_stride : 0u;
  bool live = false;
  Pcg32 rng;
  rng.state = valid ? ps.rng[slot] : 0ull;
  bool need_new = false;

  uint32_t ei_next = kNone;
  if (valid)
    shade_slot<GUIDED, BATCH>(sc0, P0, ps, q, G, B, slot, stripe, train, hslot, s_pre, rng, ei, ei_next,
                              live, need_new);

  // ---- 3. next sample / ray generation (wave-uniform control flow)
  next_sample<BATCH>(ps, q, P0, B, slot, valid && need_new, rng, live, stripe, ei_next);
  if (valid) ps.rng[slot] = rng.state;
  // the slots left live (a path, pending light or a sample to start) are
  // the next iteration's list, each with its extension ray (the list
  // carries it, so shading reads it with the slot instead of after it)
  const uint32_t li = q_append(valid && live, q.lv_count_out, q.lv_len, stripe);
  if (valid && live) {
    q.lv_out[li] = slot;
    q.lv_ei_out[li] = ei_next;
  }
  const bool any_live = __ballot(valid && live) != 0ull;
  if ((threadIdx.x & 63) == 0 && any_live) *q.live = 1u;
}

// Step 1 of shading, finishing the previous bounce's direct lighting for
// every slot of the live-slot list, as a kernel of its own: it needs about
// 50 registers and runs eight waves per SIMD, where inside k_shade it ran
// at k_shade's three (heavy mesh shading 48.7 -> 45.7 ms per render, both
// kernels). k_shade runs steps 2 and 3 after it on the same stream.
template <bool GUIDED, bool BATCH>
__global__ void __launch_bounds__(256)
k_shade_pend(DevScene sc0, RenderParams P0, PathState ps, Queues q, GuideParams G, BatchParams B) {
  const uint32_t total = *q.lv_total;
  if (blockIdx.x * 256 >= total) return;
  __shared__ uint32_t s_lv[kQStripes + 1];
  q_prefix(q.lv_q, s_lv);
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v < total) {
    const uint32_t slot = q.lv_in[q_entry(s_lv, q.lv_q.R, v)];
    const bool train = GUIDED && G.train_stride && slot % G.train_stride == 0;
    const uint32_t hslot = train ? slot / G.train_stride : 0u;
    shade_pend<GUIDED, BATCH>(sc0, P0, ps, q, G, B, slot, train, hslot);
  }
}

// AJG_RENDER_INLINE: the whole render as one kernel, one thread per work
// item tracing its samples' paths to their ends with single-ray traversal
// (no queues, no per-bounce launches). Each step is the wavefront's own
// arithmetic in the wavefront's order -- the same random draws, the same
// hits (trace_closest / trace_any visit nodes as the queue kernels do), the
// direct light summed light by light and then the MIS-BSDF term, the same
// Kahan accumulation -- so films and per-sample radiance are bitwise those
// of the wavefront render. For renders with few paths in flight (per-pixel
// streams on small images), where each wavefront iteration costs its
// kernels' latency floors. One-level scenes, plain (unguided) sampling.
// A tree of at most kInlineLdsTris triangles (its nodes and leaf triangles,
// 112 bytes per triangle) is copied into LDS first and traversed there.
static constexpr uint32_t kInlineLdsTris = 256;

template <int SPILL>
__global__ void __launch_bounds__(256)
k_pt_inline(DevScene sc, RenderParams P, float *partial, uint32_t *partial_n, float *per_sample,
            unsigned long long *totals, uint32_t lds_tris) {
  __shared__ int s_node[kLdsStack * 256];
  __shared__ float s_entry[kLdsStack * 256];
  extern __shared__ uint4 s_tree[];
  int *ln = s_node + threadIdx.x;
  float *le = s_entry + threadIdx.x;
  if (lds_tris) {
    // nodes (lds_tris - 1) x 64 B, then leaf triangles lds_tris x 48 B
    const uint32_t nw = (lds_tris - 1) * 4, nl = lds_tris * 3;
    const uint4 *gw = reinterpret_cast<const uint4 *>(sc.bvh.wide);
    const uint4 *gl = reinterpret_cast<const uint4 *>(sc.bvh.leaf_tris);
    for (uint32_t i = threadIdx.x; i < nw + nl; i += 256) s_tree[i] = i < nw ? gw[i] : gl[i - nw];
    __syncthreads();
    sc.bvh.wide = reinterpret_cast<const WideNode *>(s_tree);
    sc.bvh.leaf_tris = reinterpret_cast<const float4 *>(s_tree + nw);
  }
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  uint32_t n_closest = 0, n_any = 0;
  uint32_t nn = 0, nt = 0;
  if (w < P.total_work) {
    const float kInf = __int_as_float(0x7f800000);
    const uint32_t chunk = w / P.n_pix;
    const uint32_t p = pixel_of(w - chunk * P.n_pix, P.tile, P.tiles_x, P.Wp);
    const uint32_t x = p % P.Wp, y = p / P.Wp;
    const uint32_t s0 = chunk * P.chunk, s1 = min(P.spp, s0 + P.chunk);
    Pcg32 rng;
    rng.state = 0ull;
    if (P.seed_mode == 0) pcg_seed(rng, (uint64_t)(int64_t)(int)(x + y * P.W));
    float a[12];
    for (int k = 0; k < 12; ++k) a[k] = 0.f;
    uint32_t na = 0;
    const bool lit = sc.env_index >= 0 || sc.n_area > 0;
    // Each lane moves on to its next sample as soon as its path ends (one
    // vertex per loop pass), so a wave waits for the longest sum of path
    // lengths among its lanes, not for the longest path of every sample.
    // A lane's random draws stay in their sequential order.
    auto finish = [&](uint32_t smp, f3 Ls) {
      kahan_add(a[0], a[3], Ls.x);
      kahan_add(a[1], a[4], Ls.y);
      kahan_add(a[2], a[5], Ls.z);
      kahan_add(a[6], a[9], Ls.x * Ls.x);
      kahan_add(a[7], a[10], Ls.y * Ls.y);
      kahan_add(a[8], a[11], Ls.z * Ls.z);
      ++na;
      if (per_sample) {
        float *o = per_sample + 3 * ((size_t)p * P.spp + smp);
        o[0] = Ls.x;
        o[1] = Ls.y;
        o[2] = Ls.z;
      }
    };
    uint32_t sample = s0, b = 0;
    bool start = true;
    DevRay ray;
    f3 L = mk3(0.f, 0.f, 0.f), beta = mk3(1.f, 1.f, 1.f);
    for (;;) {
      if (start) {
        if (sample >= s1) break;
        if (P.seed_mode != 0) pcg_seed(rng, sample_seed(P.seed_base, p, sample));
        // TiledIntegrator::Render per-sample draws (pt.cpp:58-61)
        const float u = (float)x + pcg_next_float(rng);
        const float v = (float)y + pcg_next_float(rng);
        const float l0 = pcg_next_float(rng);
        const float l1 = pcg_next_float(rng);
        ray = camera_ray(P.cam, u, v, l0, l1);
        L = mk3(0.f, 0.f, 0.f);
        beta = mk3(1.f, 1.f, 1.f);
        b = 0;
        bool alive = true;
        if (0 > P.min_bounces) {
          if (pcg_next_float(rng) > P.cont_prob)
            alive = false;
          else
            beta = divs(mk3(1.f, 1.f, 1.f), P.cont_prob);
        }
        if (P.max_bounces <= 0) alive = false;
        if (!alive) {  // a zero-radiance sample
          finish(sample, mk3(0.f, 0.f, 0.f));
          ++sample;
          continue;
        }
        start = false;
      }
      ++n_closest;
      const THit h = trace_closest<SPILL, false>(sc.bvh, make_tray(ray.o, ray.d, ray.tmin, ray.tmax),
                                                 ln, le, &nn, &nt);
      if (h.slot < 0) {
        if ((b == 0 || !P.use_mis) && sc.env_index >= 0)
          L = add(L, env_radiance(sc.lights[sc.env_index].env, ray.d));
        finish(sample, L);
        ++sample;
        start = true;
        continue;
      }
      const Inct it = post_intersect(sc, ray, h.t, h.slot, h.u, h.v, 0u);
      const Bsdf bs = make_bsdf(sc.mats[it.material], it.ng, it.ns);
      if ((b == 0 || !P.use_mis) && it.light && dot(it.ng, it.wr) > 0)
        L = add(L, mul(beta, sc.lights[it.light - 1].radiance));
      // NEE (MISSampleLight, helper.h:100-133), visibility traced at once
      f3 direct = mk3(0.f, 0.f, 0.f);
      const uint32_t n_nee = P.use_mis ? sc.n_lights : 0u;
      for (uint32_t k = 0; k < sc.n_lights; ++k) {
        f3 c = mk3(0.f, 0.f, 0.f);
        if (k < n_nee) {
          const DevLight &l = sc.lights[k];
          LightSample lsmp;
          if (l.type == AJG_LIGHT_ENV) {
            const float first = pcg_next_float(rng);
            const float second = pcg_next_float(rng);
            lsmp = env_sample(l.env, it.pos, second, first, sc.radius);
          } else if (l.type == kLightArea) {
            f3 sam;
            sam.x = pcg_next_float(rng);
            sam.y = pcg_next_float(rng);
            sam.z = pcg_next_float(rng);
            lsmp = area_sample(l, it.pos, sam);
          } else {
            const float u0 = pcg_next_float(rng);
            const float u1 = pcg_next_float(rng);
            lsmp = dir_sample(l, it.pos, u0, u1, sc.radius);
          }
          if (!(is_black(lsmp.radiance) || lsmp.pdf <= 0.f)) {
            const float len = length(sub(lsmp.pos, it.pos)) - AJG_EPS;
            if (!(len < AJG_EPS)) {
              const f3 to_light = normalize(sub(lsmp.pos, it.pos));
              const f3 bf = bsdf_eval_all(bs, to_light, it.wr);
              if (!is_black(bf)) {
                const f3 nrm = normalize(it.ns);
                const f3 f = scale(mul(lsmp.radiance, bf),
                                   fabsf(dot(to_light, l.type == kLightArea ? it.ns : nrm)));
                const float bpdf = bsdf_pdf_all(bs, to_light, it.wr);
                const float wl = power_heuristic(lsmp.pdf, bpdf);
                const f3 cl = mul(beta, scale(divs(f, lsmp.pdf), wl));
                const DevRay sr = make_ray(it.pos, to_light, AJG_EPS, 0.99f * len);
                ++n_any;
                if (!trace_any<SPILL>(sc.bvh, make_tray(sr.o, sr.d, sr.tmin, sr.tmax), ln, le))
                  c = cl;
              }
            }
          }
        }
        direct = add(direct, c);
      }
      // the BSDF sampler's two draws: MIS-BSDF (helper.h:209-246), then
      // the continuation (pt.cpp:215-223)
      BsdfSample g_mis, g_cont;
      g_mis.valid = false;
      for (int k = P.use_mis ? 0 : 1; k < 2; ++k) {
        f3 sam;
        sam.x = pcg_next_float(rng);
        sam.y = pcg_next_float(rng);
        sam.z = pcg_next_float(rng);
        const BsdfSample smp = bsdf_sample_all(bs, it.wr, sam);
        if (k == 0)
          g_mis = smp;
        else
          g_cont = smp;
      }
      f3 mis = mk3(0.f, 0.f, 0.f);
      if (g_mis.valid && !(is_black(g_mis.f) || g_mis.pdf == 0.f) && lit) {
        const f3 dn = normalize(g_mis.dir);
        const DevRay mr = make_ray(eps_offset(it, dn), dn, AJG_EPS, kInf);
        bool escaped;
        if (sc.n_area > 0) {
          ++n_closest;
          const THit mh = trace_closest<SPILL, false>(sc.bvh, make_tray(mr.o, mr.d, mr.tmin, mr.tmax),
                                                      ln, le, &nn, &nt);
          escaped = mh.slot < 0;
          if (!escaped) {
            const Inct hi = post_intersect(sc, mr, mh.t, mh.slot, mh.u, mh.v, 0u);
            if (hi.light && dot(hi.ng, hi.wr) > 0) {  // AreaLight::Radiance
              const DevLight &l = sc.lights[hi.light - 1];
              if (!is_black(l.radiance)) {
                const f3 f = scale(mul(l.radiance, g_mis.f), fabsf(dot(it.ns, mr.d)));
                const float lpdf = area_pdf(l, mr.o, hi.pos, hi.ng);
                mis = scale(divs(f, g_mis.pdf), power_heuristic(g_mis.pdf, lpdf));
              }
            }
          }
        } else {
          ++n_any;
          escaped = !trace_any<SPILL>(sc.bvh, make_tray(mr.o, mr.d, mr.tmin, mr.tmax), ln, le);
        }
        if (escaped && sc.env_index >= 0) {
          const DevEnv &e = sc.lights[sc.env_index].env;
          const f3 lev = env_radiance(e, mr.d);
          if (!is_black(lev)) {
            const f3 f = scale(mul(lev, g_mis.f), fabsf(dot(it.ns, mr.d)));
            mis = scale(divs(f, g_mis.pdf), power_heuristic(g_mis.pdf, env_pdf(e, mr.d)));
          }
        }
      }
      direct = add(direct, mul(beta, mis));
      // continuation and Russian roulette of the next bounce (pt.cpp:
      // 107-110, 215-223); beta above is the vertex's own throughput
      bool cont = g_cont.valid && !(is_black(g_cont.f) || g_cont.pdf < 0.f);
      if ((P.flags & AJG_RENDER_FINITE_PDF) && g_cont.pdf == 0.f) cont = false;
      f3 nbeta = beta;
      if (cont) {
        const f3 dir = g_cont.dir;
        nbeta = mul(beta, divs(scale(g_cont.f, fabsf(dot(it.ng, dir))), g_cont.pdf));
        ray = make_ray(eps_offset(it, dir), dir, AJG_EPS, kInf);
        const int nb = (int)b + 1;
        if (nb >= P.max_bounces) {
          cont = false;
        } else if (nb > P.min_bounces) {
          if (pcg_next_float(rng) > P.cont_prob)
            cont = false;
          else
            nbeta = divs(nbeta, P.cont_prob);
        }
      }
      L = add(L, scale(direct, 1.f));
      if (!cont) {
        finish(sample, L);
        ++sample;
        start = true;
        continue;
      }
      beta = nbeta;
      ++b;
    }
    const size_t cell = (size_t)chunk * P.n_pix + p;  // chunk-major, pixel-minor
    float *dst = partial + cell * 6;
    dst[0] = a[0] - a[3];
    dst[1] = a[1] - a[4];
    dst[2] = a[2] - a[5];
    dst[3] = a[6] - a[9];
    dst[4] = a[7] - a[10];
    dst[5] = a[8] - a[11];
    partial_n[cell] = na;
  }
  wave_count_add(n_closest, totals);
  wave_count_add(n_any, totals + 1);
}

// Guided rendering: position and unit shading normal of each closest-queue
// hit (the guide network's input), by dense queue index; misses get a zero
// record. Writes the queue's total to *total (the guide's batch size).
__global__ void __launch_bounds__(256)
k_hit_points(DevScene sc, const DevRay *rays, const float4 *hits, const uint32_t *hit_inst,
             QIn qin, float *pts, uint32_t *total) {
  __shared__ uint32_t s_pre[kQStripes + 1];
  const uint32_t n = q_prefix(qin, s_pre);
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n;
  for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
    const uint32_t i = q_entry(s_pre, qin.R, v);
    const float4 h = hits[i];
    const int ls = __float_as_int(h.y);
    f3 p = mk3(0.f, 0.f, 0.f), nn = mk3(0.f, 0.f, 1.f);
    if (ls >= 0) {
      const DevRay r = load_ray(rays, i);
      const Inct it = post_intersect(sc, r, h.x, ls, h.z, h.w, sc.inst ? hit_inst[i] : 0u);
      p = it.pos;
      nn = normalize(it.ns);
    }
    float *o = pts + (size_t)v * 6;
    st3(o, p);
    st3(o + 3, nn);
  }
}

// Guided rendering: clear the training-record counter after the host has
// consumed the records, adding the count to the running total.
__global__ void k_rec_reset(uint32_t *rec_count, unsigned long long *total) {
  if (threadIdx.x == 0) {
    *total += *rec_count;
    *rec_count = 0;
  }
}

// Per-chunk partial sums -> film (Kahan over chunks in chunk order).
__global__ void __launch_bounds__(256)
k_reduce_chunks(const float *partial, const uint32_t *partial_n, uint32_t n_pix,
                uint32_t n_chunks, float *sum, float *sum_sq, uint32_t *count) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_pix) return;
  float s[6] = {0, 0, 0, 0, 0, 0}, c[6] = {0, 0, 0, 0, 0, 0};
  uint32_t n = 0;
  for (uint32_t k = 0; k < n_chunks; ++k) {
    const float *src = partial + ((size_t)k * n_pix + p) * 6;
    for (int j = 0; j < 6; ++j) kahan_add(s[j], c[j], src[j]);
    n += partial_n[(size_t)k * n_pix + p];
  }
  for (int j = 0; j < 3; ++j) {
    sum[3 * (size_t)p + j] = s[j] - c[j];
    sum_sq[3 * (size_t)p + j] = s[3 + j] - c[3 + j];
  }
  count[p] = n;
}

__global__ void __launch_bounds__(256)
k_unpack_single(const float *partial, const uint32_t *partial_n, uint32_t n_pix, float *sum,
                float *sum_sq, uint32_t *count) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_pix) return;
  for (int j = 0; j < 3; ++j) {
    sum[3 * (size_t)p + j] = partial[(size_t)p * 6 + j];
    sum_sq[3 * (size_t)p + j] = partial[(size_t)p * 6 + 3 + j];
  }
  count[p] = partial_n[p];
}

// Film resolve (film.cpp:27-63, image.h:38-48): each traced pixel's sums
// splat with weight 1 into {p-1, p} x {p-1, p}, clamped to the film edge.
__global__ void __launch_bounds__(256)
k_resolve(uint32_t w, uint32_t h, uint32_t Wp, uint32_t Hp, const float *sum,
          const uint32_t *count, float *rgb) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= w * h) return;
  const int qx = (int)(q % w), qy = (int)(q / w);
  // A traced pixel p adds into clamp(p - 1) and clamp(p) per axis; the
  // multiplicity m(p -> q) counts how many of the two land on q. Traced
  // pixels beyond the film edge (whole tiles) all clamp into the last row or
  // column.
  const int x_hi = qx == (int)w - 1 ? (int)Wp - 1 : qx + 1;
  const int y_hi = qy == (int)h - 1 ? (int)Hp - 1 : qy + 1;
  double acc[3] = {0, 0, 0}, wt = 0;
  for (int py = qy; py <= y_hi; ++py) {
    int my = 0;
    for (int dy = -1; dy <= 0; ++dy) my += (clampi(py + dy, 0, (int)h - 1) == qy);
    if (!my) continue;
    for (int px = qx; px <= x_hi; ++px) {
      int mx = 0;
      for (int dx = -1; dx <= 0; ++dx) mx += (clampi(px + dx, 0, (int)w - 1) == qx);
      if (!mx) continue;
      const size_t p = (size_t)px + (size_t)py * Wp;
      const double m = (double)(mx * my);
      for (int k = 0; k < 3; ++k) acc[k] += m * sum[3 * p + k];
      wt += m * count[p];
    }
  }
  for (int k = 0; k < 3; ++k) rgb[3 * (size_t)q + k] = wt != 0 ? (float)(acc[k] / wt) : 0.f;
}

// Filtered resolve (film.cpp:27-63 with any Filter): a traced pixel p adds
// its sums into targets j = p + o, o in [omin, omax] per axis, with weight
// w[o - omin] (the filter at offset o + 0.5), clamped to the film. Output x
// gathers from every p whose footprint clamps onto x; per axis that weight is
// the sum of w over the offsets of p that land on x.
struct FilterAxis {
  int omin, omax;
  float w[16];
};

__device__ __forceinline__ float axis_weight(const FilterAxis &f, int p, int x, int n) {
  float a = 0.f;
  for (int o = f.omin; o <= f.omax; ++o)
    if (clampi(p + o, 0, n - 1) == x) a += f.w[o - f.omin];
  return a;
}

__global__ void __launch_bounds__(256)
k_resolve_filter(uint32_t w, uint32_t h, uint32_t Wp, uint32_t Hp, FilterAxis fx, FilterAxis fy,
                 const float *sum, const uint32_t *count, float *rgb) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= w * h) return;
  const int qx = (int)(q % w), qy = (int)(q / w);
  const int x_lo = max(0, qx - fx.omax);
  const int x_hi = qx == (int)w - 1 ? (int)Wp - 1 : min((int)Wp - 1, qx - fx.omin);
  const int y_lo = max(0, qy - fy.omax);
  const int y_hi = qy == (int)h - 1 ? (int)Hp - 1 : min((int)Hp - 1, qy - fy.omin);
  double acc[3] = {0, 0, 0}, wt = 0;
  for (int py = y_lo; py <= y_hi; ++py) {
    const float ay = axis_weight(fy, py, qy, (int)h);
    if (ay == 0.f) continue;
    for (int px = x_lo; px <= x_hi; ++px) {
      const float ax = axis_weight(fx, px, qx, (int)w);
      if (ax == 0.f) continue;
      const double k = (double)ax * (double)ay;
      const size_t p = (size_t)px + (size_t)py * Wp;
      for (int c = 0; c < 3; ++c) acc[c] += k * sum[3 * p + c];
      wt += k * count[p];
    }
  }
  for (int c = 0; c < 3; ++c) rgb[3 * (size_t)q + c] = wt != 0 ? (float)(acc[c] / wt) : 0.f;
}

// ImageWriter::Write pixel path (image.cpp:32-62, image.h:73-84,
// color.h:54-64): sRGB encode, clamp to [0, 1], round(v * 255.5) in double,
// rows bottom to top when flipped; NaN components become 0.
__global__ void __launch_bounds__(256)
k_srgb8(uint32_t w, uint32_t h, const float *rgb, int flip, uint8_t *out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= w * h) return;
  const uint32_t x = i % w, y = i / w;
  const uint32_t sy = flip ? h - 1 - y : y;
  for (int k = 0; k < 3; ++k) {
    const float v = rgb[3 * ((size_t)x + (size_t)sy * w) + k];
    float sv = v > 0.0031308f ? 1.055f * powf(v, 1.0f / 2.4f) - 0.055f : v * 12.92f;
    sv = sv < 0.f ? 0.f : (sv > 1.f ? 1.f : sv);
    if (sv != sv) sv = 0.f;
    const double r = round((double)sv * 255.5);
    out[3 * (size_t)i + k] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
  }
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Device buffers of a render: path state, double-buffered ray queues,
// counters and per-work-item partial sums. Kept on the scene and reused by
// later renders that fit.
struct RenderBuffers {
  uint32_t slots = 0;
  uint64_t total = 0;
  std::vector<void *> bufs;
  PathState ps;
  DevRay *c_rays[2] = {nullptr, nullptr};
  DevRay *a_rays[2] = {nullptr, nullptr};
  float4 *c_hits = nullptr;
  uint32_t *c_inst = nullptr;  // instance of each closest hit (instanced scenes)
  uint8_t *a_occ = nullptr;
  uint32_t *counters = nullptr;
  uint32_t *qctr = nullptr;  // striped queue counters: closest 0, 1, shadow 0, 1
  uint32_t *lv[2] = {nullptr, nullptr};  // live-slot lists (kQStripes * lv_stripe_len(S))
  uint32_t *lv_ei[2] = {nullptr, nullptr};  // their entries' extension rays
  float *partial = nullptr;
  uint32_t *partial_n = nullptr;
  // shadow rays are traced on a second stream, concurrently with the
  // closest-hit queue (fork after shade, join before the next shade)
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void release() {
    for (void *p : bufs) (void)hipFree(p);
    bufs.clear();
    slots = 0;
    total = 0;
  }
  ~RenderBuffers() {
    release();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (aux) (void)hipStreamDestroy(aux);
  }
};

static int ensure_aux_stream(ajg_ctx *ctx, RenderBuffers *rb) {
  if (rb->aux) return AJG_OK;
  hipError_t e = hipStreamCreateWithFlags(&rb->aux, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&rb->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&rb->ev_join, hipEventDisableTiming);
  return e == hipSuccess ? AJG_OK : hip_fail(ctx, e, "render streams");
}

// Guided-rendering buffers (kept on the scene next to RenderBuffers).
struct GuideBuffers {
  uint32_t slots = 0, train_slots = 0, depth = 0;
  std::vector<void *> bufs;
  float *pts = nullptr, *pend_pos = nullptr, *pend_tw = nullptr;
  _Float16 *mix = nullptr;
  float *hist = nullptr;
  uint32_t *hist_n = nullptr;
  void release() {
    for (void *p : bufs) (void)hipFree(p);
    bufs.clear();
    slots = train_slots = depth = 0;
  }
  ~GuideBuffers() { release(); }
};

static int ensure_guide_buffers(ajg_ctx *ctx, GuideBuffers *gb, uint32_t S, uint32_t T,
                                uint32_t D) {
  if (gb->slots >= S && gb->train_slots >= T && gb->depth >= D && !gb->bufs.empty())
    return AJG_OK;
  gb->release();
  auto al = [&](void **p, size_t bytes, const char *what) {
    int r = dev_alloc(ctx, p, bytes, what);
    if (r == AJG_OK) gb->bufs.push_back(*p);
    return r;
  };
  int st = 0;
  if ((st = al((void **)&gb->pts, 24 * cq_capacity(S), "guide hit points")) ||
      (st = al((void **)&gb->mix, 2 * (size_t)kGMix * cq_capacity(S), "guide mixtures")) ||
      (st = al((void **)&gb->pend_pos, 12 * (size_t)S, "guide pending positions")) ||
      (st = al((void **)&gb->pend_tw, 4 * (size_t)kMaxLights * S, "guide NEE weights")) ||
      (st = al((void **)&gb->hist, 4 * (size_t)kGHist * T * D, "guide vertex history")) ||
      (st = al((void **)&gb->hist_n, 4 * (size_t)T, "guide history sizes"))) {
    gb->release();
    return st;
  }
  gb->slots = S;
  gb->train_slots = T;
  gb->depth = D;
  return AJG_OK;
}

static constexpr size_t kCounterWords = 64;  // 256 bytes
// Entries per stripe of a live-slot list (one per slot).
static inline uint32_t lv_stripe_len(uint32_t S) { return stripe_len(S, 1); }

static int ensure_buffers(ajg_ctx *ctx, RenderBuffers *rb, uint32_t S, uint64_t total) {
  if (rb->slots >= S && rb->total >= total && !rb->bufs.empty()) return AJG_OK;
  rb->release();
  auto al = [&](void **p, size_t bytes, const char *what) {
    int r = dev_alloc(ctx, p, bytes, what);
    if (r == AJG_OK) rb->bufs.push_back(*p);
    return r;
  };
  PathState &ps = rb->ps;
  const size_t cq_cap = cq_capacity(S), aq_cap = aq_capacity(S);
  int st = 0;
  if ((st = al((void **)&ps.work, 4 * (size_t)S, "state")) || (st = al((void **)&ps.sample, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.sample_end, 4 * (size_t)S, "state")) || (st = al((void **)&ps.bounce, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.bank, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.rng, 8 * (size_t)S, "state")) || (st = al((void **)&ps.L, 12 * (size_t)S, "state")) ||
      (st = al((void **)&ps.beta, 12 * (size_t)S, "state")) || (st = al((void **)&ps.pend_flags, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_mis, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_shadow, 4 * (size_t)S * kMaxLights, "state")) ||
      (st = al((void **)&ps.pend_work, 4 * (size_t)S, "state")) || (st = al((void **)&ps.pend_sample, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_bank, 4 * (size_t)S, "state")) || (st = al((void **)&ps.pend_L, 12 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_beta, 12 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_c, 12 * (size_t)S * kMaxLights, "state")) ||
      (st = al((void **)&ps.pend_f, 12 * (size_t)S, "state")) || (st = al((void **)&ps.pend_pdf, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_pb, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_ns, 12 * (size_t)S, "state")) || (st = al((void **)&ps.acc, 4 * 24 * (size_t)S, "state")) ||
      (st = al((void **)&ps.acc_n, 8 * (size_t)S, "state")) ||
      (st = al((void **)&ps.job, 4 * (size_t)S, "state")) ||
      (st = al((void **)&ps.pend_job, 4 * (size_t)S, "state")) ||
      (st = al((void **)&rb->c_rays[0], sizeof(DevRay) * cq_cap, "closest queue")) ||
      (st = al((void **)&rb->c_rays[1], sizeof(DevRay) * cq_cap, "closest queue")) ||
      (st = al((void **)&rb->c_hits, sizeof(float4) * cq_cap, "closest hits")) ||
      (st = al((void **)&rb->c_inst, 4 * cq_cap, "closest hit instances")) ||
      (st = al((void **)&rb->a_rays[0], sizeof(DevRay) * aq_cap, "shadow queue")) ||
      (st = al((void **)&rb->a_rays[1], sizeof(DevRay) * aq_cap, "shadow queue")) ||
      (st = al((void **)&rb->a_occ, aq_cap, "shadow results")) ||
      (st = al((void **)&rb->lv[0], 4 * (size_t)kQStripes * lv_stripe_len(S), "live-slot list")) ||
      (st = al((void **)&rb->lv[1], 4 * (size_t)kQStripes * lv_stripe_len(S), "live-slot list")) ||
      (st = al((void **)&rb->lv_ei[0], 4 * (size_t)kQStripes * lv_stripe_len(S), "live-slot list")) ||
      (st = al((void **)&rb->lv_ei[1], 4 * (size_t)kQStripes * lv_stripe_len(S), "live-slot list")) ||
      (st = al((void **)&rb->counters, 4 * kCounterWords, "counters")) ||
      (st = al((void **)&rb->qctr, 4 * kQctrWords, "queue counters")) ||
      (st = al((void **)&rb->partial, sizeof(float) * 6 * total, "partial sums")) ||
      (st = al((void **)&rb->partial_n, 4 * total, "partial counts"))) {
    rb->release();
    return st;
  }
  ps.n_slots = S;
  rb->slots = S;
  rb->total = total;
  return AJG_OK;
}

}  // namespace ajg

using namespace ajg;


static void scene_free(ajg_scene *s) {
  if (!s) return;
  destroy_accel(s->accel);
  (void)hipFree(s->tri_shade);
  (void)hipFree(s->mats);
  for (void *p : s->light_bufs) (void)hipFree(p);
  delete (RenderBuffers *)s->render_cache;
  delete (GuideBuffers *)s->guide_cache;
  sppm_free(s->sppm_cache);
  vcm_free(s->vcm_cache);
  for (ajg_accel *b : s->blas) destroy_accel(b);
  (void)hipFree(s->insts);
  (void)hipFree(s->blases);
  (void)hipFree(s->trace_ctr);
  delete s;
}


// Argument checks shared by one-level and instanced scenes; counts the
// mesh normals the primitives bring.
// An emissive primitive (non-black emission, GeometricPrimitive,
// geometric.cpp:35-51) carries an area light.
static bool emissive(const ajg_primitive &p) {
  return std::max(std::max(p.emission[0], p.emission[1]), p.emission[2]) > 0.f;
}

static int check_scene_inputs(ajg_ctx *ctx, const ajg_primitive *prims, uint32_t n_prims,
                              const ajg_material *mats, uint32_t n_mats, const ajg_light *lights,
                              uint32_t n_lights, size_t *n_normals, bool allow_emission = true) {
  uint32_t n_area = 0;
  for (uint32_t i = 0; i < n_prims; ++i) n_area += emissive(prims[i]) ? 1u : 0u;
  if (n_area && !allow_emission)
    return fail(ctx, AJG_ERR_INVALID, "instanced scene: emissive meshes are not supported");
  if (n_lights + n_area > (uint32_t)kMaxLights)
    return fail(ctx, AJG_ERR_INVALID,
                "scene: at most 4 lights (emissive primitives included) are supported");
  if (n_mats == 0) return fail(ctx, AJG_ERR_INVALID, "scene: no materials");
  if (n_mats > kMatMask + 1) return fail(ctx, AJG_ERR_INVALID, "scene: too many materials");
  for (uint32_t i = 0; i < n_mats; ++i)
    if (mats[i].type != AJG_MAT_DIFFUSE && mats[i].type != AJG_MAT_DISNEY)
      return fail(ctx, AJG_ERR_INVALID, "scene: unsupported material type");
  *n_normals = 0;
  for (uint32_t i = 0; i < n_prims; ++i) {
    if (prims[i].kind > AJG_GEOM_MESH)
      return fail(ctx, AJG_ERR_INVALID, "scene: unsupported geometry kind");
    if (prims[i].material >= n_mats)
      return fail(ctx, AJG_ERR_INVALID, "scene: material index out of range");
    if (prims[i].kind == AJG_GEOM_MESH) {
      if (!prims[i].normal_and_vs)
        return fail(ctx, AJG_ERR_INVALID, "scene: mesh without normal_and_vs");
      *n_normals += prims[i].view.position_size;
    }
    if (emissive(prims[i])) {
      const ajg_mesh_view &v = prims[i].view;
      if (v.index_size == 0 || !v.index_buffer || !v.position_buffer)
        return fail(ctx, AJG_ERR_INVALID, "scene: emissive primitive without triangles");
      if (prims[i].kind == AJG_GEOM_QUAD && (v.position_size < 4 || v.index_size != 2))
        return fail(ctx, AJG_ERR_INVALID, "scene: emissive quad needs 4 positions, 2 triangles");
      if (prims[i].kind == AJG_GEOM_CUBE && v.index_size != 12)
        return fail(ctx, AJG_ERR_INVALID, "scene: emissive cube needs 12 triangles");
    }
  }
  int env_count = 0;
  for (uint32_t i = 0; i < n_lights; ++i) {
    if (lights[i].type == AJG_LIGHT_ENV) {
      ++env_count;
      if (!lights[i].env_texels || lights[i].env_w == 0 || lights[i].env_h == 0)
        return fail(ctx, AJG_ERR_INVALID, "scene: env light without texels");
    } else if (lights[i].type != AJG_LIGHT_DIRECTIONAL) {
      return fail(ctx, AJG_ERR_INVALID, "scene: unsupported light type");
    }
  }
  if (env_count > 1) return fail(ctx, AJG_ERR_INVALID, "scene: more than one env light");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, AJG_ERR_HIP, "hipSetDevice");
  return AJG_OK;
}

// PostIntersect records of primitives [p0, p1) in flatten order (flags |
// material, then the three normal indices of a mesh triangle into the
// shared normal array, whose mesh normals are appended to `normals`).
static void shade_records(const ajg_primitive *prims, uint32_t p0, uint32_t p1,
                          std::vector<uint4> *by_id, std::vector<float4> *normals) {
  uint32_t area = 0;  // emissive primitives before p0 (none in instanced scenes)
  for (uint32_t p = 0; p < p0; ++p) area += emissive(prims[p]) ? 1u : 0u;
  for (uint32_t p = p0; p < p1; ++p) {
    const ajg_primitive &pr = prims[p];
    const size_t noff = normals->size();
    const uint32_t light = emissive(pr) ? ++area : 0u;  // light index + 1
    const uint32_t flags = (pr.kind << kKindShift) | (pr.two_sided ? kTwoSidedBit : 0u) |
                           (light << kLightShift) | (pr.material & kMatMask);
    for (size_t i = 0; i < pr.view.index_size; ++i) {
      uint4 rec = make_uint4(flags, 0, 0, 0);
      if (pr.kind == AJG_GEOM_MESH) {
        const char *base = (const char *)pr.view.index_buffer + pr.view.index_offset +
                           i * pr.view.index_stride;
        uint32_t w3[3];
        memcpy(w3, base, 12);
        rec.y = (uint32_t)(noff + w3[0]);
        rec.z = (uint32_t)(noff + w3[1]);
        rec.w = (uint32_t)(noff + w3[2]);
      }
      by_id->push_back(rec);
    }
    if (pr.kind == AJG_GEOM_MESH) {
      for (size_t v = 0; v < pr.view.position_size; ++v) {
        const float *src = pr.normal_and_vs + 4 * v;
        normals->push_back(make_float4(src[0], src[1], src[2], src[3]));
      }
    }
  }
}

// The leaf slot's record: the shade flags and, for a mesh triangle, its
// vertex normals copied from the concatenated normal array.
static TriShade tri_record(const uint4 &info, const std::vector<float4> &normals) {
  TriShade t;
  t.info = info;
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool mesh = (info.x >> kKindShift) == AJG_GEOM_MESH;
  t.n[0] = mesh ? normals[info.y] : z;
  t.n[1] = mesh ? normals[info.z] : z;
  t.n[2] = mesh ? normals[info.w] : z;
  return t;
}

static int download_perm(ajg_ctx *ctx, const ajg_accel *a, std::vector<uint32_t> *perm) {
  perm->resize(a->n_tri);
  if (hipMemcpy(perm->data(), a->perm, 4 * (size_t)a->n_tri, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(ctx, AJG_ERR_HIP, "perm download");
  return AJG_OK;
}

static ajg_scene *new_scene(ajg_ctx *ctx) {
  ajg_scene *s = new ajg_scene();
  s->ctx = ctx;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
    s->cu_count = prop.multiProcessorCount;
  return s;
}

// World bound and the light radius |max - center| (DirectionalLight/EnvLight
// Process, default_scene.cpp:40-48).
static void set_world(ajg_scene *s, const float *box) {
  for (int i = 0; i < 6; ++i) s->world[i] = box[i];
  box3 wb;
  wb.mn = ld3(box);
  wb.mx = ld3(box + 3);
  s->radius = length(sub(wb.mx, box_center(wb)));
}

static f3 view_position(const ajg_mesh_view &v, size_t i) {
  float p[3];
  memcpy(p, (const char *)v.position_buffer + v.position_offset + i * v.position_stride, 12);
  return mk3(p[0], p[1], p[2]);
}
static uint32_t view_word(const ajg_mesh_view &v, size_t tri, int k) {
  uint32_t w;
  memcpy(&w, (const char *)v.index_buffer + v.index_offset + tri * v.index_stride + 4 * k, 4);
  return w;
}

// AreaLight of an emissive primitive: the triangles its geometry's Sample
// draws from, in the order it indexes them, and the areas its pdfs use
// (Quad ctor, quad.cpp:35-50; Cube, cube.cpp:31-75 with the tuple words
// stored in reverse; MeshGeo ctor surface area, mesh.cpp:53-70, accumulated
// in a float with a double addend). `flat` holds the primitive's triangles
// in index-word order.
static int area_light_init(ajg_ctx *ctx, ajg_scene *s, DevLight *l, const ajg_primitive &pr,
                           const float *flat) {
  memset(l, 0, sizeof(*l));
  l->type = kLightArea;
  l->radiance = ld3(pr.emission);
  l->kind = pr.kind;
  l->two_sided = pr.two_sided ? 1u : 0u;
  const ajg_mesh_view &v = pr.view;
  const uint32_t n = (uint32_t)v.index_size;
  std::vector<float> st;
  auto push = [&](f3 p) {
    st.push_back(p.x);
    st.push_back(p.y);
    st.push_back(p.z);
  };
  if (pr.kind == AJG_GEOM_QUAD) {
    const f3 a = view_position(v, 0), b = view_position(v, 1), c = view_position(v, 2),
             d = view_position(v, 3);
    push(a), push(b), push(c), push(a), push(c), push(d);
    const float abc = length(cross(sub(b, a), sub(c, a))) * 0.5f;
    const float acd = length(cross(sub(c, a), sub(d, a))) * 0.5f;
    l->area = abc + acd;
    l->prob = abc / l->area;
    l->n_stri = 2;
  } else if (pr.kind == AJG_GEOM_CUBE) {
    for (uint32_t t = 0; t < n; ++t)
      for (int k = 0; k < 3; ++k) push(ld3(flat + 9 * (size_t)t + 3 * (2 - k)));
    const f3 a = ld3(st.data()), b = ld3(st.data() + 3), c = ld3(st.data() + 6);
    l->cube_pdf0 = 1.f / (12.f * length(cross(sub(b, a), sub(c, a))) * 0.5f);
    l->n_stri = n;
  } else {
    for (uint32_t t = 0; t < n; ++t)
      for (uint32_t k = 0; k < 3; ++k) {
        const uint32_t f = t + k;  // the index array read at the triangle number
        const uint32_t w = view_word(v, f / 3, (int)(f % 3));
        if (w >= v.position_size)
          return fail(ctx, AJG_ERR_INVALID, "scene: emissive mesh index out of range");
        push(view_position(v, w));
      }
    float area = 0.f;
    for (uint32_t t = 0; t < n; ++t) {
      const float *p = flat + 9 * (size_t)t;
      const f3 e0 = sub(ld3(p + 3), ld3(p)), e1 = sub(ld3(p + 6), ld3(p));
      area = (float)((double)area + 0.5 * (double)length(cross(e0, e1)));
    }
    l->area = area;
    l->n_stri = n;
  }
  float *d = nullptr;
  if (int r = dev_alloc(ctx, (void **)&d, sizeof(float) * st.size(), "area light triangles")) return r;
  s->light_bufs.push_back(d);
  const hipError_t e = hipMemcpy(d, st.data(), sizeof(float) * st.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(ctx, e, "area light upload");
  l->stris = d;
  return AJG_OK;
}

// Materials and lights of a scene whose geometry is in place (the env
// Distribution2D tables are built on the device, EnvLight ctor,
// env_light.cpp:29-48). The light list is DefaultScene's
// (default_scene.cpp:34-48): the emissive primitives' area lights in
// primitive order, then the scene's lights. `flat` holds the primitives'
// triangles in flatten order (one-level scenes; instanced scenes have no
// emissive primitives).
static int finish_scene(ajg_ctx *ctx, ajg_scene *s, const ajg_material *mats, uint32_t n_mats,
                        const ajg_light *lights_in, uint32_t n_lights_in, int32_t env_cdf_fp64,
                        const ajg_primitive *prims = nullptr, uint32_t n_prims = 0,
                        const float *flat = nullptr) {
  std::vector<DevMaterial> dm(n_mats);
  for (uint32_t i = 0; i < n_mats; ++i) {
    const ajg_material &m = mats[i];
    dm[i].type = m.type;
    for (int k = 0; k < 3; ++k) dm[i].base_color[k] = m.base_color[k];
    dm[i].subsurface = m.subsurface;
    dm[i].metallic = m.metallic;
    dm[i].specular = m.specular;
    dm[i].specular_tint = m.specular_tint;
    dm[i].roughness = m.roughness;
    dm[i].anisotropic = m.anisotropic;
    dm[i].sheen = m.sheen;
    dm[i].sheen_tint = m.sheen_tint;
    dm[i].clearcoat = m.clearcoat;
    dm[i].clearcoat_gloss = m.clearcoat_gloss;
  }
  int st = 0;
  if ((st = dev_alloc(ctx, (void **)&s->mats, sizeof(DevMaterial) * n_mats, "materials"))) return st;
  hipError_t e = hipMemcpy(s->mats, dm.data(), sizeof(DevMaterial) * n_mats, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(ctx, e, "material upload");
  s->n_mats = n_mats;
  s->n_area = 0;
  for (uint32_t p = 0, t0 = 0; p < n_prims; t0 += (uint32_t)prims[p].view.index_size, ++p) {
    if (!emissive(prims[p])) continue;
    if ((st = area_light_init(ctx, s, &s->lights[s->n_area], prims[p], flat + 9 * (size_t)t0)))
      return st;
    ++s->n_area;
  }
  const uint32_t n_lights = s->n_area + n_lights_in;
  s->n_lights = n_lights;
  for (uint32_t i = s->n_area; i < n_lights; ++i) {
    DevLight &l = s->lights[i];
    const ajg_light &src = lights_in[i - s->n_area];
    memset(&l, 0, sizeof(l));
    l.type = src.type;
    if (l.type == AJG_LIGHT_DIRECTIONAL) {
      const f3 d = normalize(ld3(src.dir));
      l.fr = make_frame(d);
      l.angle_cos = std::cos(src.angle);
      l.radiance = ld3(src.radiance);
      continue;
    }
    s->env_index = (int)i;
    const uint32_t w = src.env_w, h = src.env_h;
    float *tex, *wts, *cf, *cc, *ci, *mf, *mc, *mi;
    auto al = [&](float **p, size_t cnt, const char *what) {
      int r = dev_alloc(ctx, (void **)p, sizeof(float) * cnt, what);
      if (r == AJG_OK) s->light_bufs.push_back(*p);
      return r;
    };
    if ((st = al(&tex, 3 * (size_t)w * h, "env texels")) || (st = al(&wts, (size_t)w * h, "env weights")) ||
        (st = al(&cf, (size_t)w * h, "env cond func")) || (st = al(&cc, (size_t)(w + 1) * h, "env cond cdf")) ||
        (st = al(&ci, h, "env cond int")) || (st = al(&mf, h, "env marg func")) ||
        (st = al(&mc, h + 1, "env marg cdf")) || (st = al(&mi, 1, "env marg int")))
      return st;
    e = hipMemcpy(tex, src.env_texels, sizeof(float) * 3 * (size_t)w * h, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "env upload");
    l.env.w = w;
    l.env.h = h;
    l.env.tex = tex;
    l.env.cond_cdf = cc;
    l.env.cond_func = cf;
    l.env.cond_int = ci;
    l.env.marg_cdf = mc;
    l.env.marg_func = mf;
    l.env.rot = src.rot;
    l.env.scaler = src.scaler;
    hipLaunchKernelGGL(k_env_weights, dim3(cdiv((uint64_t)w * h, 256)), dim3(256), 0, nullptr,
                       l.env, wts);
    hipLaunchKernelGGL(k_env_rows, dim3(cdiv(h, 64)), dim3(64), 0, nullptr, (const float *)wts, w, h,
                       cf, cc, ci, (int)env_cdf_fp64);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(64), 0, nullptr, (const float *)ci, h, mf, mc, mi,
                       (int)env_cdf_fp64);
    e = hipMemcpy(&l.env.marg_int, mi, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "env tables");
  }
  if ((st = dev_alloc(ctx, (void **)&s->trace_ctr, 16, "trace cursor"))) return st;
  return AJG_OK;
}

// Instance boxes and maps (host, once per scene). The world->object map is
// the inverse of the given object->world map, computed in double and rounded
// to float; the normal map is cbrt|det A| (A^-1)^T for the linear part A, so a
// rotation with uniform scale maps normals by the rotation alone. The world
// box of an instance is the box of its mesh's root box corners mapped to world
// space, widened by 2^-18 of its largest coordinate magnitude (so that the
// rounding of corner and ray transforms cannot cull a hit).
static bool instance_maps(const float *o2w, const float *root_box, float *w2o, float *nrm,
                          float *box) {
  double A[3][3], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) A[r][c] = o2w[4 * r + c];
    t[r] = o2w[4 * r + 3];
  }
  const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                     A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                     A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) return false;
  double inv[3][3];
  inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  const double sc = std::cbrt(std::fabs(det));
  for (int r = 0; r < 3; ++r) {
    double tr = 0;
    for (int c = 0; c < 3; ++c) {
      w2o[4 * r + c] = (float)inv[r][c];
      tr -= inv[r][c] * t[c];
      nrm[4 * r + c] = (float)(inv[c][r] * sc);
    }
    w2o[4 * r + 3] = (float)tr;
    nrm[4 * r + 3] = 0.f;
  }
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < 8; ++k) {
    const float p[3] = {root_box[(k & 1) ? 3 : 0], root_box[(k & 2) ? 4 : 1], root_box[(k & 4) ? 5 : 2]};
    for (int r = 0; r < 3; ++r) {
      const float v = ((o2w[4 * r] * p[0] + o2w[4 * r + 1] * p[1]) + o2w[4 * r + 2] * p[2]) + o2w[4 * r + 3];
      mn[r] = std::min(mn[r], v);
      mx[r] = std::max(mx[r], v);
    }
  }
  float mag = 0.f;
  for (int r = 0; r < 3; ++r) mag = std::max(mag, std::max(std::fabs(mn[r]), std::fabs(mx[r])));
  const float pad = std::ldexp(mag, -18) + 1e-30f;
  for (int r = 0; r < 3; ++r) {
    box[r] = mn[r] - pad;
    box[3 + r] = mx[r] + pad;
    if (!std::isfinite(box[r]) || !std::isfinite(box[3 + r])) return false;
  }
  return true;
}

extern "C" {

int ajg_scene_create(ajg_ctx *ctx, const ajg_primitive *prims, uint32_t n_prims,
                     const ajg_material *mats, uint32_t n_mats, const ajg_light *lights,
                     uint32_t n_lights, uint32_t search_radius, int32_t env_cdf_fp64,
                     ajg_scene **out) {
  if (!ctx || !out || (n_prims && !prims) || (n_mats && !mats) || (n_lights && !lights))
    return AJG_ERR_INVALID;
  *out = nullptr;
  if (int st = injected(ctx)) return st;
  size_t n_normals = 0;
  if (int st = check_scene_inputs(ctx, prims, n_prims, mats, n_mats, lights, n_lights, &n_normals))
    return st;
  std::vector<ajg_mesh_view> views(n_prims);
  for (uint32_t i = 0; i < n_prims; ++i) views[i] = prims[i].view;
  std::vector<float> tri;
  float box[6];
  if (int st = flatten_views(ctx, views.data(), n_prims, &tri, box)) return st;
  ajg_scene *s = new_scene(ctx);
  auto bail = [&](int st) {
    scene_free(s);
    return st;
  };
  if (int st = upload_and_build(ctx, tri, box, search_radius, nullptr, &s->accel)) return bail(st);
  set_world(s, box);
  // per-triangle shading records, permuted into leaf-slot order
  const uint32_t n = s->accel->n_tri;
  std::vector<uint32_t> perm;
  if (int st = download_perm(ctx, s->accel, &perm)) return bail(st);
  std::vector<uint4> by_id;
  std::vector<float4> normals;
  by_id.reserve(n);
  normals.reserve(n_normals);
  shade_records(prims, 0, n_prims, &by_id, &normals);
  std::vector<TriShade> by_slot(n);
  for (uint32_t i = 0; i < n; ++i) by_slot[i] = tri_record(by_id[perm[i]], normals);
  int st = 0;
  if ((st = dev_alloc(ctx, (void **)&s->tri_shade, sizeof(TriShade) * n, "tri shade")))
    return bail(st);
  hipError_t e = hipMemcpy(s->tri_shade, by_slot.data(), sizeof(TriShade) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(hip_fail(ctx, e, "scene upload"));
  if ((st = finish_scene(ctx, s, mats, n_mats, lights, n_lights, env_cdf_fp64, prims, n_prims,
                          tri.data())))
    return bail(st);
  *out = s;
  return AJG_OK;
}

int ajg_scene_create_instanced(ajg_ctx *ctx, const ajg_primitive *meshes, uint32_t n_meshes,
                               const ajg_instance *instances, uint32_t n_instances,
                               const ajg_material *mats, uint32_t n_mats, const ajg_light *lights,
                               uint32_t n_lights, uint32_t search_radius, int32_t env_cdf_fp64,
                               ajg_scene **out) {
  if (!ctx || !out || !meshes || n_meshes == 0 || !instances || n_instances == 0 ||
      (n_mats && !mats) || (n_lights && !lights))
    return AJG_ERR_INVALID;
  *out = nullptr;
  if (int st = injected(ctx)) return st;
  size_t n_normals = 0;
  if (int st = check_scene_inputs(ctx, meshes, n_meshes, mats, n_mats, lights, n_lights, &n_normals,
                                  false))
    return st;
  for (uint32_t i = 0; i < n_instances; ++i) {
    if (instances[i].mesh >= n_meshes)
      return fail(ctx, AJG_ERR_INVALID, "instanced scene: mesh index out of range");
    if (instances[i].material != AJG_INSTANCE_KEEP_MATERIAL && instances[i].material >= n_mats)
      return fail(ctx, AJG_ERR_INVALID, "instanced scene: material index out of range");
  }
  ajg_scene *s = new_scene(ctx);
  auto bail = [&](int st) {
    scene_free(s);
    return st;
  };
  // bottom level: one PLOC tree per mesh, in the mesh's own space
  std::vector<TriShade> by_slot;
  std::vector<float4> normals;
  normals.reserve(n_normals);
  std::vector<size_t> shade_off(n_meshes);
  uint32_t max_depth = 0;
  for (uint32_t m = 0; m < n_meshes; ++m) {
    std::vector<float> tri;
    float box[6];
    if (int st = flatten_views(ctx, &meshes[m].view, 1, &tri, box)) return bail(st);
    ajg_accel *acc = nullptr;
    if (int st = upload_and_build(ctx, tri, box, search_radius, nullptr, &acc)) return bail(st);
    s->blas.push_back(acc);
    std::vector<uint32_t> perm;
    if (int st = download_perm(ctx, acc, &perm)) return bail(st);
    std::vector<uint4> by_id;
    by_id.reserve(acc->n_tri);
    shade_records(meshes, m, m + 1, &by_id, &normals);
    shade_off[m] = by_slot.size();
    for (uint32_t i = 0; i < acc->n_tri; ++i) by_slot.push_back(tri_record(by_id[perm[i]], normals));
    max_depth = std::max(max_depth, acc->depth);
  }
  // top level: PLOC over the instance boxes, each given to the builder as the
  // degenerate triangle (min, max, min), whose box is exactly the instance box
  s->n_inst = n_instances;
  s->inst_w2o.resize(12 * (size_t)n_instances);
  s->inst_box.resize(6 * (size_t)n_instances);
  std::vector<DevInst> di(n_instances);
  std::vector<float> tl(9 * (size_t)n_instances);
  float wbox[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n_instances; ++i) {
    const ajg_instance &in = instances[i];
    float w2o[12], nrm[12], box[6];
    if (!instance_maps(in.object2world, s->blas[in.mesh]->root_box, w2o, nrm, box))
      return bail(fail(ctx, AJG_ERR_INVALID, "instanced scene: singular or non-finite object2world"));
    memcpy(&s->inst_w2o[12 * (size_t)i], w2o, sizeof(w2o));
    memcpy(&s->inst_box[6 * (size_t)i], box, sizeof(box));
    for (int r = 0; r < 3; ++r) {
      di[i].w2o[r] = make_float4(w2o[4 * r], w2o[4 * r + 1], w2o[4 * r + 2], w2o[4 * r + 3]);
      di[i].nrm[r] = make_float4(nrm[4 * r], nrm[4 * r + 1], nrm[4 * r + 2], 0.f);
    }
    di[i].info = make_uint4(in.mesh, in.material, 0u, 0u);
    float *t9 = &tl[9 * (size_t)i];
    for (int k = 0; k < 3; ++k) {
      t9[k] = box[k];
      t9[3 + k] = box[3 + k];
      t9[6 + k] = box[k];
      wbox[k] = std::min(wbox[k], box[k]);
      wbox[3 + k] = std::max(wbox[3 + k], box[3 + k]);
    }
  }
  if (int st = upload_and_build(ctx, tl, wbox, search_radius, nullptr, &s->accel)) return bail(st);
  s->depth2 = s->accel->depth + max_depth;
  set_world(s, wbox);
  std::vector<DevBlas> db(n_meshes);
  int st = 0;
  if ((st = dev_alloc(ctx, (void **)&s->tri_shade, sizeof(TriShade) * by_slot.size(), "tri shade")) ||
      (st = dev_alloc(ctx, (void **)&s->insts, sizeof(DevInst) * n_instances, "instances")) ||
      (st = dev_alloc(ctx, (void **)&s->blases, sizeof(DevBlas) * n_meshes, "mesh trees")))
    return bail(st);
  for (uint32_t m = 0; m < n_meshes; ++m) {
    const ajg_accel *a = s->blas[m];
    db[m].bvh.wide = a->wide;
    db[m].bvh.leaf_tris = a->leaf_tris;
    db[m].bvh.root_ref = a->root_ref;
    for (int k = 0; k < 3; ++k) {
      db[m].bvh.root_mn[k] = a->root_box[k];
      db[m].bvh.root_mx[k] = a->root_box[3 + k];
    }
    db[m].tri_shade = s->tri_shade + shade_off[m];
    db[m].perm = a->perm;
  }
  hipError_t e = hipMemcpy(s->tri_shade, by_slot.data(), sizeof(TriShade) * by_slot.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->insts, di.data(), sizeof(DevInst) * n_instances, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->blases, db.data(), sizeof(DevBlas) * n_meshes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(hip_fail(ctx, e, "instanced scene upload"));
  if ((st = finish_scene(ctx, s, mats, n_mats, lights, n_lights, env_cdf_fp64))) return bail(st);
  *out = s;
  return AJG_OK;
}

int ajg_scene_instances(ajg_scene *scene, float *world2object, float *boxes) {
  if (!scene || !scene->insts) return AJG_ERR_INVALID;
  if (world2object) memcpy(world2object, scene->inst_w2o.data(), 4 * scene->inst_w2o.size());
  if (boxes) memcpy(boxes, scene->inst_box.data(), 4 * scene->inst_box.size());
  return AJG_OK;
}

int ajg_scene_trace_closest(ajg_scene *scene, const ajg_ray *rays, uint32_t n, ajg_hit *hits,
                            uint32_t *instance, void *stream) {
  if (!scene || (n && (!rays || !hits))) return AJG_ERR_INVALID;
  if (int st = injected(scene->ctx)) return st;
  if (n == 0) return AJG_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!scene->insts) {
    if (int st = launch_trace_closest(scene->accel, rays, n, hits, s)) return st;
    if (instance)
      hipLaunchKernelGGL(k_one_level_instance, dim3(cdiv(n, 256)), dim3(256), 0, s,
                         (const ajg_hit *)hits, n, instance);
  } else {
    const int spill = spill_for_depth(scene->depth2);
    if (!spill) return fail(scene->ctx, AJG_ERR_INTERNAL, "trace: tree too deep for the traversal stack");
    const DevScene ds = dev_scene(scene);
    const uint32_t grid = std::min<uint32_t>(cdiv(n, kTravThreads),
                                             (uint32_t)scene->cu_count * kTravBlocksPerCU);
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, s, scene->trace_ctr);
    if (spill == kSpillS)
      hipLaunchKernelGGL(k_scene_closest2<kSpillS>, dim3(grid), dim3(kTravThreads), 0, s, ds.bvh, ds.inst,
                         ds.blas, rays, n, hits, instance, scene->trace_ctr);
    else
      hipLaunchKernelGGL(k_scene_closest2<kSpillL>, dim3(grid), dim3(kTravThreads), 0, s, ds.bvh, ds.inst,
                         ds.blas, rays, n, hits, instance, scene->trace_ctr);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(scene->ctx, e, "scene trace closest");
}

int ajg_scene_trace_any(ajg_scene *scene, const ajg_ray *rays, uint32_t n, uint8_t *occluded,
                        void *stream) {
  if (!scene || (n && (!rays || !occluded))) return AJG_ERR_INVALID;
  if (int st = injected(scene->ctx)) return st;
  if (n == 0) return AJG_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!scene->insts) return launch_trace_any(scene->accel, rays, n, occluded, s);
  const int spill = spill_for_depth(scene->depth2);
  if (!spill) return fail(scene->ctx, AJG_ERR_INTERNAL, "trace: tree too deep for the traversal stack");
  const DevScene ds = dev_scene(scene);
  if (spill == kSpillS)
    hipLaunchKernelGGL(k_scene_any2<kSpillS>, dim3(cdiv(n, kTravThreads)), dim3(kTravThreads), 0, s, ds.bvh,
                       ds.inst, ds.blas, rays, n, occluded);
  else
    hipLaunchKernelGGL(k_scene_any2<kSpillL>, dim3(cdiv(n, kTravThreads)), dim3(kTravThreads), 0, s, ds.bvh,
                       ds.inst, ds.blas, rays, n, occluded);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(scene->ctx, e, "scene trace any");
}

int ajg_scene_destroy(ajg_scene *scene) {
  scene_free(scene);
  return AJG_OK;
}

int ajg_scene_get_accel(ajg_scene *scene, ajg_accel **out) {
  if (!scene || !out) return AJG_ERR_INVALID;
  *out = scene->accel;
  return AJG_OK;
}

int ajg_scene_env_tables(ajg_scene *s, uint32_t li, float *cond_cdf, float *cond_func_int,
                         float *marg_cdf, float *marg_func_int) {
  if (!s || li >= s->n_lights - s->n_area) return AJG_ERR_INVALID;
  li += s->n_area;  // the scene's lights follow the emissive primitives' area lights
  if (s->lights[li].type != AJG_LIGHT_ENV) return AJG_ERR_INVALID;
  const DevEnv &e = s->lights[li].env;
  hipError_t r = hipSuccess;
  if (cond_cdf) r = hipMemcpy(cond_cdf, e.cond_cdf, sizeof(float) * (e.w + 1) * e.h, hipMemcpyDeviceToHost);
  if (r == hipSuccess && cond_func_int)
    r = hipMemcpy(cond_func_int, e.cond_int, sizeof(float) * e.h, hipMemcpyDeviceToHost);
  if (r == hipSuccess && marg_cdf) r = hipMemcpy(marg_cdf, e.marg_cdf, sizeof(float) * (e.h + 1), hipMemcpyDeviceToHost);
  if (r != hipSuccess) return hip_fail(s->ctx, r, "env tables download");
  if (marg_func_int) *marg_func_int = e.marg_int;
  return AJG_OK;
}

}  // extern "C"

// RenderParams of one render: PTRendererArgs semantics (renderer.h:37-45),
// traced area = whole tiles (pt.cpp:43-49). Work items are (pixel, chunk of
// samples): seed_mode 0 keeps a pixel's samples in one item (one sequential
// stream); seed_mode 1 splits them into chunks of `samples_per_item`
// (0: enough chunks to fill `target` slots).
static int setup_params(ajg_ctx *ctx, const ajg_camera *camera, const ajg_pt_args *args,
                        uint64_t target, uint32_t samples_per_item, RenderParams *out) {
  if (args->spp <= 0 || args->tile_size <= 0 || args->max_bounces < 0)
    return fail(ctx, AJG_ERR_INVALID, "render: need spp > 0, tile_size > 0, max_bounces >= 0");
  const float rx = camera->resolution[0], ry = camera->resolution[1];
  if (!(rx >= 1.f) || !(ry >= 1.f) || rx > 65536.f || ry > 65536.f)
    return fail(ctx, AJG_ERR_INVALID, "render: bad resolution");
  RenderParams &P = *out;
  memset(&P, 0, sizeof(P));
  memcpy(P.cam.r2c, camera->raster2camera, sizeof(P.cam.r2c));
  memcpy(P.cam.c2w, camera->camera2world, sizeof(P.cam.c2w));
  P.cam.lens_radius = camera->lens_radius;
  P.cam.focal_distance = camera->focal_distance;
  P.W = (uint32_t)(int)rx;
  P.H = (uint32_t)(int)ry;
  const uint32_t ts = (uint32_t)args->tile_size;
  P.Wp = (P.W + ts - 1) / ts * ts;
  const uint32_t Hp = (P.H + ts - 1) / ts * ts;
  P.n_pix = P.Wp * Hp;
  P.spp = (uint32_t)args->spp;
  P.min_bounces = args->min_bounces;
  P.max_bounces = args->max_bounces;
  P.cont_prob = args->cont_prob;
  P.seed_mode = args->seed_mode;
  P.seed_base = args->seed_base;
  P.flags = args->flags;
  P.use_mis = args->use_mis ? 1u : 0u;
  if (P.seed_mode == 0) {
    P.chunk = P.spp;
    P.n_chunks = 1;
  } else if (samples_per_item > 0) {
    P.chunk = std::min(P.spp, samples_per_item);
    P.n_chunks = cdiv(P.spp, P.chunk);
  } else {
    uint32_t nc = (uint32_t)std::min<uint64_t>(P.spp, std::max<uint64_t>(1, cdiv(target, P.n_pix)));
    P.chunk = cdiv(P.spp, nc);
    P.n_chunks = cdiv(P.spp, P.chunk);
  }
  const uint64_t total = (uint64_t)P.n_pix * P.n_chunks;
  if (total > 0xFFFFFFF0ull) return fail(ctx, AJG_ERR_INVALID, "render: too much work");
  P.total_work = (uint32_t)total;
  return AJG_OK;
}

DevScene dev_scene(const ajg_scene *sc) {
  const ajg_accel *a = sc->accel;
  DevScene ds;
  memset(&ds, 0, sizeof(ds));
  ds.bvh.wide = a->wide;
  ds.bvh.leaf_tris = a->leaf_tris;
  ds.bvh.root_ref = a->root_ref;
  for (int i = 0; i < 3; ++i) {
    ds.bvh.root_mn[i] = a->root_box[i];
    ds.bvh.root_mx[i] = a->root_box[3 + i];
  }
  ds.tri_shade = sc->tri_shade;
  ds.mats = sc->mats;
  for (int i = 0; i < kMaxLights; ++i) ds.lights[i] = sc->lights[i];
  ds.n_lights = sc->n_lights;
  ds.n_area = sc->n_area;
  ds.env_index = sc->env_index;
  ds.radius = sc->radius;
  ds.inst = sc->insts;
  ds.blas = sc->blases;
  return ds;
}

static int render_impl(ajg_scene *sc, const ajg_camera *camera, const ajg_pt_args *args,
                       ajg_guide *gd, const ajg_guide_args *ga, const ajg_film *film,
                       ajg_render_stats *stats, ajg_guide_stats *gstats, void *stream) {
  if (!sc || !camera || !args || !film || !film->sum || !film->sum_sq || !film->count)
    return AJG_ERR_INVALID;
  ajg_ctx *ctx = sc->ctx;
  if (int st = injected(ctx)) return st;
  const bool guided = gd != nullptr;
  if (guided) {
    if (!ga || !(ga->alpha >= 0.f && ga->alpha <= 1.f) || (ga->train && (ga->batch == 0 || !(ga->lr > 0.f))))
      return fail(ctx, AJG_ERR_INVALID, "guided render: need 0 <= alpha <= 1, batch > 0 and lr > 0");
    if (!args->use_mis)
      return fail(ctx, AJG_ERR_INVALID, "guided render: the guide mixes into TraceMIS (use_mis 1)");
  }
  hipStream_t s = (hipStream_t)stream;
  const ajg_accel *a = sc->accel;
  const uint64_t target = args->max_slots > 0 ? (uint64_t)args->max_slots : (1ull << 20);
  RenderParams P;
  // seed_mode 1: about kItemsPerSlot work items per slot, so a slot that
  // finishes early takes more work (with one item per slot most slots sat
  // idle through the render's last iterations while the slots with the
  // longest sample sequences finished: heavy mesh, 8M slots, 76 -> 36
  // iterations, 433 -> 485 M paths/s)
  if (int st = setup_params(ctx, camera, args, kItemsPerSlot * target, 0, &P)) return st;
  const uint64_t total = P.total_work;
  P.slots = (uint32_t)std::min<uint64_t>(target, total);
  const uint32_t S = P.slots;
  const DevScene ds = dev_scene(sc);

  // ---- buffers (cached on the scene across renders)
  RenderBuffers *rb = (RenderBuffers *)sc->render_cache;
  if (!rb) {
    rb = new RenderBuffers();
    sc->render_cache = rb;
  }
  if (int st = ensure_buffers(ctx, rb, S, total)) return st;
  if (int st = ensure_aux_stream(ctx, rb)) return st;
  PathState ps = rb->ps;
  GuideParams G;
  memset(&G, 0, sizeof(G));
  BatchParams nob;
  memset(&nob, 0, sizeof(nob));
  GuideBuffers *gb = nullptr;
  uint32_t batch = 0;
  if (guided) {
    gb = (GuideBuffers *)sc->guide_cache;
    if (!gb) {
      gb = new GuideBuffers();
      sc->guide_cache = gb;
    }
    // default: every slot records, within a 256 MiB vertex-history budget
    const uint32_t D0 = (uint32_t)std::max(1, std::min(args->max_bounces, 32));
    const uint32_t by_mem = (uint32_t)((256ull << 20) / (4ull * kGHist * D0));
    const uint32_t want = ga->train ? (ga->train_slots ? std::min(ga->train_slots, S)
                                                       : std::min(S, by_mem))
                                    : 0u;
    const uint32_t stride = want ? cdiv(S, want) : 0u;
    const uint32_t T = stride ? cdiv(S, stride) : 0u;  // slots with slot % stride == 0
    const uint32_t D = (uint32_t)std::max(1, std::min(args->max_bounces, 32));
    batch = ga->train ? ga->batch : 0u;
    if (int st = ensure_guide_buffers(ctx, gb, S, std::max(T, 1u), D)) return st;
    // records accumulate in the guide's pool across renders until a batch is full
    uint32_t cap = 0;
    if (int st = guide_records(gd, ga->train ? std::max<uint32_t>(4u * batch, 1u << 16) : 1u,
                               &G.rec, &G.rec_count, &cap))
      return st;
    G.rec_cap = cap;
    G.mix = gb->mix;
    G.alpha = ga->alpha;
    G.train_stride = stride;
    G.hist_depth = D;
    G.hist = gb->hist;
    G.hist_n = gb->hist_n;
    G.pend_pos = gb->pend_pos;
    G.pend_tw = gb->pend_tw;
  }
  Queues q;
  memset(&q, 0, sizeof(q));
  const size_t cq_cap = cq_capacity(S), aq_cap = aq_capacity(S);
  DevRay **c_rays = rb->c_rays, **a_rays = rb->a_rays;
  float4 *c_hits = rb->c_hits;
  uint8_t *a_occ = rb->a_occ;
  uint32_t *counters = rb->counters;  // live, work, guide total | traversal u64 x4 | totals
  float *partial = rb->partial;
  uint32_t *partial_n = rb->partial_n;
  hipEvent_t ev0, ev1;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  auto done = [&](int code) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return code;
  };
#define AJG_RCHK(call)                                          \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return done(hip_fail(ctx, e_, #call)); \
  } while (0)
  // Slot state, per-work-item partial sums, counters, queues and the film
  // start cleared; the work counter starts at min(S, total): a slot's first
  // work item is its own index (next_sample), with no atomic.
  Fills fl;
  fl.words(ps.work, S, kNone);
  fl.words(ps.sample, S);
  fl.words(ps.sample_end, S);
  fl.words(ps.bank, S);
  fl.words(ps.rng, 2ull * S);
  fl.words(ps.pend_flags, S);
  fl.words(ps.acc, 24ull * S);
  fl.words(ps.acc_n, 2ull * S);
  fl.words(partial, 6ull * total);
  fl.words(partial_n, total);
  fl.words(counters, kCounterWords);
  fl.words(rb->qctr, kQctrWords);
  fl.bytes(a_occ, aq_cap);
  fl.words(c_hits, 4ull * cq_cap);
  const size_t n_pix_sz = P.n_pix;
  fl.words(film->sum, 3ull * n_pix_sz);
  fl.words(film->sum_sq, 3ull * n_pix_sz);
  fl.words(film->count, n_pix_sz);
  if (film->per_sample) fl.words(film->per_sample, 3ull * n_pix_sz * P.spp);
  if (guided) fl.words(gb->hist_n, gb->train_slots);
  AJG_RCHK(fl.launch(s));
  // the first shade runs over every slot (list 0 holds them all)
  const uint32_t lv_len = lv_stripe_len(S);
  hipLaunchKernelGGL(k_lv_init, dim3(cdiv(S, 256)), dim3(256), 0, s, rb->lv[0], rb->lv_ei[0],
                     rb->qctr + kListOff, S,
                     lv_len);
  // guided renders keep the fetch order through the counter for every item
  // (their training slots are every train_stride-th slot, and a slot's
  // first item equal to its index gave the guide a markedly worse fit)
  const bool first_static = !guided;
  unsigned long long rec_before = 0;
  if (guided) {
    uint32_t rc[4] = {0, 0, 0, 0};
    AJG_RCHK(hipMemcpyAsync(rc, G.rec_count, 16, hipMemcpyDeviceToHost, s));
    AJG_RCHK(hipStreamSynchronize(s));
    unsigned long long rt;
    memcpy(&rt, rc + 2, 8);
    rec_before = rt + rc[0];
  }

  q.c_hits_in = c_hits;
  q.c_inst_in = rb->c_inst;
  q.a_occ_in = a_occ;
  q.live = counters + 4;
  q.work_counter = rb->qctr + kWorkOff;
  q.work_stripes = work_stripes_for(S);
  q.first_static = first_static ? 1u : 0u;
  q.work_static = first_static ? (uint32_t)std::min<uint64_t>(S, total) : 0u;
  q.partial = partial;
  q.partial_n = partial_n;
  q.per_sample = film->per_sample;
  q.lv_total = counters + 5;
  q.lv_len = lv_len;
  const bool inst = ds.inst != nullptr;
  const int stack = spill_for_depth(inst ? sc->depth2 : a->depth);
  if (!stack) return done(fail(ctx, AJG_ERR_INTERNAL, "render: tree too deep for the traversal stack"));

  const bool counting = (args->flags & AJG_RENDER_COUNT) != 0;
  const bool inl = (args->flags & AJG_RENDER_INLINE) != 0;
  if (inl && (guided || inst || counting))
    return done(fail(ctx, AJG_ERR_INVALID,
                     "render: AJG_RENDER_INLINE takes one-level scenes, no guide, no AJG_RENDER_COUNT"));
  StageTimer timer;
  timer.on = (args->flags & AJG_RENDER_TIMING) != 0;
  // traversal counters: closest (nodes, tris) at +8, any (nodes, tris) at +12
  unsigned long long *trav_ctr = (unsigned long long *)(counters + 8);
  (void)hipEventRecord(ev0, s);
  int cur = 0;
  uint32_t iters = 0;
  uint32_t train_steps = 0;
  // an untrained guide is not sampled (alpha 0 until the first step)
  if (guided && batch && guide_steps(gd) == 0) G.alpha = 0.f;
  const uint32_t shade_grid = cdiv(S, 256);
  // entries per queue stripe: the extension ray (and, with area lights, the
  // MIS-BSDF ray) per slot; one shadow ray per light and the MIS-BSDF ray
  const uint32_t c_len = stripe_len(S, ds.n_area ? 2 : 1);
  const uint32_t a_len = stripe_len(S, ds.n_lights + 1);
  // striped fetch cursors of the closest-hit and shadow-ray traversal
  uint32_t *const c_cursor = rb->qctr + 4 * kQWords;
  uint32_t *const a_cursor = c_cursor + kCursorStripes * kCursorStride;
  const uint32_t trav_grid = trav_grid_size((uint32_t)std::max(1, sc->cu_count) * kTravBlocksPerCU);
  // counter words: 4 live flag, 6 closest-queue total for the guide, 8-15
  // u64 traversal counts, 16-19 u64 query totals (closest, shadow); the
  // queue counters, fetch cursors and work-item counters are in rb->qctr.
  unsigned long long *totals = (unsigned long long *)(counters + 16);
  if (inl) {
    const uint32_t lt = a->n_tri >= 2 && a->n_tri <= kInlineLdsTris ? a->n_tri : 0u;
    const size_t lds = lt ? 16 * (size_t)((lt - 1) * 4 + lt * 3) : 0;
    if (stack == kSpillS)
      hipLaunchKernelGGL(k_pt_inline<kSpillS>, dim3(cdiv(total, 256)), dim3(256), lds, s, ds, P, partial,
                         partial_n, film->per_sample, totals, lt);
    else
      hipLaunchKernelGGL(k_pt_inline<kSpillL>, dim3(cdiv(total, 256)), dim3(256), lds, s, ds, P, partial,
                         partial_n, film->per_sample, totals, lt);
    AJG_RCHK(hipGetLastError());
  }
  if (!inl) for (;;) {
    for (int k = 0; k < kItersPerCheck; ++k) {
      // shade reads queue `cur` (traced) and writes queue `nxt`
      const int nxt = cur ^ 1;
      q.c_rays_in = c_rays[cur];
      q.c_rays_out = c_rays[nxt];
      q.c_count_out = rb->qctr + nxt * kQWords;
      q.c_len = c_len;
      q.c_in = {rb->qctr + cur * kQWords, c_len};
      q.a_rays_in = a_rays[cur];
      q.a_rays_out = a_rays[nxt];
      q.a_count_out = rb->qctr + (2 + nxt) * kQWords;
      q.a_len = a_len;
      q.lv_in = rb->lv[cur];
      q.lv_ei_in = rb->lv_ei[cur];
      q.lv_q = {rb->qctr + kListOff + cur * kQWords, lv_len};
      q.lv_out = rb->lv[nxt];
      q.lv_ei_out = rb->lv_ei[nxt];
      q.lv_count_out = rb->qctr + kListOff + nxt * kQWords;
      const QIn cq_in = {rb->qctr + nxt * kQWords, c_len};
      const QIn aq_in = {rb->qctr + (2 + nxt) * kQWords, a_len};
      hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(64), 0, s, counters, rb->qctr, cur, nxt,
                         totals);
      timer.begin(2, s);
      if (guided)
        hipLaunchKernelGGL((k_shade_pend<true, false>), dim3(shade_grid), dim3(256), 0, s, ds, P, ps, q,
                           G, nob);
      else
        hipLaunchKernelGGL((k_shade_pend<false, false>), dim3(shade_grid), dim3(256), 0, s, ds, P, ps, q,
                           G, nob);
      if (guided)
        hipLaunchKernelGGL((k_shade<true, false>), dim3(shade_grid), dim3(256), 0, s, ds, P, ps, q, G,
                           nob);
      else
        hipLaunchKernelGGL((k_shade<false, false>), dim3(shade_grid), dim3(256), 0, s, ds, P, ps, q, G,
                           nob);
      timer.end(2, s);
      AJG_RCHK(hipEventRecord(rb->ev_fork, s));
      AJG_RCHK(hipStreamWaitEvent(rb->aux, rb->ev_fork, 0));
      timer.begin(0, s);
      if (inst && counting)
        launch_pq2<true>(stack, ds, true, c_rays[nxt], c_hits, rb->c_inst, cq_in, c_cursor,
                         trav_ctr, trav_grid, s);
      else if (inst)
        launch_pq2<false>(stack, ds, true, c_rays[nxt], c_hits, rb->c_inst, cq_in, c_cursor,
                          trav_ctr, trav_grid, s);
      else if (counting)
        launch_pq<true>(stack, ds.bvh, true, c_rays[nxt], c_hits, cq_in, c_cursor,
                        trav_ctr, trav_grid, s);
      else
        launch_pq<false>(stack, ds.bvh, true, c_rays[nxt], c_hits, cq_in, c_cursor,
                         trav_ctr, trav_grid, s);
      timer.end(0, s);
      {
        hipStream_t sa = rb->aux;
        timer.begin(1, sa);
        if (inst && counting)
          launch_pq2<true>(stack, ds, false, a_rays[nxt], a_occ, nullptr, aq_in, a_cursor,
                           trav_ctr + 2, trav_grid, sa);
        else if (inst)
          launch_pq2<false>(stack, ds, false, a_rays[nxt], a_occ, nullptr, aq_in, a_cursor,
                            trav_ctr + 2, trav_grid, sa);
        else if (counting)
          launch_pq<true>(stack, ds.bvh, false, a_rays[nxt], a_occ, aq_in,
                          a_cursor, trav_ctr + 2, trav_grid, sa);
        else
          launch_pq<false>(stack, ds.bvh, false, a_rays[nxt], a_occ, aq_in,
                           a_cursor, trav_ctr + 2, trav_grid, sa);
        timer.end(1, sa);
        AJG_RCHK(hipEventRecord(rb->ev_join, sa));
      }
      if (guided) {
        // guide mixtures of the new closest hits, read by the next shade
        timer.begin(3, s);
        // dense queue order; the queue's total goes to counter word 6
        const uint32_t n_cq = kQStripes * c_len;
        hipLaunchKernelGGL(k_hit_points, dim3(std::min<uint32_t>(cdiv(n_cq, 256), 4096)), dim3(256),
                           0, s, ds, (const DevRay *)c_rays[nxt], (const float4 *)c_hits,
                           (const uint32_t *)rb->c_inst, cq_in, gb->pts, counters + 6);
        if (int st = guide_forward_launch(gd, gb->pts, gb->pts + 3, n_cq, counters + 6, 6, nullptr,
                                          gb->mix, nullptr, s))
          return done(st);
        timer.end(3, s);
      }
      AJG_RCHK(hipStreamWaitEvent(s, rb->ev_join, 0));
      cur = nxt;
      ++iters;
    }
    uint32_t hc[8], held = 0;
    AJG_RCHK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    if (guided && batch) AJG_RCHK(hipMemcpyAsync(&held, G.rec_count, 4, hipMemcpyDeviceToHost, s));
    AJG_RCHK(hipStreamSynchronize(s));
    if (cancelled(ctx)) return done(fail(ctx, AJG_ERR_CANCELLED, "render cancelled"));
    const uint32_t live = hc[4];
    if (guided && batch) {
      const uint32_t avail = std::min(held, G.rec_cap);
      if (avail >= batch) {
        timer.begin(4, s);
        for (uint32_t j = 0; j + batch <= avail; j += batch) {
          if (ga->max_steps && train_steps >= ga->max_steps) break;
          for (uint32_t e = 0; e < std::max(1u, ga->steps_per_batch); ++e) {
            if (ga->max_steps && train_steps >= ga->max_steps) break;
            if (int st = guide_train_launch(gd, G.rec + (size_t)j * kGRec, batch, ga->lr, s))
              return done(st);
            ++train_steps;
          }
        }
        hipLaunchKernelGGL(k_rec_reset, dim3(1), dim3(64), 0, s, G.rec_count,
                           (unsigned long long *)(G.rec_count + 2));
        timer.end(4, s);
        // the guide is sampled once it has been trained at least once
        G.alpha = ga->alpha;
      }
    }
    if (live == 0) break;
    if (iters > 200000) return done(fail(ctx, AJG_ERR_INTERNAL, "render: iteration limit"));
  }
  unsigned long long tot[2] = {0, 0};
  if (P.n_chunks > 1)
    hipLaunchKernelGGL(k_reduce_chunks, dim3(cdiv(P.n_pix, 256)), dim3(256), 0, s, (const float *)partial,
                       (const uint32_t *)partial_n, P.n_pix, P.n_chunks, film->sum, film->sum_sq,
                       film->count);
  else
    hipLaunchKernelGGL(k_unpack_single, dim3(cdiv(P.n_pix, 256)), dim3(256), 0, s, (const float *)partial,
                       (const uint32_t *)partial_n, P.n_pix, film->sum, film->sum_sq, film->count);
  (void)hipEventRecord(ev1, s);
  AJG_RCHK(hipMemcpyAsync(tot, totals, 16, hipMemcpyDeviceToHost, s));
  AJG_RCHK(hipStreamSynchronize(s));
  AJG_RCHK(hipGetLastError());
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->closest_queries = tot[0];
    stats->any_queries = tot[1];
    stats->paths = (uint64_t)P.n_pix * P.spp;
    stats->iterations = iters;
    stats->slots = S;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    stats->ms_total = ms;
    stats->launches_closest = timer.launches[0];
    stats->launches_any = timer.launches[1];
    stats->launches_shade = timer.launches[2];
    if (timer.on) {
      stats->ms_trace_closest = timer.total_ms(0);
      stats->ms_trace_any = timer.total_ms(1);
      stats->ms_shade = timer.total_ms(2);
    }
    if (counting) {
      unsigned long long tc[4] = {0, 0, 0, 0};
      AJG_RCHK(hipMemcpy(tc, trav_ctr, 32, hipMemcpyDeviceToHost));
      stats->nodes_fetched = tc[0];
      stats->tris_tested = tc[1];
      stats->nodes_fetched_any = tc[2];
      stats->tris_tested_any = tc[3];
    }
  }
  if (gstats) {
    memset(gstats, 0, sizeof(*gstats));
    if (guided) {
      gstats->train_steps = train_steps;
      uint32_t rc[4] = {0, 0, 0, 0};
      AJG_RCHK(hipMemcpy(rc, G.rec_count, 16, hipMemcpyDeviceToHost));
      unsigned long long rt;
      memcpy(&rt, rc + 2, 8);
      gstats->records = rt + rc[0] - rec_before;
      if (timer.on) {
        gstats->ms_infer = timer.total_ms(3);
        gstats->ms_train = timer.total_ms(4);
      }
      if (train_steps) AJG_RCHK(hipMemcpy(&gstats->last_loss, guide_loss_dev(gd), 4, hipMemcpyDeviceToHost));
    }
  }
#undef AJG_RCHK
  return done(AJG_OK);
}

// ---------------------------------------------------------------- batches
// Buffers of batched renders, kept on the context: the render buffers of
// the shared wavefront, per-ray job ids of both queue generations, the
// device job table, job work-list prefix, completion counters and clocks,
// and the stream the iterations are captured on.
struct BatchBuffers {
  RenderBuffers rb;
  std::vector<void *> bufs;
  uint32_t slots = 0, n_jobs = 0;
  uint16_t *c_job[2] = {nullptr, nullptr};
  uint16_t *a_job[2] = {nullptr, nullptr};
  BatchJob *jobs = nullptr;
  uint32_t *job_begin = nullptr;
  uint32_t *remaining = nullptr;
  unsigned long long *clocks = nullptr;  // n_jobs completion clocks + start
  hipStream_t stream = nullptr;
  void release() {
    for (void *p : bufs) (void)hipFree(p);
    bufs.clear();
    slots = n_jobs = 0;
  }
  ~BatchBuffers() {
    release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void batch_cache_free(void *p) { delete (BatchBuffers *)p; }

static int ensure_batch_buffers(ajg_ctx *ctx, BatchBuffers *bb, uint32_t S, uint32_t n_jobs) {
  if (!bb->stream) {
    hipError_t e = hipStreamCreate(&bb->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamCreate");
  }
  if (bb->slots >= S && bb->n_jobs >= n_jobs && !bb->bufs.empty()) return AJG_OK;
  bb->release();
  auto al = [&](void **p, size_t bytes, const char *what) {
    int r = dev_alloc(ctx, p, bytes, what);
    if (r == AJG_OK) bb->bufs.push_back(*p);
    return r;
  };
  const size_t aq = aq_capacity(S);
  int st = 0;
  if ((st = al((void **)&bb->c_job[0], 2 * cq_capacity(S), "batch ray jobs")) ||
      (st = al((void **)&bb->c_job[1], 2 * cq_capacity(S), "batch ray jobs")) ||
      (st = al((void **)&bb->a_job[0], 2 * aq, "batch ray jobs")) ||
      (st = al((void **)&bb->a_job[1], 2 * aq, "batch ray jobs")) ||
      (st = al((void **)&bb->jobs, sizeof(BatchJob) * (size_t)n_jobs, "batch job table")) ||
      (st = al((void **)&bb->job_begin, 4 * ((size_t)n_jobs + 1), "batch work prefix")) ||
      (st = al((void **)&bb->remaining, 4 * (size_t)n_jobs, "batch job counters")) ||
      (st = al((void **)&bb->clocks, 8 * ((size_t)n_jobs + 1), "batch job clocks"))) {
    bb->release();
    return st;
  }
  bb->slots = S;
  bb->n_jobs = n_jobs;
  return AJG_OK;
}

// Batch arguments with defaults filled in and ranges checked.
static int batch_resolve(ajg_ctx *ctx, const ajg_batch_args *in, ajg_batch_args *out) {
  ajg_batch_args a;
  memset(&a, 0, sizeof(a));
  if (in) a = *in;
  if (a.max_slots <= 0) a.max_slots = 1 << 20;
  if (a.samples_per_item < 0) a.samples_per_item = 0;
  if (a.trav_blocks_per_cu <= 0) a.trav_blocks_per_cu = kTravBlocksPerCU;
  if (a.lds_stack <= 0) a.lds_stack = kLdsStack;
  if (a.iters_per_check <= 0) a.iters_per_check = kItersPerCheck;
  if (a.lds_stack != 4 && a.lds_stack != 8 && a.lds_stack != 16)
    return fail(ctx, AJG_ERR_INVALID, "batch: lds_stack must be 4, 8 or 16");
  if (a.trav_blocks_per_cu > 64 || a.iters_per_check > 64 || a.max_slots > (1 << 24))
    return fail(ctx, AJG_ERR_INVALID, "batch: trav_blocks_per_cu <= 64, iters_per_check <= 64, max_slots <= 2^24");
  if (a.pixel_tile != 0 && a.pixel_tile != 4 && a.pixel_tile != 8 && a.pixel_tile != 16)
    return fail(ctx, AJG_ERR_INVALID, "batch: pixel_tile must be 0, 4, 8 or 16");
  if (a.flags & AJG_RENDER_COUNT) return fail(ctx, AJG_ERR_INVALID, "batch: AJG_RENDER_COUNT is not supported");
  a.iters_per_check += a.iters_per_check & 1;  // even: a captured block ends on queue 0
  *out = a;
  return AJG_OK;
}

static int render_batch_impl(ajg_ctx *ctx, const ajg_job *jobs, uint32_t n_jobs,
                             const ajg_batch_args *bargs, ajg_batch_stats *stats,
                             float *job_done_ms, void *stream) {
  if (!ctx || !jobs || n_jobs == 0) return AJG_ERR_INVALID;
  if (int st = injected(ctx)) return st;
  if (n_jobs > 65535) return fail(ctx, AJG_ERR_INVALID, "batch: at most 65535 jobs");
  ajg_batch_args ba;
  if (int st = batch_resolve(ctx, bargs, &ba)) return st;
  const uint64_t target = (uint64_t)ba.max_slots;

  // ---- job table: each job's render parameters and scene, work prefix
  std::vector<BatchJob> hj(n_jobs);
  std::vector<uint32_t> begin(n_jobs + 1, 0u);
  std::vector<uint32_t> work_of(n_jobs);
  uint32_t max_depth = 0;
  uint64_t paths = 0, total = 0;
  int cu = 256;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const ajg_job &J = jobs[j];
    if (!J.scene || J.scene->ctx != ctx || !J.film.sum || !J.film.sum_sq || !J.film.count)
      return fail(ctx, AJG_ERR_INVALID, "batch: every job needs a scene of this context and a film");
    if (J.scene->insts)
      return fail(ctx, AJG_ERR_INVALID, "batch: instanced scenes render with ajg_render_pt");
    RenderParams P;
    // automatic chunking as a lone render with the same slot target picks
    if (int st = setup_params(ctx, &J.camera, &J.args, kItemsPerSlot * target,
                              (uint32_t)ba.samples_per_item, &P))
      return st;
    const uint32_t T = (uint32_t)ba.pixel_tile;
    const uint32_t Hp = P.n_pix / P.Wp;
    if (T && P.Wp % T == 0 && Hp % T == 0) {
      P.tile = T;
      P.tiles_x = P.Wp / T;
    }
    memset(&hj[j], 0, sizeof(BatchJob));
    hj[j].sc = dev_scene(J.scene);
    hj[j].P = P;
    work_of[j] = P.total_work;
    begin[j] = (uint32_t)total;
    total += P.total_work;
    if (total > 0xFFFFFFF0ull) return fail(ctx, AJG_ERR_INVALID, "batch: too much work");
    max_depth = std::max(max_depth, J.scene->accel->depth);
    paths += (uint64_t)P.n_pix * P.spp;
    cu = J.scene->cu_count;
  }
  begin[n_jobs] = (uint32_t)total;
  const uint32_t S = (uint32_t)std::min<uint64_t>(target, total);
  const int spill = spill_for_depth(max_depth, ba.lds_stack);
  if (!spill) return fail(ctx, AJG_ERR_INTERNAL, "batch: tree too deep for the traversal stack");

  BatchBuffers *bb = (BatchBuffers *)ctx->batch_cache;
  if (!bb) {
    bb = new BatchBuffers();
    ctx->batch_cache = bb;
    ctx->batch_free = batch_cache_free;
  }
  if (int st = ensure_buffers(ctx, &bb->rb, S, total)) return st;
  if (int st = ensure_batch_buffers(ctx, bb, S, n_jobs)) return st;
  if (int st = ensure_aux_stream(ctx, &bb->rb)) return st;
  RenderBuffers *rb = &bb->rb;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    hj[j].partial = rb->partial + 6 * (size_t)begin[j];
    hj[j].partial_n = rb->partial_n + begin[j];
    hj[j].per_sample = jobs[j].film.per_sample;
  }

  // The batch runs on its own stream (capturable), after the caller's work.
  hipStream_t us = (hipStream_t)stream;
  hipStream_t s = bb->stream;
  hipEvent_t ev0, ev1;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto done = [&](int code) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return code;
  };
#define AJG_BCHK(call)                                             \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) return done(hip_fail(ctx, e_, #call));   \
  } while (0)
  AJG_BCHK(hipStreamSynchronize(us));
  PathState ps = rb->ps;
  AJG_BCHK(hipMemcpyAsync(bb->jobs, hj.data(), sizeof(BatchJob) * n_jobs, hipMemcpyHostToDevice, s));
  AJG_BCHK(hipMemcpyAsync(bb->job_begin, begin.data(), 4 * ((size_t)n_jobs + 1), hipMemcpyHostToDevice, s));
  AJG_BCHK(hipMemcpyAsync(bb->remaining, work_of.data(), 4 * (size_t)n_jobs, hipMemcpyHostToDevice, s));
  // clears as in a single render (one launch per table)
  {
    Fills fl;
    fl.words(bb->clocks, 2ull * ((size_t)n_jobs + 1));
    fl.words(ps.work, S, kNone);
    fl.words(ps.sample, S);
    fl.words(ps.sample_end, S);
    fl.words(ps.bank, S);
    fl.words(ps.rng, 2ull * S);
    fl.words(ps.pend_flags, S);
    fl.words(ps.acc, 24ull * S);
    fl.words(ps.acc_n, 2ull * S);
    fl.words(ps.job, S);
    fl.words(ps.pend_job, S);
    fl.words(rb->partial, 6ull * total);
    fl.words(rb->partial_n, total);
    fl.words(rb->counters, kCounterWords);
    fl.bytes(rb->a_occ, aq_capacity(S));
    fl.words(rb->c_hits, 4ull * cq_capacity(S));
    fl.words(rb->qctr, kQctrWords);
    for (uint32_t j = 0; j < n_jobs; ++j) {
      if (fl.full()) AJG_BCHK(fl.launch(s));
      const ajg_film &f = jobs[j].film;
      const size_t np = hj[j].P.n_pix;
      fl.words(f.sum, 3ull * np);
      fl.words(f.sum_sq, 3ull * np);
      fl.words(f.count, np);
      if (f.per_sample) fl.words(f.per_sample, 3ull * np * hj[j].P.spp);
    }
    AJG_BCHK(fl.launch(s));
    // the first shade runs over every slot (list 0 holds them all)
    hipLaunchKernelGGL(k_lv_init, dim3(cdiv(S, 256)), dim3(256), 0, s, rb->lv[0], rb->lv_ei[0],
                       rb->qctr + kListOff,
                       S, lv_stripe_len(S));
  }

  uint32_t *counters = rb->counters;  // word layout as in a single render
  unsigned long long *totals = (unsigned long long *)(counters + 16);
  DevScene sc0;
  memset(&sc0, 0, sizeof(sc0));
  RenderParams P0;
  memset(&P0, 0, sizeof(P0));
  GuideParams G;
  memset(&G, 0, sizeof(G));
  BatchParams B;
  memset(&B, 0, sizeof(B));
  B.jobs = bb->jobs;
  B.job_begin = bb->job_begin;
  B.n_jobs = n_jobs;
  B.slots = S;
  B.total_work = (uint32_t)total;
  B.remaining = bb->remaining;
  B.done_clock = bb->clocks;
  Queues q;
  memset(&q, 0, sizeof(q));
  q.c_hits_in = rb->c_hits;
  q.c_inst_in = rb->c_inst;
  q.a_occ_in = rb->a_occ;
  q.live = counters + 4;
  q.work_counter = rb->qctr + kWorkOff;
  q.work_stripes = work_stripes_for(S);
  q.first_static = 1u;
  q.work_static = S;
  const uint32_t lv_len = lv_stripe_len(S);
  q.lv_total = counters + 5;
  q.lv_len = lv_len;
  const uint32_t shade_grid = cdiv(S, 256);
  // queue stripes sized for the most any job can append
  const uint32_t c_len = stripe_len(S, kCqPerSlot), a_len = stripe_len(S, kMaxLights + 1);
  uint32_t *const c_cursor = rb->qctr + 4 * kQWords;
  uint32_t *const a_cursor = c_cursor + kCursorStripes * kCursorStride;
  q.c_len = c_len;
  q.a_len = a_len;
  const uint32_t trav_grid = trav_grid_size((uint32_t)std::max(1, cu) * (uint32_t)ba.trav_blocks_per_cu);
  const int K = ba.iters_per_check;
  // One block of K iterations starting and ending on queue generation 0;
  // shadow rays are traced on the aux stream concurrently with the closest
  // queue (under capture the fork/join becomes two graph branches).
  hipError_t qerr = hipSuccess;
  auto q_ok = [&](hipError_t e) {
    if (qerr == hipSuccess) qerr = e;
  };
  auto enqueue_block = [&]() {
    int cur = 0;
    for (int k = 0; k < K; ++k) {
      const int nxt = cur ^ 1;
      q.c_rays_in = rb->c_rays[cur];
      q.c_rays_out = rb->c_rays[nxt];
      q.c_count_out = rb->qctr + nxt * kQWords;
      q.a_rays_in = rb->a_rays[cur];
      q.a_rays_out = rb->a_rays[nxt];
      q.a_count_out = rb->qctr + (2 + nxt) * kQWords;
      q.lv_in = rb->lv[cur];
      q.lv_ei_in = rb->lv_ei[cur];
      q.lv_q = {rb->qctr + kListOff + cur * kQWords, lv_len};
      q.lv_out = rb->lv[nxt];
      q.lv_ei_out = rb->lv_ei[nxt];
      q.lv_count_out = rb->qctr + kListOff + nxt * kQWords;
      const QIn cq_in = {rb->qctr + nxt * kQWords, c_len};
      const QIn aq_in = {rb->qctr + (2 + nxt) * kQWords, a_len};
      B.c_job_out = bb->c_job[nxt];
      B.a_job_out = bb->a_job[nxt];
      B.c_job_in = bb->c_job[cur];
      B.a_job_in = bb->a_job[cur];
      hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(64), 0, s, counters, rb->qctr, cur, nxt,
                         totals);
      hipLaunchKernelGGL((k_shade_pend<false, true>), dim3(shade_grid), dim3(256), 0, s, sc0, P0, ps, q,
                         G, B);
      hipLaunchKernelGGL((k_shade<false, true>), dim3(shade_grid), dim3(256), 0, s, sc0, P0, ps, q, G,
                         B);
      q_ok(hipEventRecord(rb->ev_fork, s));
      q_ok(hipStreamWaitEvent(rb->aux, rb->ev_fork, 0));
      if (spill == kSpillS) {
        launch_pq_b<kSpillS>(ba.lds_stack, true, bb->jobs, bb->c_job[nxt], rb->c_rays[nxt], rb->c_hits,
                        cq_in, c_cursor, trav_grid, s);
        launch_pq_b<kSpillS>(ba.lds_stack, false, bb->jobs, bb->a_job[nxt], rb->a_rays[nxt], rb->a_occ,
                        aq_in, a_cursor, trav_grid, rb->aux);
      } else {
        launch_pq_b<kSpillL>(ba.lds_stack, true, bb->jobs, bb->c_job[nxt], rb->c_rays[nxt], rb->c_hits,
                          cq_in, c_cursor, trav_grid, s);
        launch_pq_b<kSpillL>(ba.lds_stack, false, bb->jobs, bb->a_job[nxt], rb->a_rays[nxt], rb->a_occ,
                          aq_in, a_cursor, trav_grid, rb->aux);
      }
      q_ok(hipEventRecord(rb->ev_join, rb->aux));
      q_ok(hipStreamWaitEvent(s, rb->ev_join, 0));
      cur = nxt;
    }
  };
  if (ba.use_graph) {
    AJG_BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_block();
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess) return done(hip_fail(ctx, e, "hipStreamEndCapture"));
    if (qerr != hipSuccess) return done(hip_fail(ctx, qerr, "batch stream fork/join"));
    AJG_BCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, s, bb->clocks + n_jobs);
  AJG_BCHK(hipEventRecord(ev0, s));
  uint32_t iters = 0, graph_launches = 0;
  uint32_t hc[8];
  for (;;) {
    if (exec) {
      AJG_BCHK(hipGraphLaunch(exec, s));
      ++graph_launches;
    } else {
      enqueue_block();
      if (qerr != hipSuccess) return done(hip_fail(ctx, qerr, "batch stream fork/join"));
    }
    iters += (uint32_t)K;
    AJG_BCHK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    AJG_BCHK(hipStreamSynchronize(s));
    if (cancelled(ctx)) return done(fail(ctx, AJG_ERR_CANCELLED, "batch cancelled"));
    if (hc[4] == 0) break;
    if (iters > 400000) return done(fail(ctx, AJG_ERR_INTERNAL, "batch: iteration limit"));
  }
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const RenderParams &P = hj[j].P;
    const ajg_film &f = jobs[j].film;
    if (P.n_chunks > 1)
      hipLaunchKernelGGL(k_reduce_chunks, dim3(cdiv(P.n_pix, 256)), dim3(256), 0, s,
                         (const float *)hj[j].partial, (const uint32_t *)hj[j].partial_n, P.n_pix,
                         P.n_chunks, f.sum, f.sum_sq, f.count);
    else
      hipLaunchKernelGGL(k_unpack_single, dim3(cdiv(P.n_pix, 256)), dim3(256), 0, s,
                         (const float *)hj[j].partial, (const uint32_t *)hj[j].partial_n, P.n_pix,
                         f.sum, f.sum_sq, f.count);
  }
  AJG_BCHK(hipEventRecord(ev1, s));
  unsigned long long tot[2] = {0, 0};
  std::vector<unsigned long long> clk(n_jobs + 1);
  AJG_BCHK(hipMemcpyAsync(tot, totals, 16, hipMemcpyDeviceToHost, s));
  AJG_BCHK(hipMemcpyAsync(clk.data(), bb->clocks, 8 * ((size_t)n_jobs + 1), hipMemcpyDeviceToHost, s));
  AJG_BCHK(hipStreamSynchronize(s));
  AJG_BCHK(hipGetLastError());
  if (job_done_ms) {
    int khz = 0;
    AJG_BCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    const double per_ms = khz > 0 ? (double)khz : 1e5;
    for (uint32_t j = 0; j < n_jobs; ++j)
      job_done_ms[j] = clk[j] >= clk[n_jobs] ? (float)((double)(clk[j] - clk[n_jobs]) / per_ms) : -1.f;
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    stats->ms_total = ms;
    stats->iterations = iters;
    stats->slots = S;
    stats->graph_launches = graph_launches;
    stats->paths = paths;
    stats->closest_queries = tot[0];
    stats->any_queries = tot[1];
  }
#undef AJG_BCHK
  return done(AJG_OK);
}

extern "C" {

int ajg_render_pt(ajg_scene *sc, const ajg_camera *camera, const ajg_pt_args *args,
                  const ajg_film *film, ajg_render_stats *stats, void *stream) {
  return render_impl(sc, camera, args, nullptr, nullptr, film, stats, nullptr, stream);
}

int ajg_render_pt_guided(ajg_scene *sc, const ajg_camera *camera, const ajg_pt_args *args,
                         ajg_guide *guide, const ajg_guide_args *gargs, const ajg_film *film,
                         ajg_render_stats *stats, ajg_guide_stats *gstats, void *stream) {
  if (!sc || !args || !guide || !gargs) return AJG_ERR_INVALID;
  ajg_guide_stats total;
  memset(&total, 0, sizeof(total));
  if (gargs->warmup_passes) {
    if (gargs->warmup_spp == 0 || gargs->batch == 0 || !(gargs->lr > 0.f))
      return fail(sc->ctx, AJG_ERR_INVALID, "guided render: warm-up needs spp, batch and lr");
    // Warm-up passes render into the caller's film (cleared again by the
    // final render) from seed_mode 1 streams whose base no render uses.
    ajg_pt_args wa = *args;
    wa.spp = (int32_t)gargs->warmup_spp;
    wa.seed_mode = 1;
    ajg_guide_args wg = *gargs;
    wg.train = 1;
    wg.steps_per_batch = gargs->warmup_steps_per_batch;
    wg.warmup_passes = 0;
    for (uint32_t k = 0; k < gargs->warmup_passes; ++k) {
      wa.seed_base = (args->seed_base ^ 0x5bd1e995u) + 0x9e3779b9u * (k + 1);
      ajg_render_stats ws;
      ajg_guide_stats wgs;
      if (int st = render_impl(sc, camera, &wa, guide, &wg, film, &ws, &wgs, stream)) return st;
      total.train_steps += wgs.train_steps;
      total.records += wgs.records;
      total.ms_infer += wgs.ms_infer;
      total.ms_train += wgs.ms_train;
      total.ms_warmup += ws.ms_total;
      if (wgs.train_steps) total.last_loss = wgs.last_loss;
    }
    total.warmup_steps = total.train_steps;
  }
  ajg_guide_stats fs;
  if (int st = render_impl(sc, camera, args, guide, gargs, film, stats, &fs, stream)) return st;
  if (gstats) {
    *gstats = total;
    gstats->train_steps += fs.train_steps;
    gstats->records += fs.records;
    gstats->ms_infer += fs.ms_infer;
    gstats->ms_train += fs.ms_train;
    if (fs.train_steps) gstats->last_loss = fs.last_loss;
  }
  return AJG_OK;
}

int ajg_render_batch(ajg_ctx *ctx, const ajg_job *jobs, uint32_t n_jobs,
                     const ajg_batch_args *bargs, ajg_batch_stats *stats, float *job_done_ms,
                     void *stream) {
  return render_batch_impl(ctx, jobs, n_jobs, bargs, stats, job_done_ms, stream);
}

int ajg_batch_autotune(ajg_ctx *ctx, const ajg_job *jobs, uint32_t n_jobs, uint32_t trial_spp,
                       ajg_batch_args *inout, float *best_paths_per_s) {
  if (!ctx || !jobs || n_jobs == 0 || !inout || trial_spp == 0) return AJG_ERR_INVALID;
  ajg_batch_args cur;
  if (int st = batch_resolve(ctx, inout, &cur)) return st;
  std::vector<ajg_job> tj(jobs, jobs + n_jobs);
  for (auto &j : tj) j.args.spp = std::min<int32_t>(j.args.spp, (int32_t)trial_spp);
  auto rate = [&](const ajg_batch_args &a, double *r) {
    ajg_batch_stats st;
    if (int e = render_batch_impl(ctx, tj.data(), n_jobs, &a, &st, nullptr, nullptr)) return e;
    *r = st.ms_total > 0.f ? (double)st.paths / ((double)st.ms_total * 1e-3) : 0.0;
    return (int)AJG_OK;
  };
  double best = 0.0;
  if (int st = rate(cur, &best)) return st;  // first run also allocates
  if (int st = rate(cur, &best)) return st;
  // Coordinate descent, one pass: queue size (slots in flight), work-item
  // shape (samples per item, pixel tile order), traversal stack in LDS,
  // traversal grid, iterations per host check.
  struct Knob {
    int32_t ajg_batch_args::*field;
    std::vector<int32_t> values;
  };
  const Knob knobs[] = {
      {&ajg_batch_args::max_slots, {1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23}},
      {&ajg_batch_args::samples_per_item, {0, 1, 2, 4, 8}},
      {&ajg_batch_args::pixel_tile, {0, 4, 8, 16}},
      {&ajg_batch_args::lds_stack, {4, 8, 16}},
      {&ajg_batch_args::trav_blocks_per_cu, {4, 6, 8, 12, 16}},
      {&ajg_batch_args::iters_per_check, {2, 4, 8}},
  };
  bool any_chunked = false;  // samples_per_item only shapes seed_mode 1 jobs
  for (const auto &j : tj) any_chunked |= j.args.seed_mode != 0;
  for (const Knob &k : knobs) {
    if (k.field == &ajg_batch_args::samples_per_item && !any_chunked) continue;
    const int32_t keep = cur.*(k.field);
    int32_t pick = keep;
    for (int32_t v : k.values) {
      if (v == keep) continue;
      ajg_batch_args a = cur;
      a.*(k.field) = v;
      double r = 0.0;
      if (int st = rate(a, &r)) return st;
      if (r > best) {
        best = r;
        pick = v;
      }
    }
    cur.*(k.field) = pick;
  }
  *inout = cur;
  if (best_paths_per_s) *best_paths_per_s = (float)best;
  return AJG_OK;
}

int ajg_film_resolve_filter(ajg_ctx *ctx, uint32_t w, uint32_t h, uint32_t tile_size,
                            const ajg_filter *filter, const float *sum, const uint32_t *count,
                            float *rgb, void *stream) {
  if (!ctx || !filter || !sum || !count || !rgb || w == 0 || h == 0 || tile_size == 0)
    return AJG_ERR_INVALID;
  if (int st = injected(ctx)) return st;
  if (filter->type != AJG_FILTER_BOX && filter->type != AJG_FILTER_GAUSSIAN)
    return fail(ctx, AJG_ERR_INVALID, "film: unknown filter type");
  FilterAxis ax[2];
  for (int a = 0; a < 2; ++a) {
    const float r = filter->radius[a];
    if (!(r > 0.f) || r > 7.5f) return fail(ctx, AJG_ERR_INVALID, "film: filter radius must be in (0, 7.5]");
    // targets of pixel p: [ceil(p - 0.5 - r), floor(p - 0.5 + r)] (film.cpp:30-33)
    ax[a].omin = (int)std::ceil(-0.5f - r);
    ax[a].omax = (int)std::floor(-0.5f + r);
    if (ax[a].omax - ax[a].omin + 1 > 16) return fail(ctx, AJG_ERR_INVALID, "film: filter too wide");
    // Filter::Eval per axis at o + 0.5, in float with the host's expf
    // (box_filter.cpp:35; gaussian_filter.cpp:38-45)
    const float ev = std::exp(-filter->alpha * r * r);
    for (int o = ax[a].omin; o <= ax[a].omax; ++o) {
      const float d = (float)o + 0.5f;
      ax[a].w[o - ax[a].omin] =
          filter->type == AJG_FILTER_BOX ? 1.f : std::max(0.f, (float)std::exp(-filter->alpha * d * d) - ev);
    }
  }
  const uint32_t Wp = (w + tile_size - 1) / tile_size * tile_size;
  const uint32_t Hp = (h + tile_size - 1) / tile_size * tile_size;
  hipLaunchKernelGGL(k_resolve_filter, dim3(cdiv((uint64_t)w * h, 256)), dim3(256), 0, (hipStream_t)stream,
                     w, h, Wp, Hp, ax[0], ax[1], sum, count, rgb);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(ctx, e, "k_resolve_filter");
}

int ajg_image_srgb8(ajg_ctx *ctx, uint32_t w, uint32_t h, const float *rgb, int32_t flip,
                    uint8_t *out, void *stream) {
  if (!ctx || !rgb || !out || w == 0 || h == 0) return AJG_ERR_INVALID;
  if (int st = injected(ctx)) return st;
  hipLaunchKernelGGL(k_srgb8, dim3(cdiv((uint64_t)w * h, 256)), dim3(256), 0, (hipStream_t)stream, w, h,
                     rgb, (int)flip, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(ctx, e, "k_srgb8");
}

int ajg_film_resolve(ajg_ctx *ctx, uint32_t w, uint32_t h, uint32_t tile_size, const float *sum,
                     const uint32_t *count, float *rgb, void *stream) {
  if (!ctx || !sum || !count || !rgb || w == 0 || h == 0 || tile_size == 0) return AJG_ERR_INVALID;
  if (int st = injected(ctx)) return st;
  const uint32_t Wp = (w + tile_size - 1) / tile_size * tile_size;
  const uint32_t Hp = (h + tile_size - 1) / tile_size * tile_size;
  hipLaunchKernelGGL(k_resolve, dim3(cdiv((uint64_t)w * h, 256)), dim3(256), 0, (hipStream_t)stream, w, h,
                     Wp, Hp, sum, count, rgb);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(ctx, e, "k_resolve");
}

}  // extern "C"
