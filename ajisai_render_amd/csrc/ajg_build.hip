// ajg_build.hip — PLOC BVH construction on gfx950.
//
// Restates PLOCAggregate::Build (ajisai/core/aggregate/ploc_aggregate.cpp:
// 252-327) for the device, producing the node array the reference produces
// bit for bit:
//   Morton codes   MortonEncoder (:56-82) on triangle AABB centres (:530-541)
//   sort           RadixSort<10> (:84-132) — any stable sort yields the same
//                  permutation; here 4 passes of an 8-bit LSD radix sort with
//                  LDS digit histograms and wave-ballot stable ranking
//   leaves         (:301-311)
//   Cluster        nearest neighbour (:400-459, ties to the lowest index),
//                  mutual-pair flags and inclusive scan (:461-471), emit with
//                  the reference's compacted layout (:473-508)
// The reference ping-pongs whole node arrays and copies finished nodes
// forward (:503-508); here finished children are written once into the final
// array and only live clusters ping-pong, which leaves the same final array.
// Iterations with more than kTail live clusters run as kernels on the whole
// chip; the tail runs in one LDS-resident workgroup.
// Finally the binary tree is re-laid into 64-byte fat nodes (both child boxes
// in the parent) for traversal.
#include <algorithm>

#include "ajg_internal.h"

namespace ajg {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t morton_split(uint32_t x) {
  // MortonSplit<uint32_t> (ploc_aggregate.cpp:37-49)
  uint32_t mask = 0xFFFFFFFFu >> 16;
  x &= mask;
#pragma unroll
  for (unsigned i = 4, n = 16; i > 0; --i, n >>= 1) {
    mask = (mask | (mask << n)) & ~(mask << (n / 2));
    x = (x | (x << n)) & mask;
  }
  return x;
}

// float -> u32 as baseline x86-64 converts a non-negative float through a
// 64-bit cvttss2si: NaN, +inf and values >= 2^63 give 0.
__device__ __forceinline__ uint32_t f2u_trunc(float g) {
  if (!(g < 9.2233720e18f)) return 0u;
  return (uint32_t)(uint64_t)g;
}

__device__ __forceinline__ f3 tri_vertex(const float *tri, uint32_t t, int k) {
  const float *p = tri + 9 * (size_t)t + 3 * k;
  return mk3(p[0], p[1], p[2]);
}

// ------------------------------------------------------------- scene box
__global__ void __launch_bounds__(256)
k_box_partial(const float *tri, uint32_t n_vert, float *partial) {
  __shared__ float smn[3][256], smx[3][256];
  box3 b = box_empty();
  for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n_vert; v += gridDim.x * 256) {
    f3 p = mk3(tri[3 * (size_t)v], tri[3 * (size_t)v + 1], tri[3 * (size_t)v + 2]);
    b.mn = vmin(b.mn, p);
    b.mx = vmax(b.mx, p);
  }
  smn[0][threadIdx.x] = b.mn.x;
  smn[1][threadIdx.x] = b.mn.y;
  smn[2][threadIdx.x] = b.mn.z;
  smx[0][threadIdx.x] = b.mx.x;
  smx[1][threadIdx.x] = b.mx.y;
  smx[2][threadIdx.x] = b.mx.z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        smn[k][threadIdx.x] = fmin_r(smn[k][threadIdx.x], smn[k][threadIdx.x + s]);
        smx[k][threadIdx.x] = fmax_r(smx[k][threadIdx.x], smx[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) {
      partial[6 * blockIdx.x + k] = smn[k][0];
      partial[6 * blockIdx.x + 3 + k] = smx[k][0];
    }
}

// ----------------------------------------------------------------- morton
__global__ void __launch_bounds__(256)
k_morton(const float *tri, uint32_t n, f3 w2g, f3 off, uint32_t *codes,
         uint32_t *vals) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  box3 b = tri_box(tri_vertex(tri, i, 0), tri_vertex(tri, i, 1), tri_vertex(tri, i, 2));
  f3 c = box_center(b);
  f3 g = add(mul(c, w2g), off);
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float gk = comp(g, k);
    uint32_t u = f2u_trunc(gk < 0.f ? 0.f : gk);
    q[k] = u < 1023u ? u : 1023u;
  }
  codes[i] = morton_split(q[0]) | (morton_split(q[1]) << 1) | (morton_split(q[2]) << 2);
  vals[i] = i;
}

// ------------------------------------------------------------- radix sort
static constexpr int kSortThreads = 256;
static constexpr int kSortItems = 16;
static constexpr uint32_t kSortTile = kSortThreads * kSortItems;

__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const uint32_t *keys, uint32_t n, int shift, uint32_t *counts,
             uint32_t nblocks) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    uint32_t idx = base + r * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// Stable scatter: elements of a tile are ranked in their original order —
// within a wave by a ballot-built peer mask, across waves and rounds by
// per-digit counters in LDS.
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const uint32_t *keys, const uint32_t *vals, uint32_t *keys_out,
                uint32_t *vals_out, uint32_t n, int shift,
                const uint32_t *offsets, uint32_t nblocks) {
  __shared__ uint32_t base_off[256];
  __shared__ uint32_t wave_cnt[kSortThreads / 64][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  base_off[t] = offsets[t * nblocks + blockIdx.x];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t tile = blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    for (int k = 0; k < kSortThreads / 64; ++k) wave_cnt[k][t] = 0;
    __syncthreads();
    const uint32_t idx = tile + r * kSortThreads + t;
    const bool valid = idx < n;
    uint32_t key = valid ? keys[idx] : 0u;
    uint32_t val = valid ? vals[idx] : 0u;
    uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    const uint32_t cnt = __popcll(peers);
    if (valid && rank == 0) wave_cnt[w][d] = cnt;
    __syncthreads();
    if (valid) {
      uint32_t o = base_off[d] + rank;
      for (int k = 0; k < w; ++k) o += wave_cnt[k][d];
      keys_out[o] = key;
      vals_out[o] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int k = 0; k < kSortThreads / 64; ++k) add += wave_cnt[k][t];
    base_off[t] += add;
    __syncthreads();
  }
}

// ------------------------------------------------------------ PLOC kernels
__global__ void __launch_bounds__(256)
k_leaf_init(const float *tri, const uint32_t *perm, uint32_t n, PlocNode *live,
            float4 *leaf_tris) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t p = perm[i];
  f3 a = tri_vertex(tri, p, 0), b = tri_vertex(tri, p, 1), c = tri_vertex(tri, p, 2);
  PlocNode nd;
  node_set_box(nd, tri_box(a, b, c));
  nd.count = 1;
  nd.first = i;
  live[i] = nd;
  leaf_tris[3 * (size_t)i + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(p));
  leaf_tris[3 * (size_t)i + 1] = make_float4(b.x, b.y, b.z, 0.f);
  leaf_tris[3 * (size_t)i + 2] = make_float4(c.x, c.y, c.z, 0.f);
}

// Distance of a pair: Join(lower index, higher index).area(), as the
// reference's distance cache computes it (ploc_aggregate.cpp:416-442).
__device__ __forceinline__ float pair_dist(const box3 &lo, const box3 &hi) {
  return box_area(box_join(lo, hi));
}

static constexpr int kNNThreads = 256;
static constexpr int kMaxRadius = 64;

__global__ void __launch_bounds__(kNNThreads)
k_nn(const PlocNode *live, uint32_t n, uint32_t r, uint32_t *nb) {
  __shared__ float sb[6][kNNThreads + 2 * kMaxRadius];
  const int64_t k0 = (int64_t)blockIdx.x * kNNThreads;
  const int64_t lo = k0 - (int64_t)r;
  const int span = kNNThreads + 2 * (int)r;
  for (int s = threadIdx.x; s < span; s += kNNThreads) {
    int64_t j = lo + s;
    if (j >= 0 && j < (int64_t)n) {
      const PlocNode &nd = live[j];
      sb[0][s] = nd.mn[0];
      sb[1][s] = nd.mn[1];
      sb[2][s] = nd.mn[2];
      sb[3][s] = nd.mx[0];
      sb[4][s] = nd.mx[1];
      sb[5][s] = nd.mx[2];
    }
  }
  __syncthreads();
  const uint32_t k = (uint32_t)(k0 + threadIdx.x);
  if (k >= n) return;
  const uint32_t sbeg = k > r ? k - r : 0u;
  const uint32_t send = (k + r + 1 < n) ? k + r + 1 : n;
  const int sk = (int)(k - lo);
  box3 bk;
  bk.mn = mk3(sb[0][sk], sb[1][sk], sb[2][sk]);
  bk.mx = mk3(sb[3][sk], sb[4][sk], sb[5][sk]);
  float best = 3.40282347e+38f;
  uint32_t best_j = kNone;
  for (uint32_t j = sbeg; j < send; ++j) {
    if (j == k) continue;
    const int sj = (int)((int64_t)j - lo);
    box3 bj;
    bj.mn = mk3(sb[0][sj], sb[1][sj], sb[2][sj]);
    bj.mx = mk3(sb[3][sj], sb[4][sj], sb[5][sj]);
    const float d = j < k ? pair_dist(bj, bk) : pair_dist(bk, bj);
    if (d < best) {
      best = d;
      best_j = j;
    }
  }
  nb[k] = best_j;
}

__global__ void __launch_bounds__(256)
k_flags(const uint32_t *nb, uint32_t n, uint32_t *m) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  uint32_t j = nb[k];
  m[k] = (j != kNone && nb[j] == k && k < j) ? 1u : 0u;
}

// Emit (ploc_aggregate.cpp:473-501) in local indices: survivors and merged
// nodes go to live_out (position P - next_begin), children to the final
// array at their global positions.
__global__ void __launch_bounds__(256)
k_emit(const PlocNode *live_in, const uint32_t *nb, const uint32_t *m, uint32_t n,
       uint32_t merged_count, uint32_t end_global, PlocNode *live_out,
       PlocNode *final_nodes) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const uint32_t children_begin = end_global - 2 * merged_count;
  uint32_t j = nb[k];
  bool mutual = j != kNone && nb[j] == k;
  if (mutual) {
    if (k < j) {
      uint32_t first = children_begin + (m[k] - 1) * 2;
      PlocNode a = live_in[k], b = live_in[j];
      final_nodes[first + 0] = a;
      final_nodes[first + 1] = b;
      PlocNode u;
      node_set_box(u, box_join(node_box(b), node_box(a)));
      u.count = 0;
      u.first = first;
      live_out[j - m[j]] = u;
    }
  } else {
    live_out[k - m[k]] = live_in[k];
  }
}

// Tail: all remaining iterations in one workgroup with LDS-resident state.
static constexpr int kTail = 512;

__global__ void __launch_bounds__(kTail)
k_tail(const PlocNode *live_in, uint32_t n0, uint32_t end0, uint32_t r,
       PlocNode *final_nodes, uint32_t *iters_out, uint32_t *status) {
  __shared__ float bx[2][6][kTail];
  __shared__ uint32_t cnt[2][kTail], fst[2][kTail];
  __shared__ uint32_t nb[kTail], m[kTail];
  __shared__ uint32_t wsum[kTail / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if ((uint32_t)t < n0) {
    const PlocNode &nd = live_in[t];
    for (int c = 0; c < 3; ++c) {
      bx[0][c][t] = nd.mn[c];
      bx[0][3 + c][t] = nd.mx[c];
    }
    cnt[0][t] = nd.count;
    fst[0][t] = nd.first;
  }
  __syncthreads();
  uint32_t n = n0, end = end0;
  int cur = 0;
  uint32_t iters = 0;
  while (n > 1) {
    box3 bk;
    if ((uint32_t)t < n) {
      bk.mn = mk3(bx[cur][0][t], bx[cur][1][t], bx[cur][2][t]);
      bk.mx = mk3(bx[cur][3][t], bx[cur][4][t], bx[cur][5][t]);
      const uint32_t k = t;
      const uint32_t sbeg = k > r ? k - r : 0u;
      const uint32_t send = (k + r + 1 < n) ? k + r + 1 : n;
      float best = 3.40282347e+38f;
      uint32_t best_j = kNone;
      for (uint32_t j = sbeg; j < send; ++j) {
        if (j == k) continue;
        box3 bj;
        bj.mn = mk3(bx[cur][0][j], bx[cur][1][j], bx[cur][2][j]);
        bj.mx = mk3(bx[cur][3][j], bx[cur][4][j], bx[cur][5][j]);
        const float d = j < k ? pair_dist(bj, bk) : pair_dist(bk, bj);
        if (d < best) {
          best = d;
          best_j = j;
        }
      }
      nb[t] = best_j;
    }
    __syncthreads();
    uint32_t flag = 0;
    if ((uint32_t)t < n) {
      uint32_t j = nb[t];
      flag = (j != kNone && nb[j] == (uint32_t)t && (uint32_t)t < j) ? 1u : 0u;
    }
    // block inclusive scan of flag
    uint32_t x = flag;
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t woff = 0;
    for (int i = 0; i < w; ++i) woff += wsum[i];
    m[t] = woff + x;
    __syncthreads();
    const uint32_t mc = m[n - 1];
    if (mc == 0) {
      if (t == 0) *status = 1u;
      return;
    }
    const uint32_t children_begin = end - 2 * mc;
    const int nxt = cur ^ 1;
    if ((uint32_t)t < n) {
      uint32_t j = nb[t];
      bool mutual = j != kNone && nb[j] == (uint32_t)t;
      if (mutual) {
        if ((uint32_t)t < j) {
          uint32_t first = children_begin + (m[t] - 1) * 2;
          PlocNode a, b;
          for (int c = 0; c < 3; ++c) {
            a.mn[c] = bx[cur][c][t];
            a.mx[c] = bx[cur][3 + c][t];
            b.mn[c] = bx[cur][c][j];
            b.mx[c] = bx[cur][3 + c][j];
          }
          a.count = cnt[cur][t];
          a.first = fst[cur][t];
          b.count = cnt[cur][j];
          b.first = fst[cur][j];
          final_nodes[first + 0] = a;
          final_nodes[first + 1] = b;
          box3 u = box_join(node_box(b), node_box(a));
          uint32_t P = j - m[j];
          bx[nxt][0][P] = u.mn.x;
          bx[nxt][1][P] = u.mn.y;
          bx[nxt][2][P] = u.mn.z;
          bx[nxt][3][P] = u.mx.x;
          bx[nxt][4][P] = u.mx.y;
          bx[nxt][5][P] = u.mx.z;
          cnt[nxt][P] = 0;
          fst[nxt][P] = first;
        }
      } else {
        uint32_t P = t - m[t];
        for (int c = 0; c < 6; ++c) bx[nxt][c][P] = bx[cur][c][t];
        cnt[nxt][P] = cnt[cur][t];
        fst[nxt][P] = fst[cur][t];
      }
    }
    __syncthreads();
    end = children_begin;
    n = n - mc;
    cur = nxt;
    ++iters;
  }
  if (t == 0) {
    PlocNode root;
    for (int c = 0; c < 3; ++c) {
      root.mn[c] = bx[cur][c][0];
      root.mx[c] = bx[cur][3 + c][0];
    }
    root.count = cnt[cur][0];
    root.first = fst[cur][0];
    final_nodes[end - 1] = root;
    *iters_out = iters;
    *status = (end == 1) ? 0u : 2u;
  }
}

// ---------------------------------------------------------------- finalize
__global__ void __launch_bounds__(256)
k_mark_interior(const PlocNode *fin, uint32_t nn, uint32_t *flag) {
  uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f < nn) flag[f] = fin[f].count == 0 ? 1u : 0u;
}

__device__ __forceinline__ int child_ref(const PlocNode *fin, const uint32_t *fat_idx,
                                         uint32_t c) {
  const PlocNode &nd = fin[c];
  return nd.count != 0 ? ~(int)nd.first : (int)fat_idx[c];
}

__global__ void __launch_bounds__(256)
k_build_fat(const PlocNode *fin, const uint32_t *fat_idx, uint32_t nn, FatNode *fat,
            int *parent_of_fat, int *parent_of_leaf) {
  uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nn) return;
  const PlocNode nd = fin[f];
  if (nd.count != 0) return;
  const uint32_t c0 = nd.first, c1 = nd.first + 1;
  const PlocNode a = fin[c0], b = fin[c1];
  FatNode o;
  o.a = make_float4(a.mn[0], a.mn[1], a.mn[2], a.mx[0]);
  o.b = make_float4(a.mx[1], a.mx[2], b.mn[0], b.mn[1]);
  o.c = make_float4(b.mn[2], b.mx[0], b.mx[1], b.mx[2]);
  const int r0 = child_ref(fin, fat_idx, c0), r1 = child_ref(fin, fat_idx, c1);
  o.d = make_int4(r0, r1, -1, 0);
  const int me = (int)fat_idx[f];
  fat[me] = o;
  if (r0 >= 0) parent_of_fat[r0] = me; else parent_of_leaf[~r0] = me;
  if (r1 >= 0) parent_of_fat[r1] = me; else parent_of_leaf[~r1] = me;
}

// Quantisation grid of one axis of a node: step 2^e and origin o (a
// multiple of 2^e) such that [lo, hi] lies within o + [0, 253] 2^e and every
// plane o + q 2^e, q = 0..255, is an exact float: 2^e is at least the ulp of
// the largest magnitude on the axis (e >= exponent - 23), so |o + q 2^e| /
// 2^e stays below 2^24 + 256 on a grid of 2^e. e in [-126, 127].
__device__ void qaxis(float lo, float hi, float *origin, int *e_out) {
  const float M = fmaxf(fabsf(lo), fabsf(hi));
  int e = -126;
  const int em = (int)((__float_as_uint(M) >> 23) & 0xffu) - 127;
  if (M > 0.f) e = max(e, em - 23);
  const double ext = (double)hi - (double)lo;
  while (e < 127 && ldexp(253.0, e) < ext) ++e;
  const double step = ldexp(1.0, e);
  *origin = (float)(floor((double)lo / step) * step);
  *e_out = e;
}

// Outward rounding of one plane: the largest q with o + q s <= v (lower
// planes) or the smallest with o + q s >= v (upper planes), checked in the
// float decode itself.
__device__ uint32_t q_down(float v, float o, float s) {
  const double t = floor(((double)v - (double)o) / (double)s);
  int q = (int)fmin(fmax(t, 0.0), 255.0);
  while (q > 0 && fmaf((float)q, s, o) > v) --q;
  return (uint32_t)q;
}
__device__ uint32_t q_up(float v, float o, float s) {
  const double t = ceil(((double)v - (double)o) / (double)s);
  int q = (int)fmin(fmax(t, 0.0), 255.0);
  while (q < 255 && fmaf((float)q, s, o) < v) ++q;
  return (uint32_t)q;
}

// The six plane bytes of box (mn, mx) in the node's grid; *bad is raised
// if the decoded box does not contain it (non-finite coordinates).
__device__ void q_box(const float *mn, const float *mx, const float *o, const float *st, uint8_t *out,
                      uint32_t *bad) {
  for (int a = 0; a < 3; ++a) {
    const uint32_t lo = q_down(mn[a], o[a], st[a]), hi = q_up(mx[a], o[a], st[a]);
    out[a] = (uint8_t)lo;
    out[3 + a] = (uint8_t)hi;
    if (!(fmaf((float)lo, st[a], o[a]) <= mn[a] && fmaf((float)hi, st[a], o[a]) >= mx[a])) atomicAdd(bad, 1u);
  }
}

// FatNode child boxes: c0 = (a.xyz, (a.w, b.x, b.y)), c1 = ((b.z, b.w, c.x), c.yzw).
__device__ void fat_boxes(const FatNode &f, float (*mn)[3], float (*mx)[3]) {
  const float v[12] = {f.a.x, f.a.y, f.a.z, f.a.w, f.b.x, f.b.y, f.b.z, f.b.w, f.c.x, f.c.y, f.c.z, f.c.w};
  for (int k = 0; k < 2; ++k)
    for (int a = 0; a < 3; ++a) {
      mn[k][a] = v[6 * k + a];
      mx[k][a] = v[6 * k + 3 + a];
    }
}

__global__ void __launch_bounds__(256)
k_build_wide(const FatNode *fat, uint32_t n_int, WideNode *wide, uint32_t *bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_int) return;
  const FatNode f = fat[i];
  float cmn[2][3], cmx[2][3];
  fat_boxes(f, cmn, cmx);
  // slots: the grandchildren in left-first order (a leaf child stands for itself)
  float mn[4][3], mx[4][3];
  int ref[4] = {kNoRef, kNoRef, kNoRef, kNoRef};
  int ns = 0;
  const int child[2] = {f.d.x, f.d.y};
  for (int c = 0; c < 2; ++c) {
    if (child[c] >= 0) {
      const FatNode g = fat[child[c]];
      float gmn[2][3], gmx[2][3];
      fat_boxes(g, gmn, gmx);
      const int gr[2] = {g.d.x, g.d.y};
      for (int k = 0; k < 2; ++k, ++ns) {
        for (int a = 0; a < 3; ++a) {
          mn[ns][a] = gmn[k][a];
          mx[ns][a] = gmx[k][a];
        }
        ref[ns] = gr[k];
      }
    } else {
      for (int a = 0; a < 3; ++a) {
        mn[ns][a] = cmn[c][a];
        mx[ns][a] = cmx[c][a];
      }
      ref[ns++] = child[c];
    }
  }
  // grid over the node's box (every slot box lies inside it)
  float o[3], st[3];
  uint32_t ebits = 0;
  for (int a = 0; a < 3; ++a) {
    int e;
    qaxis(fmin_r(cmn[0][a], cmn[1][a]), fmax_r(cmx[0][a], cmx[1][a]), &o[a], &e);
    st[a] = ldexpf(1.f, e);
    ebits |= (uint32_t)(e + 127) << (8 * a);
  }
  // plane-major words: word j (lo x, lo y, lo z, hi x, hi y, hi z) holds
  // slot k's plane in byte k; empty slots stay 0
  uint8_t q[4][6] = {};
  for (int k = 0; k < ns; ++k) q_box(mn[k], mx[k], o, st, q[k], bad);
  uint32_t wd[6];
  for (int j = 0; j < 6; ++j)
    wd[j] = (uint32_t)q[0][j] | (uint32_t)q[1][j] << 8 | (uint32_t)q[2][j] << 16 |
            (uint32_t)q[3][j] << 24;
  WideNode w;
  w.o = make_float4(o[0], o[1], o[2], __uint_as_float(ebits));
  w.qa = make_uint4(wd[0], wd[1], wd[2], wd[3]);
  w.qb = make_uint4(wd[4], wd[5], 0u, 0u);
  w.r = make_int4(ref[0], ref[1], ref[2], ref[3]);
  wide[i] = w;
}

__global__ void __launch_bounds__(256)
k_set_parent(FatNode *fat, const int *parent_of_fat, uint32_t n_int) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n_int) fat[i].d.z = i == 0 ? -1 : parent_of_fat[i];
}

__global__ void __launch_bounds__(256)
k_depth(const int *parent_of_fat, const int *parent_of_leaf, uint32_t n_leaf,
        uint32_t n_int, uint32_t *depth) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_leaf) return;
  uint32_t d = 1;  // the leaf
  int p = parent_of_leaf[i];
  while (p >= 0 && d <= n_int + 1) {
    ++d;
    p = p == 0 ? -1 : parent_of_fat[p];
  }
  atomicMax(depth, d);
}

// ---------------------------------------------------------------- host side
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct DevBuf {
  ajg_ctx *ctx;
  std::vector<void *> ptrs;
  explicit DevBuf(ajg_ctx *c) : ctx(c) {}
  ~DevBuf() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T **p, size_t count, const char *what) {
    void *q = nullptr;
    int st = dev_alloc(ctx, &q, count * sizeof(T) + 16, what);
    if (st != AJG_OK) return st;
    ptrs.push_back(q);
    *p = (T *)q;
    return AJG_OK;
  }
};

void destroy_accel(ajg_accel *a) {
  if (!a) return;
  (void)hipFree(a->final_nodes);
  (void)hipFree(a->perm);
  (void)hipFree(a->leaf_tris);
  (void)hipFree(a->fat);
  (void)hipFree(a->wide);
  (void)hipFree(a->cursor);
  delete a;
}

int build_accel(ajg_ctx *ctx, const float *tri, uint32_t n, const float *scene_box,
                uint32_t r, hipStream_t s, ajg_accel **out) {
  *out = nullptr;
  if (n == 0) return fail(ctx, AJG_ERR_INVALID, "accel build: no triangles");
  if (n > 0x7FFFFFF0u / 2) return fail(ctx, AJG_ERR_INVALID, "accel build: too many triangles");
  if (r == 0 || r > (uint32_t)kMaxRadius)
    return fail(ctx, AJG_ERR_INVALID, "accel build: search_radius must be in [1, 64]");

  ajg_accel *a = new ajg_accel();
  a->ctx = ctx;
  a->n_tri = n;
  const uint32_t nn = 2 * n - 1;
  int st;
  auto bail = [&](int code) {
    destroy_accel(a);
    return code;
  };
  if ((st = dev_alloc(ctx, (void **)&a->final_nodes, sizeof(PlocNode) * nn, "final nodes"))) return bail(st);
  if ((st = dev_alloc(ctx, (void **)&a->perm, sizeof(uint32_t) * n, "perm"))) return bail(st);
  if ((st = dev_alloc(ctx, (void **)&a->leaf_tris, sizeof(float4) * 3 * (size_t)n, "leaf tris"))) return bail(st);
  if ((st = dev_alloc(ctx, (void **)&a->cursor, 16, "trace cursor"))) return bail(st);
  if (n > 1 &&
      ((st = dev_alloc(ctx, (void **)&a->fat, sizeof(FatNode) * (size_t)(n - 1), "fat nodes")) ||
       (st = dev_alloc(ctx, (void **)&a->wide, sizeof(WideNode) * (size_t)(n - 1), "traversal nodes"))))
    return bail(st);
  a->bytes = sizeof(PlocNode) * (uint64_t)nn + 4ull * n + 48ull * n +
             (uint64_t)sizeof(WideNode) * (n - 1);

  DevBuf tmp(ctx);
  // ---- scene box
  float box[6];
  if (scene_box) {
    for (int i = 0; i < 6; ++i) box[i] = scene_box[i];
  } else {
    const uint32_t nv = 3 * n;
    const uint32_t pb = std::min<uint32_t>(cdiv(nv, 256), 1024);
    float *partial;
    if ((st = tmp.alloc(&partial, 6 * (size_t)pb, "box partial"))) return bail(st);
    hipLaunchKernelGGL(k_box_partial, dim3(pb), dim3(256), 0, s, tri, nv, partial);
    std::vector<float> hp(6 * (size_t)pb);
    hipError_t e = hipMemcpyAsync(hp.data(), partial, sizeof(float) * 6 * pb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "box partial copy"));
    box3 b = box_empty();
    for (uint32_t i = 0; i < pb; ++i) {
      b.mn = vmin(b.mn, mk3(hp[6 * i], hp[6 * i + 1], hp[6 * i + 2]));
      b.mx = vmax(b.mx, mk3(hp[6 * i + 3], hp[6 * i + 4], hp[6 * i + 5]));
    }
    st3(box, b.mn);
    st3(box + 3, b.mx);
  }
  for (int i = 0; i < 6; ++i) a->scene_box[i] = box[i];

  // ---- Morton codes: world2grid = 1024 * (1 / size); offset = -min * w2g
  const f3 mn = ld3(box), mx = ld3(box + 3);
  const f3 size = sub(mx, mn);
  const f3 inv = mk3(1.f / size.x, 1.f / size.y, 1.f / size.z);
  const f3 w2g = scale(inv, 1024.f);
  const f3 off = mul(neg(mn), w2g);
  uint32_t *keys, *keys2, *vals2;
  if ((st = tmp.alloc(&keys, n, "keys"))) return bail(st);
  if ((st = tmp.alloc(&keys2, n, "keys2"))) return bail(st);
  if ((st = tmp.alloc(&vals2, n, "vals2"))) return bail(st);
  uint32_t *vals = a->perm;
  hipLaunchKernelGGL(k_morton, dim3(cdiv(n, 256)), dim3(256), 0, s, tri, n, w2g, off, keys, vals);

  // ---- stable radix sort, 4 x 8 bits
  const uint32_t sblocks = cdiv(n, kSortTile);
  const uint32_t ncnt = 256 * sblocks;
  uint32_t *counts, *offs, *scan_tmp;
  if ((st = tmp.alloc(&counts, ncnt, "sort counts"))) return bail(st);
  if ((st = tmp.alloc(&offs, ncnt, "sort offsets"))) return bail(st);
  const uint32_t scan_n = std::max<uint32_t>(ncnt, nn);
  if ((st = tmp.alloc(&scan_tmp, scan_tmp_words(scan_n), "scan tmp"))) return bail(st);
  uint32_t *kin = keys, *vin = vals, *kout = keys2, *vout = vals2;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(sblocks), dim3(kSortThreads), 0, s, kin, n, shift, counts, sblocks);
    scan_u32(counts, offs, ncnt, false, scan_tmp, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(sblocks), dim3(kSortThreads), 0, s, kin, vin, kout, vout, n, shift,
                       offs, sblocks);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // 4 passes: the sorted values are back in `vals` == a->perm.

  // ---- leaves
  PlocNode *live_a, *live_b;
  uint32_t *nb, *m;
  if ((st = tmp.alloc(&live_a, n, "live a"))) return bail(st);
  if ((st = tmp.alloc(&live_b, n, "live b"))) return bail(st);
  if ((st = tmp.alloc(&nb, n, "neighbour"))) return bail(st);
  if ((st = tmp.alloc(&m, n, "merge flags"))) return bail(st);
  hipLaunchKernelGGL(k_leaf_init, dim3(cdiv(n, 256)), dim3(256), 0, s, tri, a->perm, n, live_a, a->leaf_tris);

  // ---- clustering
  uint32_t live_n = n, end = nn, iters = 0;
  PlocNode *lin = live_a, *lout = live_b;
  while (live_n > (uint32_t)kTail) {
    const uint32_t g = cdiv(live_n, 256);
    hipLaunchKernelGGL(k_nn, dim3(cdiv(live_n, kNNThreads)), dim3(kNNThreads), 0, s, lin, live_n, r, nb);
    hipLaunchKernelGGL(k_flags, dim3(g), dim3(256), 0, s, nb, live_n, m);
    scan_u32(m, m, live_n, true, scan_tmp, s);
    uint32_t mc = 0;
    hipError_t e = hipMemcpyAsync(&mc, m + live_n - 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "merge count copy"));
    if (mc == 0) return bail(fail(ctx, AJG_ERR_INTERNAL, "PLOC iteration merged no pair"));
    hipLaunchKernelGGL(k_emit, dim3(g), dim3(256), 0, s, lin, nb, m, live_n, mc, end, lout, a->final_nodes);
    end -= 2 * mc;
    live_n -= mc;
    std::swap(lin, lout);
    ++iters;
  }
  uint32_t *dstat;
  if ((st = tmp.alloc(&dstat, 4, "tail status"))) return bail(st);
  AJG_HIP(ctx, hipMemsetAsync(dstat, 0, 16, s));
  hipLaunchKernelGGL(k_tail, dim3(1), dim3(kTail), 0, s, lin, live_n, end, r, a->final_nodes, dstat, dstat + 1);
  uint32_t hstat[2] = {0, 0};
  {
    hipError_t e = hipMemcpyAsync(hstat, dstat, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "PLOC tail"));
  }
  if (hstat[1] != 0) return bail(fail(ctx, AJG_ERR_INTERNAL, "PLOC tail failed to merge"));
  a->iterations = iters + hstat[0];

  // ---- fat nodes
  if (n == 1) {
    a->root_ref = ~0;
  } else {
    uint32_t *flag, *fat_idx;
    int *pof, *pol;
    if ((st = tmp.alloc(&flag, nn, "interior flags"))) return bail(st);
    if ((st = tmp.alloc(&fat_idx, nn, "fat index"))) return bail(st);
    if ((st = tmp.alloc(&pof, n, "parent of fat"))) return bail(st);
    if ((st = tmp.alloc(&pol, n, "parent of leaf"))) return bail(st);
    const uint32_t gn = cdiv(nn, 256);
    hipLaunchKernelGGL(k_mark_interior, dim3(gn), dim3(256), 0, s, a->final_nodes, nn, flag);
    scan_u32(flag, fat_idx, nn, false, scan_tmp, s);
    hipLaunchKernelGGL(k_build_fat, dim3(gn), dim3(256), 0, s, a->final_nodes, fat_idx, nn, a->fat, pof, pol);
    hipLaunchKernelGGL(k_set_parent, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, a->fat, pof, n - 1);
    uint32_t *ddepth;  // [0] tree depth, [1] boxes the quantised grid failed to contain
    if ((st = tmp.alloc(&ddepth, 2, "depth"))) return bail(st);
    AJG_HIP(ctx, hipMemsetAsync(ddepth, 0, 8, s));
    hipLaunchKernelGGL(k_build_wide, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, (const FatNode *)a->fat,
                       n - 1, a->wide, ddepth + 1);
    hipLaunchKernelGGL(k_depth, dim3(cdiv(n, 256)), dim3(256), 0, s, pof, pol, n, n - 1, ddepth);
    uint32_t hd[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(hd, ddepth, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "fat layout"));
    a->depth = hd[0];
    if (hd[1] != 0)
      return bail(fail(ctx, AJG_ERR_INVALID, "accel build: non-finite triangle coordinates"));
    // traversal reads the wide nodes only
    (void)hipFree(a->fat);
    a->fat = nullptr;
    a->root_ref = 0;
  }
  if (n == 1) a->depth = 1;
  PlocNode root;
  {
    hipError_t e = hipMemcpyAsync(&root, a->final_nodes, sizeof(PlocNode), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "root copy"));
  }
  for (int c = 0; c < 3; ++c) {
    a->root_box[c] = root.mn[c];
    a->root_box[3 + c] = root.mx[c];
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bail(hip_fail(ctx, e, "build kernels"));
  *out = a;
  return AJG_OK;
}

}  // namespace ajg
