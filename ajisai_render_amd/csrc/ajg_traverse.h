// ajg_traverse.h — device ray/BVH traversal shared by the batch trace
// kernels and the wavefront renderer.
//
// Closest hit restates PLOCAggregate::Intersect/IntersectAux
// (ploc_aggregate.cpp:329-335, 367-391): a recursive left-first DFS whose
// t_max shrinks to each accepted hit. The device loop keeps the same visit
// order with an explicit stack. The slab test of a node (IntersectP,
// :154-162) splits into a part that does not depend on t_max
// (entry <= exit and exit > t_min) and the comparison entry < t_max; a node
// evaluates both child boxes when it is fetched, pushes the right child
// with its entry distance, and compares that entry with the t_max current at
// pop time — exactly when the recursive reference would test it. A fetched
// WideNode also carries its left child's child boxes, so entering an interior
// left child performs that child's tests without another fetch (same order,
// same t_max). The node's boxes are quantised outward (ajg_common.h), so an
// interior test passes wherever the reference's passes and the traversal
// visits a superset of its nodes; a leaf is tested against its exact box
// (the triangle's AABB, as the PLOC leaf node holds it) at visit time, with
// the t_max current then, as the reference tests it. A leaf the reference
// never reaches lies inside an ancestor whose exact test failed, so its own
// exact test (an interval inside that ancestor's) fails as well: ids, t, u,
// v are bit-identical to the reference, ties included.
// Any hit restates OccludeAux (:345-365); its result does not depend on the
// visit order because t_max never changes.
#pragma once

#include "ajg_common.h"

namespace ajg {

struct DevBVH {
  const WideNode *wide;
  const float4 *leaf_tris;
  int root_ref;
  float root_mn[3], root_mx[3];
};

struct TRay {
  f3 o, d, inv;
  bool neg[3];
  bool fin;  // every component of inv is finite (no zero direction component)
  float tmin, tmax;
};

__device__ __forceinline__ TRay make_tray(f3 o, f3 d, float tmin, float tmax) {
  TRay r;
  r.o = o;
  r.d = d;
  // inv_dir = 1 / d; dir_is_neg = inv_dir < 0 (ploc_aggregate.cpp:331-332)
  r.inv = mk3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  r.neg[0] = r.inv.x < 0.f;
  r.neg[1] = r.inv.y < 0.f;
  r.neg[2] = r.inv.z < 0.f;
  r.fin = __builtin_isfinite(r.inv.x) && __builtin_isfinite(r.inv.y) && __builtin_isfinite(r.inv.z);
  r.tmin = tmin;
  r.tmax = tmax;
  return r;
}

// Static part of IntersectP. Returns false if the box is rejected
// independently of t_max; otherwise *entry = tmin3.max().
__device__ __forceinline__ bool slab_static(const TRay &r, float mnx, float mny,
                                            float mnz, float mxx, float mxy,
                                            float mxz, float *entry) {
  const float lx = ((r.neg[0] ? mxx : mnx) - r.o.x) * r.inv.x;
  const float ly = ((r.neg[1] ? mxy : mny) - r.o.y) * r.inv.y;
  const float lz = ((r.neg[2] ? mxz : mnz) - r.o.z) * r.inv.z;
  const float hx = ((r.neg[0] ? mnx : mxx) - r.o.x) * r.inv.x;
  const float hy = ((r.neg[1] ? mny : mxy) - r.o.y) * r.inv.y;
  const float hz = ((r.neg[2] ? mnz : mxz) - r.o.z) * r.inv.z;
  const float tmin_max = fmax_r(fmax_r(lx, ly), lz);
  const float tmax_min = fmin_r(fmin_r(hx, hy), hz);
  *entry = tmin_max;
  if (tmin_max > tmax_min) return false;
  return tmax_min > r.tmin;
}

// Triangle::Intersect / Occlude (ploc_aggregate.cpp:184-242) without the
// t < best part: returns true with t, u, v when tmin < t < tmax (t <= tmax
// with `incl`, which only two-level traversal asks for, see below).
__device__ __forceinline__ bool mt_test(f3 v0, f3 v1, f3 v2, const TRay &r,
                                        float tmax, float *t_out, float *u_out,
                                        float *v_out, bool incl = false) {
  const f3 e0 = sub(v1, v0);
  const f3 e1 = sub(v2, v0);
  const f3 h = cross(r.d, e1);
  const float a = dot(e0, h);
  if (a > -1e-6f && a < 1e-6f) return false;
  const float f = 1.f / a;
  const f3 s = sub(r.o, v0);
  const float u = f * dot(s, h);
  if (u < 0.f || u > 1.f) return false;
  const f3 q = cross(s, e0);
  const float v = f * dot(r.d, q);
  if (v < 0.f || u + v > 1.f) return false;
  const float t = f * dot(e1, q);
  if (t > r.tmin && (t < tmax || (incl && t == tmax))) {
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
  }
  return false;
}

struct THit {
  float t, u, v;
  int slot;  // leaf slot, -1 on a miss
};

__device__ __forceinline__ void load_leaf(const float4 *lt, int slot, f3 *a, f3 *b,
                                          f3 *c) {
  const float4 p0 = lt[3 * slot + 0];
  const float4 p1 = lt[3 * slot + 1];
  const float4 p2 = lt[3 * slot + 2];
  *a = mk3(p0.x, p0.y, p0.z);
  *b = mk3(p1.x, p1.y, p1.z);
  *c = mk3(p2.x, p2.y, p2.z);
}

// The exact leaf box test (IntersectP on the triangle's AABB,
// ploc_aggregate.cpp:154-162, 168-182) with the given t_max.
// Where every component of 1/d is finite no NaN can arise (the builder
// rejects non-finite vertices): the reference's compare-select Min/Max
// then equal fminf/fmaxf up to the sign of a zero, which no comparison
// below can see, and the plane its sign of 1/d selects is the nearer
// (farther) of the two, since (lo - o) * inv and (hi - o) * inv keep the
// order of lo <= hi under rounding; so min/max give the reference's
// distances bit for bit with fewer instructions. A ray with a zero
// direction component takes the reference's own sequence.
__device__ __forceinline__ bool leaf_box_test(const TRay &r, f3 a, f3 b, f3 c, float tmax,
                                              bool incl = false) {
  if (r.fin) {
    const float t0x = (fminf(fminf(a.x, b.x), c.x) - r.o.x) * r.inv.x;
    const float t0y = (fminf(fminf(a.y, b.y), c.y) - r.o.y) * r.inv.y;
    const float t0z = (fminf(fminf(a.z, b.z), c.z) - r.o.z) * r.inv.z;
    const float t1x = (fmaxf(fmaxf(a.x, b.x), c.x) - r.o.x) * r.inv.x;
    const float t1y = (fmaxf(fmaxf(a.y, b.y), c.y) - r.o.y) * r.inv.y;
    const float t1z = (fmaxf(fmaxf(a.z, b.z), c.z) - r.o.z) * r.inv.z;
    const float e = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    const float x = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    return !(e > x) && x > r.tmin && (e < tmax || (incl && e == tmax));
  }
  const box3 bx = tri_box(a, b, c);
  float e;
  return slab_static(r, bx.mn.x, bx.mn.y, bx.mn.z, bx.mx.x, bx.mx.y, bx.mx.z, &e) &&
         (e < tmax || (incl && e == tmax));
}

// The leaf box test at visit time + MT, updating the closest hit.
__device__ __forceinline__ void visit_leaf_closest(const DevBVH &bvh, const TRay &r,
                                                   int slot, THit *h, float *tmax) {
  f3 a, b, c;
  load_leaf(bvh.leaf_tris, slot, &a, &b, &c);
  float t, u, v;
  if (leaf_box_test(r, a, b, c, *tmax) && mt_test(a, b, c, r, *tmax, &t, &u, &v)) {
    *tmax = t;
    h->t = t;
    h->u = u;
    h->v = v;
    h->slot = slot;
  }
}

// Traversal stack: entry k (k = 0 .. d-1, (node ref, entry t)) lives in the
// lane's LDS ring slot k & (kLdsStack - 1) while k >= d - kLdsStack; older
// entries are moved to a private spill array (index k), which only deep
// paths touch. The depth d is a plain register; the ring is one LDS column
// per lane (lanes on consecutive banks).
static constexpr int kLdsStack = 8;
static constexpr int kTravThreads = 256;

struct StackRef {
  int *ln;    // LDS node refs: slot k at ln[k * kTravThreads]
  float *le;  // LDS entry distances
  int *sn;    // spill (private)
  float *se;
};

// LDS: ring entries per lane (a power of two; kLdsStack unless tuned).
template <int LDS = kLdsStack>
__device__ __forceinline__ void stack_push(const StackRef &s, int &d, int n, float e) {
  const int slot = (d & (LDS - 1)) * kTravThreads;
  if (d >= LDS) {  // the slot holds entry d - LDS: move it to the spill
    s.sn[d - LDS] = s.ln[slot];
    s.se[d - LDS] = s.le[slot];
  }
  s.ln[slot] = n;
  s.le[slot] = e;
  ++d;
}

template <int LDS = kLdsStack>
__device__ __forceinline__ void stack_pop(const StackRef &s, int &d, int *n, float *e) {
  --d;
  const int slot = (d & (LDS - 1)) * kTravThreads;
  *n = s.ln[slot];
  *e = s.le[slot];
  if (d >= LDS) {  // bring entry d - LDS back into the freed slot
    s.ln[slot] = s.sn[d - LDS];
    s.le[slot] = s.se[d - LDS];
  }
}

// Static slab part of quantised slot k (0..3) of a node, from the node's
// plane words chosen for the ray: near[a] holds, per slot, the plane the
// ray enters axis a's slab through (the upper plane where the direction is
// negative), far[a] the plane it leaves through; byte k is slot k's. The
// planes and (plane - origin) * inv_dir are computed exactly as the
// reference computes its slab test, and a quantised box contains the exact
// one, so per axis the near distance is at most the exact box's and the far
// distance at least its (rounding is monotone). The three axes combine with
// fmaxf / fminf, which ignore a NaN operand (a zero direction component
// with the ray origin on a plane), where the reference's compare-select
// chain would carry a NaN from the first axis and cull the box: the test
// passes wherever the reference's test of the exact box passes, with an
// entry distance no larger. It only ever visits more nodes; the exact leaf
// box test at visit time decides what is hit.
struct QBoxes {
  float ox, oy, oz;
  float sx, sy, sz;
  uint32_t near[3], far[3];
};
__device__ __forceinline__ bool qslab(const TRay &r, const QBoxes &q, int k, float *entry) {
  const float nx = (qplane(q.near[0], k, q.sx, q.ox) - r.o.x) * r.inv.x;
  const float ny = (qplane(q.near[1], k, q.sy, q.oy) - r.o.y) * r.inv.y;
  const float nz = (qplane(q.near[2], k, q.sz, q.oz) - r.o.z) * r.inv.z;
  const float fx = (qplane(q.far[0], k, q.sx, q.ox) - r.o.x) * r.inv.x;
  const float fy = (qplane(q.far[1], k, q.sy, q.oy) - r.o.y) * r.inv.y;
  const float fz = (qplane(q.far[2], k, q.sz, q.oz) - r.o.z) * r.inv.z;
  const float tmin_max = fmaxf(fmaxf(nx, ny), nz);
  const float tmax_min = fminf(fminf(fx, fy), fz);
  *entry = tmin_max;
  return !(tmin_max > tmax_min) && tmax_min > r.tmin;
}
__device__ __forceinline__ QBoxes qload(const WideNode *w, const TRay &r) {
  QBoxes q;
  const float4 o = w->o;
  q.ox = o.x;
  q.oy = o.y;
  q.oz = o.z;
  q.sx = qstep(o, 0);
  q.sy = qstep(o, 1);
  q.sz = qstep(o, 2);
  const uint4 a = w->qa;
  const uint4 b = w->qb;
  q.near[0] = r.neg[0] ? a.w : a.x;
  q.far[0] = r.neg[0] ? a.x : a.w;
  q.near[1] = r.neg[1] ? b.x : a.y;
  q.far[1] = r.neg[1] ? a.y : b.x;
  q.near[2] = r.neg[2] ? b.y : a.z;
  q.far[2] = r.neg[2] ? a.z : b.y;
  return q;
}

// Closest-hit visit of interior node `node` from one fetch: slots 3, 2, 1
// (when present and their static slab part passes) are pushed with their
// entry distances, to be compared with t_max when popped — when the
// recursive reference would test them; slot 0 is entered if its entry is
// below t_max. Returns false when the stack must be popped; otherwise `node`
// is the next node (interior or ~leaf).
template <int LDS>
__device__ __forceinline__ bool closest_step(const WideNode *wide, int &node, const TRay &r,
                                             float tmax, const StackRef &st, int &d,
                                             bool incl = false) {
  const WideNode *w = wide + node;
  const QBoxes q = qload(w, r);
  const int4 ref = w->r;
  float e;
  if (ref.w != kNoRef && qslab(r, q, 3, &e)) stack_push<LDS>(st, d, ref.w, e);
  if (ref.z != kNoRef && qslab(r, q, 2, &e)) stack_push<LDS>(st, d, ref.z, e);
  if (qslab(r, q, 1, &e)) stack_push<LDS>(st, d, ref.y, e);
  if (qslab(r, q, 0, &e) && (e < tmax || (incl && e == tmax))) {
    node = ref.x;
    return true;
  }
  return false;
}

// Top-level visit of the two-level closest-hit traversal: the slots that
// pass their static slab part with entry <= t_max are visited nearest entry
// first (equal entries in slot order); the others are pushed farthest first.
// Entries equal to t_max still count: a tie with the current hit may go to a
// lower instance number (the two-level tie rule, below).
__device__ __forceinline__ void near_cs(float &ka, int &na, int &sa, float &kb, int &nb,
                                        int &sb) {
  if (kb < ka || (kb == ka && sb < sa)) {
    const float k = ka;
    const int n = na, s = sa;
    ka = kb, na = nb, sa = sb;
    kb = k, nb = n, sb = s;
  }
}
template <int LDS>
__device__ __forceinline__ bool closest_step_near(const WideNode *wide, int &node,
                                                  const TRay &r, float tmax,
                                                  const StackRef &st, int &d) {
  const WideNode *w = wide + node;
  const QBoxes q = qload(w, r);
  const int4 ref = w->r;
  const float inf = __int_as_float(0x7f800000);
  float k0, k1, k2, k3;
  int n0 = ref.x, n1 = ref.y, n2 = ref.z, n3 = ref.w;
  int s0 = 0, s1 = 1, s2 = 2, s3 = 3;
  if (!(qslab(r, q, 0, &k0) && k0 <= tmax)) k0 = inf, n0 = kNoRef;
  if (!(qslab(r, q, 1, &k1) && k1 <= tmax)) k1 = inf, n1 = kNoRef;
  if (!(n2 != kNoRef && qslab(r, q, 2, &k2) && k2 <= tmax)) k2 = inf, n2 = kNoRef;
  if (!(n3 != kNoRef && qslab(r, q, 3, &k3) && k3 <= tmax)) k3 = inf, n3 = kNoRef;
  // a sorting network on (entry, slot); rejected slots carry no reference
  near_cs(k0, n0, s0, k1, n1, s1);
  near_cs(k2, n2, s2, k3, n3, s3);
  near_cs(k0, n0, s0, k2, n2, s2);
  near_cs(k1, n1, s1, k3, n3, s3);
  near_cs(k1, n1, s1, k2, n2, s2);
  if (n3 != kNoRef) stack_push<LDS>(st, d, n3, k3);
  if (n2 != kNoRef) stack_push<LDS>(st, d, n2, k2);
  if (n1 != kNoRef) stack_push<LDS>(st, d, n1, k1);
  if (n0 != kNoRef) {
    node = n0;
    return true;
  }
  return false;  // (whatever was pushed is popped next)
}

// Any-hit visit (t_max fixed, so culling happens at push time and the visit
// order does not change the result).
template <int LDS>
__device__ __forceinline__ bool any_step(const WideNode *wide, int &node, const TRay &r,
                                         float tmax, const StackRef &st, int &d) {
  const WideNode *w = wide + node;
  const QBoxes q = qload(w, r);
  const int4 ref = w->r;
  float e;
  if (ref.w != kNoRef && qslab(r, q, 3, &e) && e < tmax) stack_push<LDS>(st, d, ref.w, e);
  if (ref.z != kNoRef && qslab(r, q, 2, &e) && e < tmax) stack_push<LDS>(st, d, ref.z, e);
  if (qslab(r, q, 1, &e) && e < tmax) stack_push<LDS>(st, d, ref.y, e);
  if (qslab(r, q, 0, &e) && e < tmax) {
    node = ref.x;
    return true;
  }
  return false;
}

// Stack bound: spill_for_depth.
template <int SPILL, bool COUNT>
__device__ __forceinline__ THit trace_closest(const DevBVH &bvh, const TRay &r, int *lds_node,
                                              float *lds_entry, uint32_t *n_nodes,
                                              uint32_t *n_tris) {
  THit h;
  h.t = __int_as_float(0x7f800000);
  h.u = 0.f;
  h.v = 0.f;
  h.slot = -1;
  float tmax = r.tmax;
  float e;
  if (!slab_static(r, bvh.root_mn[0], bvh.root_mn[1], bvh.root_mn[2], bvh.root_mx[0],
                   bvh.root_mx[1], bvh.root_mx[2], &e) ||
      !(e < tmax))
    return h;
  int node = bvh.root_ref;
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0;
  for (;;) {
    if (node < 0) {
      if (COUNT) ++*n_tris;
      visit_leaf_closest(bvh, r, ~node, &h, &tmax);
    } else {
      if (COUNT) ++*n_nodes;
      // the right child is visited after the left subtree: its entry
      // distance is kept and compared with t_max when it is popped
      if (closest_step<kLdsStack>(bvh.wide, node, r, tmax, st, d)) continue;
    }
    bool found = false;
    while (d > 0) {
      int n;
      float en;
      stack_pop(st, d, &n, &en);
      if (en < tmax) {
        node = n;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  return h;
}

template <int SPILL, bool COUNT = false, int LDS = kLdsStack>
__device__ __forceinline__ bool trace_any(const DevBVH &bvh, const TRay &r, int *lds_node,
                                          float *lds_entry, uint32_t *n_nodes = nullptr,
                                          uint32_t *n_tris = nullptr) {
  float e;
  if (!slab_static(r, bvh.root_mn[0], bvh.root_mn[1], bvh.root_mn[2], bvh.root_mx[0],
                   bvh.root_mx[1], bvh.root_mx[2], &e) ||
      !(e < r.tmax))
    return false;
  int node = bvh.root_ref;
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0;
  for (;;) {
    if (node < 0) {
      if (COUNT) ++*n_tris;
      f3 a, b, c;
      load_leaf(bvh.leaf_tris, ~node, &a, &b, &c);
      float t, u, v;
      if (leaf_box_test(r, a, b, c, r.tmax) && mt_test(a, b, c, r, r.tmax, &t, &u, &v)) return true;
    } else {
      if (COUNT) ++*n_nodes;
      if (any_step<LDS>(bvh.wide, node, r, r.tmax, st, d)) continue;
    }
    if (d == 0) break;
    int n;
    float en;
    stack_pop<LDS>(st, d, &n, &en);
    node = n;
  }
  return false;
}

// ---- Persistent closest-hit traversal with lane refill.
// A wave keeps one ray per lane and advances every active lane by one node
// (or one leaf) per step; when at least kRefillMin lanes have finished, the
// wave takes that many new rays from the batch with one atomic on a shared
// cursor, so lanes stay busy whatever the spread of path lengths within a
// wave. Each ray's traversal (order, culling, arithmetic) is the same as
// trace_closest, so results are unchanged. The queue kernels' any-hit
// traversals (persistent_any, persistent_any_2l below) refill the same way.
static constexpr int kRefillMin = 16;

// Wave-uniform refill: idle lanes take consecutive batch indices. Returns the
// lane's new index (kNone if none) and sets *drained when the batch is used up.
// With STRIPES > 1 the cursor is striped: wave w takes from cursor word
// w % STRIPES (kCursorStride words apart), which hands out the 64-ray chunks
// c = s (mod STRIPES) in order (cursor value j: chunk s + STRIPES (j / 64),
// ray j % 64). Atomics on one word are served one at a time, and where rays
// are short every wave comes back for more within a few steps.
static constexpr uint32_t kCursorStride = 64;
template <uint32_t STRIPES = 1>
__device__ __forceinline__ uint32_t refill(bool idle, uint32_t n, uint32_t *cursor, bool *drained) {
  const uint64_t mask = __ballot(idle);
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const uint32_t cnt = (uint32_t)__popcll(mask);
  const uint32_t s = STRIPES == 1 ? 0u : ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % STRIPES;
  auto at = [&](uint32_t j) { return STRIPES == 1 ? j : (s + STRIPES * (j >> 6)) * 64 + (j & 63); };
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(cursor + s * kCursorStride, cnt);
  base = __shfl(base, leader, 64);
  if (at(base + cnt) >= n) *drained = true;
  if (!idle) return kNone;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t idx = at(base + (uint32_t)__popcll(mask & lt));
  return idx < n ? idx : kNone;
}

// Drive closest-hit traversal over rays [0, n) of a batch. load(i, bvh)
// returns the TRay of ray i and, when PER_RAY_BVH, sets the BVH it is traced
// in (batched renders); store(i, hit) writes its result. The per-lane state
// is kept in plain locals (registers); only the stack spill array is
// private memory.
template <int SPILL, bool COUNT, bool PER_RAY_BVH = false, int LDS = kLdsStack,
          uint32_t STRIPES = 1, class Load, class Store>
__device__ __forceinline__ void persistent_closest(const DevBVH &bvh_uniform, uint32_t n,
                                                   uint32_t *cursor, int *lds_node,
                                                   float *lds_entry, Load load, Store store,
                                                   uint32_t *nn, uint32_t *nt) {
  DevBVH bvh_lane = bvh_uniform;
  const DevBVH &bvh = PER_RAY_BVH ? bvh_lane : bvh_uniform;
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0;
  TRay r;
  THit h;
  float tmax = 0.f;
  int node = 0;
  uint32_t ray = kNone;
  bool drained = false;
  for (;;) {
    const bool idle = ray == kNone;
    if (!drained && __popcll(__ballot(idle)) >= kRefillMin) {
      const uint32_t idx = refill<STRIPES>(idle, n, cursor, &drained);
      if (idx != kNone) {
        r = load(idx, bvh_lane);
        h.t = __int_as_float(0x7f800000);
        h.u = 0.f;
        h.v = 0.f;
        h.slot = -1;
        tmax = r.tmax;
        d = 0;
        float e;
        if (slab_static(r, bvh.root_mn[0], bvh.root_mn[1], bvh.root_mn[2], bvh.root_mx[0],
                        bvh.root_mx[1], bvh.root_mx[2], &e) &&
            e < tmax) {
          node = bvh.root_ref;
          ray = idx;
        } else {
          store(idx, h);
        }
      }
      // (no continue: the lanes that kept their rays take this step too)
    }
    if (__ballot(ray != kNone) == 0ull) {
      if (drained) break;
      continue;
    }
    if (ray == kNone) continue;
    // a leaf test (if the lane stands on a leaf) and then one interior
    // visit, in trace_closest's order
    bool next = true;
    if (node < 0) {
      if (COUNT) ++*nt;
      visit_leaf_closest(bvh, r, ~node, &h, &tmax);
      next = false;
      while (d > 0) {
        int pn;
        float en;
        stack_pop<LDS>(st, d, &pn, &en);
        if (en < tmax) {
          node = pn;
          next = true;
          break;
        }
      }
    }
    if (next && node >= 0) {
      if (COUNT) ++*nn;
      next = closest_step<LDS>(bvh.wide, node, r, tmax, st, d);
      if (!next) {
        while (d > 0) {
          int pn;
          float en;
          stack_pop<LDS>(st, d, &pn, &en);
          if (en < tmax) {
            node = pn;
            next = true;
            break;
          }
        }
      }
    }
    if (!next) {
      store(ray, h);
      ray = kNone;
    }
  }
}

// Any-hit traversal over rays [0, n) with the same lane refill. Its result
// does not depend on the order in which a ray's triangles are tested (t_max
// is fixed), so a lane that reaches a leaf parks it and goes on with the
// next stacked node; the parked leaves of the wave are tested together once
// at least half of the traversing lanes hold one (or none can go on), which
// keeps more lanes busy in both the node and the triangle code than a leaf
// test whenever one lane stands on a leaf. load(i, bvh) as for
// persistent_closest; store(i, occluded).
template <int SPILL, bool COUNT, bool PER_RAY_BVH = false, int LDS = kLdsStack,
          uint32_t STRIPES = 1, class Load, class Store>
__device__ __forceinline__ void persistent_any(const DevBVH &bvh_uniform, uint32_t n,
                                               uint32_t *cursor, int *lds_node, float *lds_entry,
                                               Load load, Store store, uint32_t *nn,
                                               uint32_t *nt) {
  DevBVH bvh_lane = bvh_uniform;
  const DevBVH &bvh = PER_RAY_BVH ? bvh_lane : bvh_uniform;
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0;
  TRay r;
  int node = kNoRef;    // next node: >= 0 interior, < 0 ~leaf slot, kNoRef none left
  int parked = kNoRef;  // a leaf waiting for the wave's next leaf tests
  uint32_t ray = kNone;
  bool drained = false;
  for (;;) {
    const bool idle = ray == kNone;
    if (!drained && __popcll(__ballot(idle)) >= kRefillMin) {
      const uint32_t idx = refill<STRIPES>(idle, n, cursor, &drained);
      if (idx != kNone) {
        r = load(idx, bvh_lane);
        d = 0;
        parked = kNoRef;
        float e;
        if (slab_static(r, bvh.root_mn[0], bvh.root_mn[1], bvh.root_mn[2], bvh.root_mx[0],
                        bvh.root_mx[1], bvh.root_mx[2], &e) &&
            e < r.tmax) {
          node = bvh.root_ref;
          ray = idx;
        } else {
          store(idx, false);
        }
      }
      // (no continue: the lanes that kept their rays take this step too)
    }
    const uint64_t act = __ballot(ray != kNone);
    if (act == 0ull) {
      if (drained) break;
      continue;
    }
    if (ray == kNone) continue;
    if (node >= 0) {
      if (COUNT) ++*nn;
      if (!any_step<LDS>(bvh.wide, node, r, r.tmax, st, d)) {
        node = kNoRef;
        if (d > 0) {
          float en;
          stack_pop<LDS>(st, d, &node, &en);
        }
      }
    }
    if (node < 0 && node != kNoRef && parked == kNoRef) {
      parked = node;
      node = kNoRef;
      if (d > 0) {
        float en;
        stack_pop<LDS>(st, d, &node, &en);
      }
    }
    const uint64_t holding = __ballot(parked != kNoRef);
    const uint64_t going = __ballot(node >= 0);
    bool hit = false;
    if (holding != 0ull && (going == 0ull || 2 * __popcll(holding) >= __popcll(act))) {
      if (parked != kNoRef) {
        if (COUNT) ++*nt;
        f3 a, b, c;
        load_leaf(bvh.leaf_tris, ~parked, &a, &b, &c);
        float t, u, v;
        hit = leaf_box_test(r, a, b, c, r.tmax) && mt_test(a, b, c, r, r.tmax, &t, &u, &v);
        parked = kNoRef;
      }
    }
    if (hit || (node == kNoRef && parked == kNoRef)) {
      store(ray, hit);
      ray = kNone;
    }
  }
}

// ---- Two-level traversal (instancing; no counterpart in the reference).
// The top-level tree's leaves are instances; an instance holds the
// world->object affine map of its mesh and the index of the mesh's own
// (bottom-level) tree. Entering an instance transforms the ray into object
// space with the direction left unnormalised, so the ray parameter t (and
// with it t_max and every stacked entry distance) means the same in both
// spaces and one stack serves both levels: `base` is the stack depth at which
// the bottom-level traversal started; popping back to it restores the world
// ray and the top-level arrays. Visit order, culling and arithmetic per level
// are those of the one-level traversal.
struct DevInst {
  float4 w2o[3];  // rows of the world->object map (3x4, row-major)
  float4 nrm[3];  // rows of the normal map (xyz): cbrt|det M| (M^-1)^T of the linear part
  uint4 info;     // x: mesh (bottom-level tree), y: material or kKeepMaterial
};
static constexpr uint32_t kKeepMaterial = 0xFFFFFFFFu;

// Per leaf slot shading record, one 64-byte line: flags | material (and
// the normal indices it was built from), then, for mesh triangles, the
// three vertex normals (normal_and_vs) in index order, so PostIntersect
// reads the record and the normals with one dependent load instead of two.
struct TriShade {
  uint4 info;
  float4 n[3];
};

struct DevBlas {
  DevBVH bvh;
  const TriShade *tri_shade;  // per leaf slot (as the one-level scene)
  const uint32_t *perm;    // leaf slot -> triangle index within the mesh
};

// ((m0 x + m1 y) + m2 z) + m3 per row: the same order on every path.
__device__ __forceinline__ f3 xform_point(const float4 *m, f3 p) {
  return mk3(((m[0].x * p.x + m[0].y * p.y) + m[0].z * p.z) + m[0].w,
             ((m[1].x * p.x + m[1].y * p.y) + m[1].z * p.z) + m[1].w,
             ((m[2].x * p.x + m[2].y * p.y) + m[2].z * p.z) + m[2].w);
}
__device__ __forceinline__ f3 xform_vec(const float4 *m, f3 v) {
  return mk3((m[0].x * v.x + m[0].y * v.y) + m[0].z * v.z,
             (m[1].x * v.x + m[1].y * v.y) + m[1].z * v.z,
             (m[2].x * v.x + m[2].y * v.y) + m[2].z * v.z);
}

__device__ __forceinline__ uint32_t tlas_leaf_instance(const DevBVH &tlas, int slot) {
  return __float_as_uint(tlas.leaf_tris[3 * slot].w);
}

// Closest hit over a two-level scene; store(i, hit, instance). The result
// is the hit with the least (t, instance number), and within an instance
// the one-level traversal's (the first of equal t in its visit order): an
// instance numbered below the current hit's also accepts t equal to t_max
// (boxes and triangles), so the top level may be visited in any order. It
// is visited nearest entry first. The world ray is not kept while a mesh's
// tree is traversed: leaving the tree loads it again (load(i), the same
// operations as at the start), which keeps six registers per lane out of
// the live state.
// COUNT: *nn += interior node fetches (both levels) and instance visits
// (an instance's map and mesh record), *nt += leaf triangles tested and
// instance boxes.
template <int SPILL, int LDS, uint32_t STRIPES = 1, bool COUNT = false, class Load, class Store>
__device__ __forceinline__ void persistent_closest_2l(const DevBVH &tlas, const DevInst *insts,
                                                      const DevBlas *blases, uint32_t n,
                                                      uint32_t *cursor, int *lds_node,
                                                      float *lds_entry, Load load, Store store,
                                                      uint32_t *nn = nullptr,
                                                      uint32_t *nt = nullptr) {
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0;
  TRay r;  // ray of the current level
  THit h;
  uint32_t h_inst = kNone, cur_inst = 0;
  float tmax = 0.f;
  int node = 0, base = -1;
  const WideNode *wn = tlas.wide;
  const float4 *leaf = tlas.leaf_tris;
  uint32_t ray = kNone;
  bool drained = false;
  for (;;) {
    const bool idle = ray == kNone;
    if (!drained && __popcll(__ballot(idle)) >= kRefillMin) {
      const uint32_t idx = refill<STRIPES>(idle, n, cursor, &drained);
      if (idx != kNone) {
        r = load(idx);
        h.t = __int_as_float(0x7f800000);
        h.u = 0.f;
        h.v = 0.f;
        h.slot = -1;
        h_inst = kNone;
        tmax = r.tmax;
        d = 0;
        base = -1;
        wn = tlas.wide;
        leaf = tlas.leaf_tris;
        float e;
        if (slab_static(r, tlas.root_mn[0], tlas.root_mn[1], tlas.root_mn[2], tlas.root_mx[0],
                        tlas.root_mx[1], tlas.root_mx[2], &e) &&
            e < tmax) {
          node = tlas.root_ref;
          ray = idx;
        } else {
          store(idx, h, h_inst);
        }
      }
      // (no continue: the lanes that kept their rays take this step too)
    }
    if (__ballot(ray != kNone) == 0ull) {
      if (drained) break;
      continue;
    }
    if (ray == kNone) continue;
    bool next = false;
    // this instance may take a tie at t_max from the current hit
    const bool tie = base >= 0 && h_inst != kNone && cur_inst < h_inst;
    if (node < 0) {
      if (base < 0) {  // an instance: enter its mesh's tree in object space
        if (COUNT) ++*nn, ++*nt;
        f3 ia, ib, ic;  // the instance box as its degenerate leaf triangle
        load_leaf(tlas.leaf_tris, ~node, &ia, &ib, &ic);
        cur_inst = tlas_leaf_instance(tlas, ~node);
        const DevInst *I = insts + cur_inst;
        const DevBlas *B = blases + I->info.x;
        const TRay ro = make_tray(xform_point(I->w2o, r.o), xform_vec(I->w2o, r.d), r.tmin, tmax);
        const bool tie_in = h_inst != kNone && cur_inst < h_inst;
        float e;
        if (leaf_box_test(r, ia, ib, ic, tmax, true) &&
            slab_static(ro, B->bvh.root_mn[0], B->bvh.root_mn[1], B->bvh.root_mn[2],
                        B->bvh.root_mx[0], B->bvh.root_mx[1], B->bvh.root_mx[2], &e) &&
            (e < tmax || (tie_in && e == tmax))) {
          r = ro;
          base = d;
          wn = B->bvh.wide;
          leaf = B->bvh.leaf_tris;
          node = B->bvh.root_ref;
          next = true;
        }
      } else {
        if (COUNT) ++*nt;
        f3 a, b, c;
        load_leaf(leaf, ~node, &a, &b, &c);
        float t, u, v;
        if (leaf_box_test(r, a, b, c, tmax, tie) && mt_test(a, b, c, r, tmax, &t, &u, &v, tie)) {
          tmax = t;
          h.t = t;
          h.u = u;
          h.v = v;
          h.slot = ~node;
          h_inst = cur_inst;
        }
      }
    } else if (base < 0) {
      if (COUNT) ++*nn;
      next = closest_step_near<LDS>(wn, node, r, tmax, st, d);
    } else {
      if (COUNT) ++*nn;
      next = closest_step<LDS>(wn, node, r, tmax, st, d, tie);
    }
    if (!next) {
      for (;;) {
        if (d == base) {  // the mesh's tree is done: back to the top level
          base = -1;
          r = load(ray);
          r.tmax = tmax;
          wn = tlas.wide;
          leaf = tlas.leaf_tris;
        }
        if (d == 0) break;
        int pn;
        float en;
        stack_pop<LDS>(st, d, &pn, &en);
        // (top-level entries and a tying instance's keep an entry equal to t_max)
        if (en < tmax ||
            (en == tmax && (base < 0 || (h_inst != kNone && cur_inst < h_inst)))) {
          node = pn;
          next = true;
          break;
        }
      }
    }
    if (!next) {
      store(ray, h, h_inst);
      ray = kNone;
    }
  }
}

// Any hit over rays [0, n) of a two-level scene with lane refill (as
// persistent_any): per step a lane enters an instance, tests a leaf or
// visits a node, in trace_any_2l's order; leaving a mesh's tree reloads the
// world ray (load(i)). store(i, occluded).
template <int SPILL, int LDS, uint32_t STRIPES = 1, bool COUNT = false, class Load, class Store>
__device__ __forceinline__ void persistent_any_2l(const DevBVH &tlas, const DevInst *insts,
                                                  const DevBlas *blases, uint32_t n,
                                                  uint32_t *cursor, int *lds_node,
                                                  float *lds_entry, Load load, Store store,
                                                  uint32_t *nn = nullptr, uint32_t *nt = nullptr) {
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0, node = 0, base = -1;
  TRay r;
  const WideNode *wn = tlas.wide;
  const float4 *leaf = tlas.leaf_tris;
  uint32_t ray = kNone;
  bool drained = false;
  for (;;) {
    const bool idle = ray == kNone;
    if (!drained && __popcll(__ballot(idle)) >= kRefillMin) {
      const uint32_t idx = refill<STRIPES>(idle, n, cursor, &drained);
      if (idx != kNone) {
        r = load(idx);
        d = 0;
        base = -1;
        wn = tlas.wide;
        leaf = tlas.leaf_tris;
        float e;
        if (slab_static(r, tlas.root_mn[0], tlas.root_mn[1], tlas.root_mn[2], tlas.root_mx[0],
                        tlas.root_mx[1], tlas.root_mx[2], &e) &&
            e < r.tmax) {
          node = tlas.root_ref;
          ray = idx;
        } else {
          store(idx, false);
        }
      }
      // (no continue: the lanes that kept their rays take this step too)
    }
    if (__ballot(ray != kNone) == 0ull) {
      if (drained) break;
      continue;
    }
    if (ray == kNone) continue;
    bool next = false;
    if (node < 0) {
      if (base < 0) {
        if (COUNT) ++*nn, ++*nt;
        f3 ia, ib, ic;  // the instance box as its degenerate leaf triangle
        load_leaf(tlas.leaf_tris, ~node, &ia, &ib, &ic);
        const DevInst *I = insts + tlas_leaf_instance(tlas, ~node);
        const DevBlas *B = blases + I->info.x;
        const TRay ro = make_tray(xform_point(I->w2o, r.o), xform_vec(I->w2o, r.d), r.tmin,
                                  r.tmax);
        float e;
        if (leaf_box_test(r, ia, ib, ic, r.tmax) &&
            slab_static(ro, B->bvh.root_mn[0], B->bvh.root_mn[1], B->bvh.root_mn[2],
                        B->bvh.root_mx[0], B->bvh.root_mx[1], B->bvh.root_mx[2], &e) &&
            e < r.tmax) {
          r = ro;
          base = d;
          wn = B->bvh.wide;
          leaf = B->bvh.leaf_tris;
          node = B->bvh.root_ref;
          next = true;
        }
      } else {
        if (COUNT) ++*nt;
        f3 a, b, c;
        load_leaf(leaf, ~node, &a, &b, &c);
        float t, u, v;
        if (leaf_box_test(r, a, b, c, r.tmax) && mt_test(a, b, c, r, r.tmax, &t, &u, &v)) {
          store(ray, true);
          ray = kNone;
          continue;
        }
      }
    } else {
      if (COUNT) ++*nn;
      next = any_step<LDS>(wn, node, r, r.tmax, st, d);
    }
    if (!next) {
      if (d == base) {
        base = -1;
        r = load(ray);
        wn = tlas.wide;
        leaf = tlas.leaf_tris;
      }
      if (d == 0) {
        store(ray, false);
        ray = kNone;
        continue;
      }
      float en;
      stack_pop<LDS>(st, d, &node, &en);
    }
  }
}

// Any hit over a two-level scene (one ray).
template <int SPILL, int LDS>
__device__ __forceinline__ bool trace_any_2l(const DevBVH &tlas, const DevInst *insts,
                                             const DevBlas *blases, const TRay &rw, int *lds_node,
                                             float *lds_entry) {
  float e;
  if (!slab_static(rw, tlas.root_mn[0], tlas.root_mn[1], tlas.root_mn[2], tlas.root_mx[0],
                   tlas.root_mx[1], tlas.root_mx[2], &e) ||
      !(e < rw.tmax))
    return false;
  int sn[SPILL];
  float se[SPILL];
  const StackRef st = {lds_node, lds_entry, sn, se};
  int d = 0, base = -1;
  TRay r = rw;
  const WideNode *wn = tlas.wide;
  const float4 *leaf = tlas.leaf_tris;
  int node = tlas.root_ref;
  for (;;) {
    bool next = false;
    if (node < 0) {
      if (base < 0) {
        f3 ia, ib, ic;  // the instance box as its degenerate leaf triangle
        load_leaf(tlas.leaf_tris, ~node, &ia, &ib, &ic);
        const DevInst *I = insts + tlas_leaf_instance(tlas, ~node);
        const DevBlas *B = blases + I->info.x;
        const TRay ro = make_tray(xform_point(I->w2o, rw.o), xform_vec(I->w2o, rw.d), rw.tmin,
                                  rw.tmax);
        if (leaf_box_test(rw, ia, ib, ic, rw.tmax) &&
            slab_static(ro, B->bvh.root_mn[0], B->bvh.root_mn[1], B->bvh.root_mn[2],
                        B->bvh.root_mx[0], B->bvh.root_mx[1], B->bvh.root_mx[2], &e) &&
            e < rw.tmax) {
          r = ro;
          base = d;
          wn = B->bvh.wide;
          leaf = B->bvh.leaf_tris;
          node = B->bvh.root_ref;
          next = true;
        }
      } else {
        f3 a, b, c;
        load_leaf(leaf, ~node, &a, &b, &c);
        float t, u, v;
        if (leaf_box_test(r, a, b, c, rw.tmax) && mt_test(a, b, c, r, rw.tmax, &t, &u, &v))
          return true;
      }
    } else {
      next = any_step<LDS>(wn, node, r, rw.tmax, st, d);
    }
    if (!next) {
      if (d == base) {
        base = -1;
        r = rw;
        wn = tlas.wide;
        leaf = tlas.leaf_tris;
      }
      if (d == 0) return false;
      float en;
      stack_pop<LDS>(st, d, &node, &en);
    }
  }
}

// Spill capacity for a tree of the given depth (for two-level scenes, the
// sum of both levels' depths) with an LDS ring of `lds` entries per lane
// (0: unsupported depth). A fetch pushes at most three entries and covers
// two tree levels, so a path below a node of depth D holds at most
// 3 ceil(D / 2) entries; +3 covers a second level's rounding.
static constexpr int kSpillS = 64, kSpillL = 1024;
inline int spill_for_depth(uint32_t depth, int lds = kLdsStack) {
  const uint32_t entries = 3 * ((depth + 1) / 2) + 3;
  if (entries <= (uint32_t)lds + kSpillS) return kSpillS;
  if (entries <= (uint32_t)lds + kSpillL) return kSpillL;
  return 0;
}

}  // namespace ajg
