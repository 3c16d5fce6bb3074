// ajg_traverse.h — device ray/BVH traversal shared by the batch trace
// kernels and the wavefront renderer.
//
// Closest hit restates PLOCAggregate::Intersect/IntersectAux
// (ploc_aggregate.cpp:329-335, 367-391): a recursive left-first DFS whose
// t_max shrinks to each accepted hit. The device loop keeps the same visit
// order with an explicit stack. The slab test of a node (IntersectP,
// :154-162) splits into a part that does not depend on t_max
// (entry <= exit and exit > t_min) and the comparison entry < t_max; a fat
// node evaluates both child boxes when it is fetched, pushes the right child
// with its entry distance, and compares that entry with the t_max current at
// pop time — exactly when the recursive reference would test it. Leaf boxes
// are exact (the PLOC leaf AABB), so ids, t, u, v are bit-identical to the
// reference, ties included.
// Any hit restates OccludeAux (:345-365); its result does not depend on the
// visit order because t_max never changes.
#pragma once

#include "ajg_common.h"

namespace ajg {

struct DevBVH {
  const FatNode *fat;
  const float4 *leaf_tris;
  int root_ref;
  float root_mn[3], root_mx[3];
};

struct TRay {
  f3 o, d, inv;
  bool neg[3];
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
// t < best part: returns true with t, u, v when tmin < t < tmax.
__device__ __forceinline__ bool mt_test(f3 v0, f3 v1, f3 v2, const TRay &r,
                                        float tmax, float *t_out, float *u_out,
                                        float *v_out) {
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
  if (t > r.tmin && t < tmax) {
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

// The leaf box test at visit time + MT, updating the closest hit.
__device__ __forceinline__ void visit_leaf_closest(const DevBVH &bvh, const TRay &r,
                                                   int slot, THit *h, float *tmax) {
  f3 a, b, c;
  load_leaf(bvh.leaf_tris, slot, &a, &b, &c);
  float t, u, v;
  if (mt_test(a, b, c, r, *tmax, &t, &u, &v)) {
    *tmax = t;
    h->t = t;
    h->u = u;
    h->v = v;
    h->slot = slot;
  }
}

// Traversal stack: the newest kLdsStack entries live in LDS (one column
// per lane, lanes on consecutive banks); older entries spill to a private
// array, which only deep paths ever touch. Entries are (node ref, entry t).
static constexpr int kLdsStack = 8;
static constexpr int kTravThreads = 256;

template <int SPILL>
struct TravStack {
  int *ln;    // LDS node refs of this lane: element k at ln[k * kTravThreads]
  float *le;  // LDS entry distances
  int top = 0, base = 0, spilled = 0;
  int sn[SPILL];
  float se[SPILL];
  __device__ __forceinline__ TravStack(int *lds_node, float *lds_entry)
      : ln(lds_node), le(lds_entry) {}
  __device__ __forceinline__ void push(int n, float e) {
    if (top == kLdsStack) {
      sn[spilled] = ln[base * kTravThreads];
      se[spilled] = le[base * kTravThreads];
      ++spilled;
      base = (base + 1) & (kLdsStack - 1);
      --top;
    }
    const int i = (base + top) & (kLdsStack - 1);
    ln[i * kTravThreads] = n;
    le[i * kTravThreads] = e;
    ++top;
  }
  __device__ __forceinline__ bool pop(int *n, float *e) {
    if (top == 0) {
      if (spilled == 0) return false;
      --spilled;
      *n = sn[spilled];
      *e = se[spilled];
      return true;
    }
    --top;
    const int i = (base + top) & (kLdsStack - 1);
    *n = ln[i * kTravThreads];
    *e = le[i * kTravThreads];
    return true;
  }
};

// Stack bound: entries <= tree depth - 1, so SPILL >= depth - kLdsStack.
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
  TravStack<SPILL> st(lds_node, lds_entry);
  for (;;) {
    if (node < 0) {
      if (COUNT) ++*n_tris;
      visit_leaf_closest(bvh, r, ~node, &h, &tmax);
    } else {
      if (COUNT) ++*n_nodes;
      const FatNode fn = bvh.fat[node];
      float e0, e1;
      const bool s0 = slab_static(r, fn.a.x, fn.a.y, fn.a.z, fn.a.w, fn.b.x, fn.b.y, &e0);
      const bool s1 = slab_static(r, fn.b.z, fn.b.w, fn.c.x, fn.c.y, fn.c.z, fn.c.w, &e1);
      // the right child is visited after the left subtree: keep its entry
      // distance and compare it with t_max when it is popped
      if (s1) st.push(fn.d.y, e1);
      if (s0 && e0 < tmax) {
        node = fn.d.x;
        continue;
      }
    }
    bool found = false;
    int n;
    float en;
    while (st.pop(&n, &en)) {
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

template <int SPILL, bool COUNT = false>
__device__ __forceinline__ bool trace_any(const DevBVH &bvh, const TRay &r, int *lds_node,
                                          float *lds_entry, uint32_t *n_nodes = nullptr,
                                          uint32_t *n_tris = nullptr) {
  float e;
  if (!slab_static(r, bvh.root_mn[0], bvh.root_mn[1], bvh.root_mn[2], bvh.root_mx[0],
                   bvh.root_mx[1], bvh.root_mx[2], &e) ||
      !(e < r.tmax))
    return false;
  int node = bvh.root_ref;
  TravStack<SPILL> st(lds_node, lds_entry);
  for (;;) {
    if (node < 0) {
      if (COUNT) ++*n_tris;
      f3 a, b, c;
      load_leaf(bvh.leaf_tris, ~node, &a, &b, &c);
      float t, u, v;
      if (mt_test(a, b, c, r, r.tmax, &t, &u, &v)) return true;
    } else {
      if (COUNT) ++*n_nodes;
      const FatNode fn = bvh.fat[node];
      float e0, e1;
      const bool s0 = slab_static(r, fn.a.x, fn.a.y, fn.a.z, fn.a.w, fn.b.x, fn.b.y, &e0) &&
                      e0 < r.tmax;
      const bool s1 = slab_static(r, fn.b.z, fn.b.w, fn.c.x, fn.c.y, fn.c.z, fn.c.w, &e1) &&
                      e1 < r.tmax;
      if (s0 && s1) {
        st.push(fn.d.y, e1);
        node = fn.d.x;
        continue;
      }
      if (s0) {
        node = fn.d.x;
        continue;
      }
      if (s1) {
        node = fn.d.y;
        continue;
      }
    }
    int n;
    float en;
    if (!st.pop(&n, &en)) break;
    node = n;
  }
  return false;
}

// Spill capacity for a tree of the given depth (0: unsupported depth).
inline int spill_for_depth(uint32_t depth) {
  if (depth <= (uint32_t)kLdsStack + 48) return 48;
  if (depth <= 1024) return 1024;
  return 0;
}

}  // namespace ajg
