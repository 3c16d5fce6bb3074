/*
 * PLOC build and BVH traversal of the oracle — TEST INFRASTRUCTURE (see orc.h).
 * Restates ajisai/core/aggregate/ploc_aggregate.cpp.
 */
#include <stdlib.h>

#include "orc.h"
#include "orc_math.h"

/* MortonSplit<uint32_t> (ploc_aggregate.cpp:37-49): spreads the low 16 bits
 * so that bit k lands at bit 3k, deriving each mask from the previous one. */
uint32_t orc_morton_split(uint32_t x) {
  const unsigned kBitCount = 32, kLogBits = 5;
  uint32_t mask = 0xFFFFFFFFu >> (kBitCount / 2);
  x &= mask;
  unsigned i = kLogBits - 1, n = 1u << i;
  for (; i > 0; --i, n >>= 1) {
    mask = (mask | (mask << n)) & ~(mask << (n / 2));
    x = (x | (x << n)) & mask;
  }
  return x;
}

/* MortonEncode (ploc_aggregate.cpp:51-54) */
uint32_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z) {
  return orc_morton_split(x) | (orc_morton_split(y) << 1) |
         (orc_morton_split(z) << 2);
}

static inline uint32_t rd_u32(const void *base, size_t off) {
  uint32_t v;
  memcpy(&v, (const char *)base + off, 4);
  return v;
}
static inline v3 rd_v3(const void *base, size_t off) {
  float f[3];
  memcpy(f, (const char *)base + off, 12);
  return V3(f[0], f[1], f[2]);
}

/* PLOCAggregate::Build flatten (ploc_aggregate.cpp:255-286). The scene box is
 * the Join of primitive AABBs; for Quad (quad.cpp:92-104), Cube
 * (cube.cpp:123-130) and MeshGeo (mesh.cpp:64-67) the primitive AABB is the
 * Min/Max fold over every position of the MeshView position buffer. */
uint32_t orc_flatten(const orc_mesh_view *views, uint32_t n_views, float *tri,
                     float *scene_box) {
  box3 global = box_empty();
  uint32_t t = 0;
  for (uint32_t p = 0; p < n_views; ++p) {
    const orc_mesh_view *mv = &views[p];
    box3 pb = box_empty();
    for (size_t v = 0; v < mv->position_size; ++v) {
      v3 pos = rd_v3(mv->position_buffer,
                     mv->position_offset + v * mv->position_stride);
      pb.mn = vmin(pb.mn, pos);
      pb.mx = vmax(pb.mx, pos);
    }
    global = box_join(global, pb);
    for (size_t i = 0; i < mv->index_size; ++i) {
      size_t base = mv->index_offset + i * mv->index_stride;
      uint32_t i1 = rd_u32(mv->index_buffer, base + 0);
      uint32_t i2 = rd_u32(mv->index_buffer, base + 4);
      uint32_t i3 = rd_u32(mv->index_buffer, base + 8);
      if (tri) {
        v3 a = rd_v3(mv->position_buffer,
                     mv->position_offset + (size_t)i1 * mv->position_stride);
        v3 b = rd_v3(mv->position_buffer,
                     mv->position_offset + (size_t)i2 * mv->position_stride);
        v3 c = rd_v3(mv->position_buffer,
                     mv->position_offset + (size_t)i3 * mv->position_stride);
        vstore(tri + 9 * (size_t)t + 0, a);
        vstore(tri + 9 * (size_t)t + 3, b);
        vstore(tri + 9 * (size_t)t + 6, c);
      }
      ++t;
    }
  }
  if (scene_box) {
    vstore(scene_box, global.mn);
    vstore(scene_box + 3, global.mx);
  }
  return t;
}

static inline box3 tri_box_at(const float *tri, uint32_t i) {
  const float *p = tri + 9 * (size_t)i;
  return tri_box(vload(p), vload(p + 3), vload(p + 6));
}

/* float -> u32 truncation as baseline x86-64 performs it for a non-negative
 * float (cvttss2si into a 64-bit register, low 32 bits kept): NaN, +inf and
 * values >= 2^63 give the "integer indefinite" 0x8000000000000000, i.e. 0. */
static inline uint32_t f2u_trunc(float g) {
  if (!(g < 9.2233720e18f)) return 0u;
  return (uint32_t)(int64_t)g;
}

/* MortonEncoder (ploc_aggregate.cpp:56-82) applied to triangle AABB centres
 * (ploc_aggregate.cpp:530-541). */
void orc_morton_codes(const float *tri, uint32_t n, const float *scene_box,
                      uint32_t *codes) {
  const float grid = 1024.f;
  v3 mn = vload(scene_box), mx = vload(scene_box + 3);
  v3 size = vsub(mx, mn);
  v3 inv = V3(1.f / size.x, 1.f / size.y, 1.f / size.z);
  v3 w2g = vscale(inv, grid);
  v3 off = vmul(vneg(mn), w2g);
  for (uint32_t i = 0; i < n; ++i) {
    v3 c = box_center(tri_box_at(tri, i));
    v3 g = vadd(vmul(c, w2g), off);
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float gk = vget(g, k);
      float m = gk < 0.f ? 0.f : gk; /* std::max(g, 0.f) */
      uint32_t u = f2u_trunc(m);
      q[k] = u < 1023u ? u : 1023u; /* std::min(grid-1, u) */
    }
    codes[i] = orc_morton_encode(q[0], q[1], q[2]);
  }
}

/* RadixSort<10>::Sort (ploc_aggregate.cpp:84-132): stable LSD passes of 10
 * bits; result is left in keys/vals. */
void orc_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t n,
                    uint32_t bit_count) {
  enum { kBits = 10, kBuckets = 1 << kBits };
  uint32_t *kc = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *vc = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *k = keys, *v = vals;
  size_t buckets[kBuckets], accum[kBuckets];
  for (uint32_t bit = 0; bit < bit_count; bit += kBits) {
    memset(buckets, 0, sizeof(buckets));
    for (uint32_t i = 0; i < n; ++i) buckets[(k[i] >> bit) & (kBuckets - 1)]++;
    size_t sum = 0;
    for (int b = 0; b < kBuckets; ++b) {
      accum[b] = sum;
      sum += buckets[b];
    }
    for (uint32_t i = 0; i < n; ++i) {
      size_t j = accum[(k[i] >> bit) & (kBuckets - 1)]++;
      kc[j] = k[i];
      vc[j] = v[i];
    }
    uint32_t *t = k;
    k = kc;
    kc = t;
    t = v;
    v = vc;
    vc = t;
  }
  if (k != keys) {
    memcpy(keys, k, sizeof(uint32_t) * n);
    memcpy(vals, v, sizeof(uint32_t) * n);
    free(k);
    free(v);
  } else {
    free(kc);
    free(vc);
  }
}

static inline box3 node_box(const orc_node *nd) {
  box3 b = {vload(nd->mn), vload(nd->mx)};
  return b;
}
static inline void node_set_box(orc_node *nd, box3 b) {
  vstore(nd->mn, b.mn);
  vstore(nd->mx, b.mx);
}

/* Cluster nearest-neighbour search (ploc_aggregate.cpp:400-459). The
 * distance of a pair is Join(lower index, higher index).area(); the scan
 * runs j ascending with a strict '<', so ties go to the lowest j. */
void orc_ploc_nearest(const orc_node *in, uint32_t begin, uint32_t end,
                      uint32_t r, uint32_t *neighbor) {
  for (uint32_t i = begin; i < end; ++i) {
    uint32_t sb = i > begin + r ? i - r : begin;
    uint32_t se = i + r + 1 < end ? i + r + 1 : end;
    float best = FLT_MAX;
    uint32_t best_j = 0xFFFFFFFFu;
    box3 bi = node_box(&in[i]);
    for (uint32_t j = sb; j < se; ++j) {
      if (j == i) continue;
      box3 bj = node_box(&in[j]);
      float d = j < i ? box_area(box_join(bj, bi)) : box_area(box_join(bi, bj));
      if (d < best) {
        best = d;
        best_j = j;
      }
    }
    neighbor[i] = best_j;
  }
}

/* PLOCAggregate::Build (ploc_aggregate.cpp:252-327) over a flattened
 * triangle array. */
int orc_ploc_build(const float *tri, uint32_t n, const float *scene_box,
                   uint32_t r, orc_node *nodes_out, uint32_t *perm) {
  if (n == 0) return -1;
  uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * n);
  orc_morton_codes(tri, n, scene_box, codes);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  orc_radix_sort(codes, perm, n, 30);
  free(codes);

  uint32_t node_count = 2 * n - 1;
  orc_node *nodes = (orc_node *)calloc(node_count, sizeof(orc_node));
  orc_node *copy = (orc_node *)calloc(node_count, sizeof(orc_node));
  uint32_t *neighbor = (uint32_t *)malloc(sizeof(uint32_t) * node_count);
  uint32_t *merged = (uint32_t *)malloc(sizeof(uint32_t) * node_count);
  uint32_t begin = node_count - n, end = node_count, pre_end = end;
  for (uint32_t i = 0; i < n; ++i) {
    orc_node *nd = &nodes[begin + i];
    node_set_box(nd, tri_box_at(tri, perm[i]));
    nd->count = 1;
    nd->first = i;
  }
  int iters = 0;
  while (end - begin > 1) {
    const orc_node *in = nodes;
    orc_node *out = copy;
    orc_ploc_nearest(in, begin, end, r, neighbor);
    for (uint32_t i = begin; i < end; ++i) {
      uint32_t j = neighbor[i];
      int mergeable = j != 0xFFFFFFFFu && neighbor[j] == i;
      merged[i] = (i < j && mergeable) ? 1u : 0u;
    }
    for (uint32_t i = begin + 1; i < end; ++i) merged[i] += merged[i - 1];
    uint32_t merged_count = merged[end - 1];
    if (merged_count == 0) {
      iters = -1;
      break;
    }
    uint32_t unmerged_count = end - begin - merged_count;
    uint32_t children_count = merged_count * 2;
    uint32_t children_begin = end - children_count;
    uint32_t unmerged_begin = end - (children_count + unmerged_count);
    for (uint32_t i = begin; i < end; ++i) {
      uint32_t j = neighbor[i];
      if (j != 0xFFFFFFFFu && neighbor[j] == i) {
        if (i < j) {
          orc_node *un = &out[unmerged_begin + j - begin - merged[j]];
          uint32_t first_child = children_begin + (merged[i] - 1) * 2;
          node_set_box(un, box_join(node_box(&in[j]), node_box(&in[i])));
          un->count = 0;
          un->first = first_child;
          out[first_child + 0] = in[i];
          out[first_child + 1] = in[j];
        }
      } else {
        out[unmerged_begin + i - begin - merged[i]] = in[i];
      }
    }
    for (uint32_t i = end; i < pre_end; ++i) out[i] = in[i];
    orc_node *t = nodes;
    nodes = copy;
    copy = t;
    pre_end = end;
    begin = unmerged_begin;
    end = children_begin;
    ++iters;
  }
  if (iters >= 0) memcpy(nodes_out, nodes, sizeof(orc_node) * node_count);
  free(nodes);
  free(copy);
  free(neighbor);
  free(merged);
  return iters;
}

/* ---- traversal ---------------------------------------------------------- */

typedef struct tray {
  v3 o, d, inv;
  int neg[3];
  float tmin, tmax;
} tray;

static inline tray make_tray(const orc_ray *r) {
  tray t;
  t.o = vload(r->o);
  t.d = vload(r->d);
  /* inv_dir = 1 / d, dir_is_neg = inv_dir < 0 (ploc_aggregate.cpp:331-332) */
  t.inv = V3(1.f / t.d.x, 1.f / t.d.y, 1.f / t.d.z);
  t.neg[0] = t.inv.x < 0.f;
  t.neg[1] = t.inv.y < 0.f;
  t.neg[2] = t.inv.z < 0.f;
  t.tmin = r->tmin;
  t.tmax = r->tmax;
  return t;
}

/* IntersectP (ploc_aggregate.cpp:154-162) */
static inline int slab(const orc_node *nd, const tray *r) {
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    float a = nd->mn[k], b = nd->mx[k];
    float near_plane = r->neg[k] ? b : a;
    float far_plane = r->neg[k] ? a : b;
    lo[k] = (near_plane - vget(r->o, k)) * vget(r->inv, k);
    hi[k] = (far_plane - vget(r->o, k)) * vget(r->inv, k);
  }
  float tmin_max = fmax_r(fmax_r(lo[0], lo[1]), lo[2]);
  float tmax_min = fmin_r(fmin_r(hi[0], hi[1]), hi[2]);
  if (tmin_max > tmax_min) return 0;
  return (tmin_max < r->tmax) && (tmax_min > r->tmin);
}

/* Triangle::Intersect without PostIntersect (ploc_aggregate.cpp:184-216):
 * returns 1 and t/u/v on an accepted hit (t in (tmin, tmax), t < best). */
static inline int mt(const float *p, const tray *r, float best, float *t_out,
                     float *u_out, float *v_out) {
  v3 v0 = vload(p), v1 = vload(p + 3), v2 = vload(p + 6);
  v3 e0 = vsub(v1, v0);
  v3 e1 = vsub(v2, v0);
  v3 h = vcross(r->d, e1);
  float a = vdot(e0, h);
  if (a > -1e-6f && a < 1e-6f) return 0;
  float f = 1.f / a;
  v3 s = vsub(r->o, v0);
  float u = f * vdot(s, h);
  if (u < 0.f || u > 1.f) return 0;
  v3 q = vcross(s, e0);
  float v = f * vdot(r->d, q);
  if (v < 0.f || u + v > 1.f) return 0;
  float t = f * vdot(e1, q);
  if (t > r->tmin && t < r->tmax && t < best) {
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return 1;
  }
  return 0;
}

typedef struct trav_ctx {
  const orc_node *nodes;
  const uint32_t *perm;
  const float *tri;
  uint32_t nv, nt;
} trav_ctx;

/* IntersectAux (ploc_aggregate.cpp:367-391): recursive left-first DFS,
 * ray.t_max shrinks to each accepted hit. */
static int intersect_aux(trav_ctx *c, tray *r, orc_hit *h, uint32_t ni) {
  const orc_node *nd = &c->nodes[ni];
  c->nv++;
  if (!slab(nd, r)) return 0;
  if (nd->count != 0) {
    uint32_t id = c->perm[nd->first];
    float t, u, v;
    c->nt++;
    if (mt(c->tri + 9 * (size_t)id, r, h->t, &t, &u, &v)) {
      h->t = t;
      h->u = u;
      h->v = v;
      h->id = id;
      r->tmax = t;
      return 1;
    }
    return 0;
  }
  int left = intersect_aux(c, r, h, nd->first + 0);
  int right = intersect_aux(c, r, h, nd->first + 1);
  return left || right;
}

/* OccludeAux (ploc_aggregate.cpp:345-365) with Triangle::Occlude (:218-242). */
static int occlude_aux(trav_ctx *c, const tray *r, uint32_t ni) {
  const orc_node *nd = &c->nodes[ni];
  if (!slab(nd, r)) return 0;
  if (nd->count != 0) {
    float t, u, v;
    return mt(c->tri + 9 * (size_t)c->perm[nd->first], r, INFINITY, &t, &u, &v);
  }
  return occlude_aux(c, r, nd->first + 0) || occlude_aux(c, r, nd->first + 1);
}

void orc_trace_closest(const orc_node *nodes, const uint32_t *perm,
                       const float *tri, const orc_ray *rays, uint32_t n,
                       orc_hit *hits) {
  trav_ctx c = {nodes, perm, tri, 0, 0};
  for (uint32_t i = 0; i < n; ++i) {
    tray r = make_tray(&rays[i]);
    orc_hit h = {INFINITY, 0xFFFFFFFFu, 0.f, 0.f};
    intersect_aux(&c, &r, &h, 0);
    hits[i] = h;
  }
}

void orc_trace_closest_counts(const orc_node *nodes, const uint32_t *perm,
                              const float *tri, const orc_ray *rays,
                              uint32_t n, uint32_t *nodes_visited,
                              uint32_t *tris_tested) {
  for (uint32_t i = 0; i < n; ++i) {
    trav_ctx c = {nodes, perm, tri, 0, 0};
    tray r = make_tray(&rays[i]);
    orc_hit h = {INFINITY, 0xFFFFFFFFu, 0.f, 0.f};
    intersect_aux(&c, &r, &h, 0);
    nodes_visited[i] = c.nv;
    tris_tested[i] = c.nt;
  }
}

void orc_trace_any(const orc_node *nodes, const uint32_t *perm,
                   const float *tri, const orc_ray *rays, uint32_t n,
                   uint8_t *occluded) {
  trav_ctx c = {nodes, perm, tri, 0, 0};
  for (uint32_t i = 0; i < n; ++i) {
    tray r = make_tray(&rays[i]);
    occluded[i] = (uint8_t)occlude_aux(&c, &r, 0);
  }
}

void orc_brute_closest(const float *tri, uint32_t n_tri, const orc_ray *rays,
                       uint32_t n, orc_hit *hits) {
  for (uint32_t i = 0; i < n; ++i) {
    tray r = make_tray(&rays[i]);
    orc_hit h = {INFINITY, 0xFFFFFFFFu, 0.f, 0.f};
    for (uint32_t k = 0; k < n_tri; ++k) {
      float t, u, v;
      if (mt(tri + 9 * (size_t)k, &r, h.t, &t, &u, &v)) {
        h.t = t;
        h.u = u;
        h.v = v;
        h.id = k;
        r.tmax = t;
      }
    }
    hits[i] = h;
  }
}

/* ---- two-level traversal (instancing) -----------------------------------
 * No reference counterpart (the reference flattens all primitives into one
 * tree). Restates the device's two-level traversal in the recursive form of
 * IntersectAux/OccludeAux: the top-level tree's leaves are instances; at an
 * instance the ray is mapped into the mesh's space by the instance's
 * world->object map with per-row order ((m0 x + m1 y) + m2 z) + m3 and the
 * direction left unnormalised, then the mesh's tree is traversed with the
 * same t_max. The closest hit is the one with the least (t, instance
 * number), and within an instance the one-level traversal's: an instance
 * numbered below the current hit's also accepts boxes and triangles at
 * t == t_max until it records a hit, and top-level boxes are kept at
 * entry == t_max, so the result does not depend on the top-level visit
 * order (the device visits it nearest first, this restatement left first).
 * Pinned by: an identity instance equals the one-level traversal bit for
 * bit; against a brute-force MT over the world-space copy of every
 * instance, hit ids agree except at near-ties and t within a relative
 * tolerance; coincident instances resolve ties to the lower number. */
typedef struct ctx2 {
  const orc_node *tlas;
  const uint32_t *tlas_perm;
  const float *w2o;
  const uint32_t *inst_mesh;
  const orc_blas *blas;
} ctx2;

static tray map_ray(const ctx2 *c, uint32_t inst, const tray *rw) {
  const float *m = c->w2o + 12 * (size_t)inst;
  orc_ray rr;
  for (int k = 0; k < 3; ++k) {
    const float *row = m + 4 * k;
    rr.o[k] = ((row[0] * rw->o.x + row[1] * rw->o.y) + row[2] * rw->o.z) + row[3];
    rr.d[k] = (row[0] * rw->d.x + row[1] * rw->d.y) + row[2] * rw->d.z;
  }
  rr.tmin = rw->tmin;
  rr.tmax = rw->tmax;
  return make_tray(&rr);
}

/* slab and mt that also accept an entry / t equal to t_max when `incl` */
static inline int slab_incl(const orc_node *nd, const tray *r, int incl) {
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    float a = nd->mn[k], b = nd->mx[k];
    float near_plane = r->neg[k] ? b : a;
    float far_plane = r->neg[k] ? a : b;
    lo[k] = (near_plane - vget(r->o, k)) * vget(r->inv, k);
    hi[k] = (far_plane - vget(r->o, k)) * vget(r->inv, k);
  }
  float tmin_max = fmax_r(fmax_r(lo[0], lo[1]), lo[2]);
  float tmax_min = fmin_r(fmin_r(hi[0], hi[1]), hi[2]);
  if (tmin_max > tmax_min) return 0;
  return (tmin_max < r->tmax || (incl && tmin_max == r->tmax)) && (tmax_min > r->tmin);
}
static inline int mt_incl(const float *p, const tray *r, float best, float *t_out,
                          float *u_out, float *v_out, int incl) {
  float t, u, v;
  tray rr = *r;
  rr.tmax = INFINITY;
  if (!mt(p, &rr, INFINITY, &t, &u, &v)) return 0;
  if (!((t < r->tmax && t < best) || (incl && t == r->tmax && t == best))) return 0;
  *t_out = t;
  *u_out = u;
  *v_out = v;
  return 1;
}
/* IntersectAux of one instance's mesh; *tie: the instance may take a tie at
 * t_max (cleared when it records a hit). */
static int intersect_tie_aux(trav_ctx *c, tray *r, orc_hit *h, uint32_t ni, int *tie) {
  const orc_node *nd = &c->nodes[ni];
  if (!slab_incl(nd, r, *tie)) return 0;
  if (nd->count != 0) {
    uint32_t id = c->perm[nd->first];
    float t, u, v;
    if (mt_incl(c->tri + 9 * (size_t)id, r, h->t, &t, &u, &v, *tie)) {
      h->t = t;
      h->u = u;
      h->v = v;
      h->id = id;
      r->tmax = t;
      *tie = 0;
      return 1;
    }
    return 0;
  }
  int left = intersect_tie_aux(c, r, h, nd->first + 0, tie);
  int right = intersect_tie_aux(c, r, h, nd->first + 1, tie);
  return left || right;
}

static int intersect2_aux(const ctx2 *c, tray *rw, orc_hit *h, uint32_t *hi, uint32_t ni) {
  const orc_node *nd = &c->tlas[ni];
  if (!slab_incl(nd, rw, 1)) return 0;
  if (nd->count != 0) {
    const uint32_t inst = c->tlas_perm[nd->first];
    tray ro = map_ray(c, inst, rw);
    const orc_blas *b = &c->blas[c->inst_mesh[inst]];
    trav_ctx bc = {b->nodes, b->perm, b->tri, 0, 0};
    int tie = *hi != 0xFFFFFFFFu && inst < *hi;
    if (intersect_tie_aux(&bc, &ro, h, 0, &tie)) {
      rw->tmax = ro.tmax;
      *hi = inst;
      return 1;
    }
    return 0;
  }
  int left = intersect2_aux(c, rw, h, hi, nd->first + 0);
  int right = intersect2_aux(c, rw, h, hi, nd->first + 1);
  return left || right;
}

static int occlude2_aux(const ctx2 *c, const tray *rw, uint32_t ni) {
  const orc_node *nd = &c->tlas[ni];
  if (!slab(nd, rw)) return 0;
  if (nd->count != 0) {
    const uint32_t inst = c->tlas_perm[nd->first];
    tray ro = map_ray(c, inst, rw);
    const orc_blas *b = &c->blas[c->inst_mesh[inst]];
    trav_ctx bc = {b->nodes, b->perm, b->tri, 0, 0};
    return occlude_aux(&bc, &ro, 0);
  }
  return occlude2_aux(c, rw, nd->first + 0) || occlude2_aux(c, rw, nd->first + 1);
}

void orc_trace_closest_2l(const orc_node *tlas, const uint32_t *tlas_perm, const float *w2o,
                          const uint32_t *inst_mesh, const orc_blas *blas, const orc_ray *rays,
                          uint32_t n, orc_hit *hits, uint32_t *inst_out) {
  const ctx2 c = {tlas, tlas_perm, w2o, inst_mesh, blas};
  for (uint32_t i = 0; i < n; ++i) {
    tray r = make_tray(&rays[i]);
    orc_hit h = {INFINITY, 0xFFFFFFFFu, 0.f, 0.f};
    uint32_t hi = 0xFFFFFFFFu;
    intersect2_aux(&c, &r, &h, &hi, 0);
    hits[i] = h;
    inst_out[i] = hi;
  }
}

void orc_trace_any_2l(const orc_node *tlas, const uint32_t *tlas_perm, const float *w2o,
                      const uint32_t *inst_mesh, const orc_blas *blas, const orc_ray *rays,
                      uint32_t n, uint8_t *occluded) {
  const ctx2 c = {tlas, tlas_perm, w2o, inst_mesh, blas};
  for (uint32_t i = 0; i < n; ++i) {
    tray r = make_tray(&rays[i]);
    occluded[i] = (uint8_t)occlude2_aux(&c, &r, 0);
  }
}
