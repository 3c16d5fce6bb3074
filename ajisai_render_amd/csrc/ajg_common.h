// ajg_common.h — shared device/host helpers of the gfx950 path-tracing hot
// path. Every arithmetic helper keeps the operation order of the reference
// math templates, because that order fixes the FP32 bits that the parity
// tests compare:
//   Min/Max compare-select           ajisai/math/vector.h:35-45
//   min()/max() fold x -> y -> z     ajisai/math/vector.h:381-401
//   dot = ((0 + a0b0) + a1b1) + a2b2 ajisai/math/vector.h:62-69
//   cross                            ajisai/math/vector3.h:89-94
//   normalized = v * (1 / sqrt(v.v)) ajisai/math/vector.h:205-216
//   Join / area / center             ajisai/math/bounds.h:94-113
// All code is compiled with -ffp-contract=off and correctly rounded FP32
// division and square root (no fast-math), so no FMA is formed and `1/x`,
// `sqrtf` round like the x86-64 reference build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define AJG_HD __host__ __device__ __forceinline__

namespace ajg {

struct f3 {
  float x, y, z;
};

AJG_HD f3 mk3(float x, float y, float z) {
  f3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
AJG_HD float fmin_r(float a, float b) { return b < a ? b : a; }
AJG_HD float fmax_r(float a, float b) { return a < b ? b : a; }
// std::max(a, b) == (a < b) ? b : a; argument order decides NaN / -0.
AJG_HD float std_max(float a, float b) { return (a < b) ? b : a; }
AJG_HD float clamp_r(float v, float lo, float hi) {
  return v < lo ? lo : (hi < v ? hi : v);
}
AJG_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }
AJG_HD f3 add(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
AJG_HD f3 sub(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
AJG_HD f3 mul(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
AJG_HD f3 scale(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
AJG_HD f3 divs(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
AJG_HD f3 neg(f3 a) { return mk3(-a.x, -a.y, -a.z); }
AJG_HD float dot(f3 a, f3 b) {
  float out = 0.f;
  out += a.x * b.x;
  out += a.y * b.y;
  out += a.z * b.z;
  return out;
}
AJG_HD f3 cross(f3 a, f3 b) {
  return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
AJG_HD float length(f3 a) { return sqrtf(dot(a, a)); }
AJG_HD f3 normalize(f3 a) { return scale(a, 1.f / length(a)); }
AJG_HD float maxc(f3 a) { return fmax_r(fmax_r(a.x, a.y), a.z); }
AJG_HD float minc(f3 a) { return fmin_r(fmin_r(a.x, a.y), a.z); }
AJG_HD f3 vmin(f3 a, f3 b) { return mk3(fmin_r(a.x, b.x), fmin_r(a.y, b.y), fmin_r(a.z, b.z)); }
AJG_HD f3 vmax(f3 a, f3 b) { return mk3(fmax_r(a.x, b.x), fmax_r(a.y, b.y), fmax_r(a.z, b.z)); }
AJG_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
AJG_HD f3 ld3(const float *p) { return mk3(p[0], p[1], p[2]); }
AJG_HD void st3(float *p, f3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

struct box3 {
  f3 mn, mx;
};

AJG_HD box3 box_empty() {
  box3 b;
  b.mn = mk3(3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f);
  b.mx = mk3(-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f);
  return b;
}
AJG_HD bool box_is_point(const box3 &b) {
  return b.mn.x == b.mx.x && b.mn.y == b.mx.y && b.mn.z == b.mx.z;
}
// Join (bounds.h:108-113): a point box yields the other operand.
AJG_HD box3 box_join(const box3 &a, const box3 &b) {
  if (box_is_point(a)) return b;
  if (box_is_point(b)) return a;
  box3 r;
  r.mn = vmin(a.mn, b.mn);
  r.mx = vmax(a.mx, b.mx);
  return r;
}
AJG_HD float box_area(const box3 &b) {
  f3 d = sub(b.mx, b.mn);
  return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
AJG_HD f3 box_center(const box3 &b) { return divs(add(b.mx, b.mn), 2.f); }
// Triangle::AABB (ploc_aggregate.cpp:168-182)
AJG_HD box3 tri_box(f3 a, f3 b, f3 c) {
  box3 r = box_empty();
  r.mn = vmin(r.mn, a);
  r.mx = vmax(r.mx, a);
  r.mn = vmin(r.mn, b);
  r.mx = vmax(r.mx, b);
  r.mn = vmin(r.mn, c);
  r.mx = vmax(r.mx, c);
  return r;
}

// Node as built by PLOC (ploc_aggregate.cpp:146-152), 32 bytes; same layout
// as ajg_bvh_node of the C-ABI.
struct PlocNode {
  float mn[3], mx[3];
  uint32_t count, first;
};

AJG_HD box3 node_box(const PlocNode &n) {
  box3 b;
  b.mn = ld3(n.mn);
  b.mx = ld3(n.mx);
  return b;
}
AJG_HD void node_set_box(PlocNode &n, const box3 &b) {
  st3(n.mn, b.mn);
  st3(n.mx, b.mx);
}

// Traversal node: both child boxes in the parent (64 bytes, one 4 x 16-byte
// fetch). child >= 0: interior fat node index; child < 0: leaf slot ~child.
struct FatNode {
  float4 a;  // c0.mn.xyz, c0.mx.x
  float4 b;  // c0.mx.yz, c1.mn.xy
  float4 c;  // c1.mn.z, c1.mx.xyz
  int4 d;    // child0, child1, parent (fat index, -1 at root), unused
};

// Traversal node (64 bytes, one per interior node, same index as the fat
// node): the node's grandchildren in left-first order — for each child, its
// two children if it is interior, else the child itself — so 2 to 4 slots
// and one fetch per two tree levels. Skipping the children's own box tests
// keeps every result: a grandchild's box lies inside its parent's, so its
// test fails wherever the parent's would (and the traversal only ever
// visits more nodes than the reference, never fewer or in another order).
//
// The slot boxes are quantised: per axis a grid of step 2^e anchored at
// o (a multiple of 2^e), planes o + q 2^e with q in 0..255, each box rounded
// outward (lower planes down, upper planes up) and checked at build time, so
// a decoded box (fmaf(q, 2^e, o), the same operation in the builder and the
// traversal) contains the exact PLOC box. A slab test of a containing box
// passes wherever the exact one does; leaves are tested against their exact
// boxes (the triangle's AABB) at visit time, which keeps hits, t, u, v
// bit-identical to the reference.
static constexpr int kNoRef = (int)0x80000000;  // empty slot
struct WideNode {
  float4 o;  // grid origin xyz; w: the bits (e_x + 127) | (e_y + 127) << 8 | (e_z + 127) << 16
  uint4 qa;  // plane words lo x, lo y, lo z, hi x (byte k of a word: slot k's plane)
  uint4 qb;  // plane words hi y, hi z in .x, .y
  int4 r;    // slot refs (fat index, ~leaf slot, or kNoRef), left-first order
};
static_assert(sizeof(WideNode) == 64, "64-byte traversal node");

// Plane k (0..23) of a node's quantised boxes: byte k of qa.xyzw, qb.xy.
__device__ __forceinline__ float qplane(uint32_t word, int byte, float step, float origin) {
  return fmaf((float)((word >> (8 * byte)) & 0xffu), step, origin);
}
__device__ __forceinline__ float qstep(float4 o, int axis) {
  return __uint_as_float(((__float_as_uint(o.w) >> (8 * axis)) & 0xffu) << 23);
}

static constexpr uint32_t kNone = 0xFFFFFFFFu;

}  // namespace ajg
