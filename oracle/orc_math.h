/*
 * Scalar/vector helpers of the oracle — TEST INFRASTRUCTURE (see orc.h).
 * Each helper restates the operation order of the reference math templates,
 * because that order fixes the FP32 results:
 *   Min/Max      compare-select          math/vector.h:35-45
 *   v.min/v.max  fold x -> y -> z        math/vector.h:381-401
 *   dot          ((0 + a0b0) + a1b1) + a2b2   math/vector.h:62-69
 *   cross        math/vector3.h:89-94
 *   normalized   v * (1 / sqrt(dot(v,v)))     math/vector.h:205-216
 *   area/center/Join                     math/bounds.h:94-113
 */
#ifndef AJG_ORC_MATH_H
#define AJG_ORC_MATH_H

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct v3 {
  float x, y, z;
} v3;

static inline v3 V3(float x, float y, float z) {
  v3 r = {x, y, z};
  return r;
}
static inline float fmin_r(float a, float b) { return b < a ? b : a; }
/* std::max(a, b) = (a < b) ? b : a — the argument order decides NaN and -0. */
static inline float std_max(float a, float b) { return (a < b) ? b : a; }
static inline float fmax_r(float a, float b) { return a < b ? b : a; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) {
  float out = 0.f;
  out += a.x * b.x;
  out += a.y * b.y;
  out += a.z * b.z;
  return out;
}
static inline v3 vcross(v3 a, v3 b) {
  return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x,
            a.x * b.y - b.x * a.y);
}
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnormalize(v3 a) { return vscale(a, 1.f / vlength(a)); }
static inline float vmaxc(v3 a) { return fmax_r(fmax_r(a.x, a.y), a.z); }
static inline float vminc(v3 a) { return fmin_r(fmin_r(a.x, a.y), a.z); }
static inline v3 vmin(v3 a, v3 b) {
  return V3(fmin_r(a.x, b.x), fmin_r(a.y, b.y), fmin_r(a.z, b.z));
}
static inline v3 vmax(v3 a, v3 b) {
  return V3(fmax_r(a.x, b.x), fmax_r(a.y, b.y), fmax_r(a.z, b.z));
}
static inline float vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline v3 vload(const float *p) { return V3(p[0], p[1], p[2]); }
static inline void vstore(float *p, v3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

typedef struct box3 {
  v3 mn, mx;
} box3;

static inline box3 box_empty(void) {
  box3 b = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
  return b;
}
static inline int box_is_point(box3 b) {
  return b.mn.x == b.mx.x && b.mn.y == b.mx.y && b.mn.z == b.mx.z;
}
/* Join (bounds.h:108-113): a point box yields the other operand. */
static inline box3 box_join(box3 a, box3 b) {
  if (box_is_point(a)) return b;
  if (box_is_point(b)) return a;
  box3 r = {vmin(a.mn, b.mn), vmax(a.mx, b.mx)};
  return r;
}
static inline float box_area(box3 b) {
  v3 d = vsub(b.mx, b.mn);
  return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
static inline v3 box_center(box3 b) { return vdivs(vadd(b.mx, b.mn), 2.f); }
/* Triangle::AABB (ploc_aggregate.cpp:168-182) */
static inline box3 tri_box(v3 a, v3 b, v3 c) {
  box3 r = box_empty();
  r.mn = vmin(r.mn, a);
  r.mx = vmax(r.mx, a);
  r.mn = vmin(r.mn, b);
  r.mx = vmax(r.mx, b);
  r.mn = vmin(r.mn, c);
  r.mx = vmax(r.mx, c);
  return r;
}

#endif
