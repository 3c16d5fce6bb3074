// ajg_scene.h — the device scene shared by the renderers: the scene handle,
// the per-triangle shading records, the scene as kernels see it, and
// PostIntersect.
#pragma once
#include <vector>

#include "ajg_internal.h"
#include "ajg_shade.h"
#include "ajg_traverse.h"

struct ajg_scene {
  ajg_ctx *ctx = nullptr;
  ajg_accel *accel = nullptr;
  ajg::TriShade *tri_shade = nullptr;  // per leaf slot: flags|material and vertex normals
  ajg::DevMaterial *mats = nullptr;
  uint32_t n_mats = 0;
  ajg::DevLight lights[ajg::kMaxLights];
  uint32_t n_lights = 0;
  uint32_t n_area = 0;               // lights[0, n_area): emissive primitives
  int env_index = -1;
  std::vector<void *> light_bufs;    // env tables, area-light triangles
  float world[6] = {0, 0, 0, 0, 0, 0};
  float radius = 0.f;
  int cu_count = 256;
  void *render_cache = nullptr;  // RenderBuffers reused across renders
  void *guide_cache = nullptr;   // GuideBuffers reused across guided renders
  void *sppm_cache = nullptr;    // SppmBuffers reused across SPPM renders
  void *vcm_cache = nullptr;     // VcmBuffers reused across VCM renders
  // instanced scenes: `accel` is the top-level tree over instance boxes
  std::vector<ajg_accel *> blas;   // one bottom-level tree per mesh
  ajg::DevInst *insts = nullptr;
  ajg::DevBlas *blases = nullptr;
  uint32_t n_inst = 0;
  uint32_t depth2 = 0;             // top depth + deepest bottom depth
  std::vector<float> inst_w2o;     // 12 floats per instance (as on the device)
  std::vector<float> inst_box;     // 6 floats per instance (padded world box)
  uint32_t *trace_ctr = nullptr;   // cursor of scene trace calls
};

namespace ajg {

// tri_shade flags
static constexpr uint32_t kKindShift = 28;  // 2 bits
static constexpr uint32_t kTwoSidedBit = 1u << 27;
static constexpr uint32_t kLightShift = 24;  // 3 bits: area light index + 1 (0: none)
static constexpr uint32_t kMatMask = (1u << 24) - 1;

struct DevScene {
  DevBVH bvh;                  // instanced: the top-level tree
  const TriShade *tri_shade;
  const DevInst *inst;         // instanced scenes only (else null)
  const DevBlas *blas;
  const DevMaterial *mats;
  DevLight lights[kMaxLights];
  uint32_t n_lights;
  uint32_t n_area;             // area lights (lights[0, n_area))
  int env_index;
  float radius;
};

struct Inct {
  f3 pos, ng, ns, wr;
  uint32_t material;
  uint32_t light;  // area light index + 1 of an emissive primitive, else 0
};

// MT normal + PostIntersect (ploc_aggregate.cpp:199-211; quad.cpp:73-90,
// cube.cpp:96-108 (tuple order reversed in the index buffer), mesh.cpp:183-204,
// two_sided.cpp:44-53).
// Instanced scenes: the triangle is read in its mesh's object space; the
// hit point comes from the world ray, and normals go to world space through
// the instance's normal map (for a rotation with uniform scale that map is the
// rotation itself, so the shading normal has the length it would have in a
// world-space copy of the mesh).
__device__ __forceinline__ Inct post_intersect(const DevScene &sc, const DevRay &r, float t, int ls,
                                               float u, float v, uint32_t inst = 0) {
  Inct it;
  f3 a, b, c;
  const float4 *leaf = sc.bvh.leaf_tris;
  const TriShade *shade = sc.tri_shade;
  const DevInst *I = nullptr;
  if (sc.inst) {
    I = sc.inst + inst;
    const DevBlas &B = sc.blas[I->info.x];
    leaf = B.bvh.leaf_tris;
    shade = B.tri_shade;
  }
  load_leaf(leaf, ls, &a, &b, &c);
  const TriShade ts = shade[ls];
  const uint4 info = ts.info;
  const uint32_t kind = info.x >> kKindShift;
  it.material = (I && I->info.y != kKeepMaterial) ? I->info.y : (info.x & kMatMask);
  it.light = (info.x >> kLightShift) & 7u;
  it.pos = add(r.o, scale(r.d, t));
  it.wr = neg(r.d);
  if (kind == AJG_GEOM_CUBE) {
    it.ng = normalize(cross(sub(b, c), sub(a, c)));
    it.ns = it.ng;
  } else {
    it.ng = normalize(cross(sub(b, a), sub(c, a)));
    if (kind == AJG_GEOM_MESH) {
      const float w = 1 - u - v;
      const float4 ni = ts.n[0], nj = ts.n[1], nk = ts.n[2];
      it.ns = add(add(scale(mk3(ni.x, ni.y, ni.z), w), scale(mk3(nj.x, nj.y, nj.z), u)),
                  scale(mk3(nk.x, nk.y, nk.z), v));
    } else {
      it.ns = it.ng;
    }
  }
  if (I) {
    const bool interp = kind == AJG_GEOM_MESH;
    it.ng = normalize(xform_vec(I->nrm, it.ng));
    it.ns = interp ? xform_vec(I->nrm, it.ns) : it.ng;
  }
  if ((info.x & kTwoSidedBit) && dot(it.ng, it.wr) < 0) {
    it.ng = scale(it.ng, -1.f);
    it.ns = scale(it.ns, -1.f);
  }
  return it;
}

// Intersection::EpsOffset (intersection.h:42-47)
__device__ __forceinline__ f3 eps_offset(const Inct &it, f3 dir) {
  if (dot(dir, it.ng) > 0.f) return add(it.pos, scale(it.ng, AJG_EPS));
  return sub(it.pos, scale(it.ng, AJG_EPS));
}

__device__ __forceinline__ float lum_f(f3 c) {
  return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
}

}  // namespace ajg

// The scene as kernels take it (by value).
ajg::DevScene dev_scene(const ajg_scene *sc);
// Frees a scene's SPPM buffers (sppm_cache) / VCM buffers (vcm_cache).
void sppm_free(void *cache);
void vcm_free(void *cache);
