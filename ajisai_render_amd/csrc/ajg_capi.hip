// ajg_capi.hip — extern "C" entry points of libajisai_gpu.so: context,
// errors, accelerator build/trace/download and camera set-up. The renderer
// entry points live in ajg_render.hip.
#include <cmath>
#include <cstring>

#include "ajg_internal.h"

namespace ajg {

int fail(ajg_ctx *ctx, int status, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return status;
}

int hip_fail(ajg_ctx *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();  // clear the sticky error of this thread
  return fail(ctx, e == hipErrorOutOfMemory ? AJG_ERR_OOM : AJG_ERR_HIP, m);
}

int dev_alloc(ajg_ctx *ctx, void **p, size_t bytes, const char *what) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ctx, e == hipErrorOutOfMemory ? AJG_ERR_OOM : AJG_ERR_HIP,
                std::string("allocation of ") + std::to_string(bytes) + " bytes for " + what +
                    " failed: " + hipGetErrorString(e));
  }
  return AJG_OK;
}

// Host fault injection (tests of error paths): returns a status to report.
int injected(ajg_ctx *ctx) {
  if (!ctx || ctx->inject_countdown <= 0) return AJG_OK;
  if (--ctx->inject_countdown == 0) {
    int st = ctx->inject_status;
    ctx->inject_status = 0;
    return fail(ctx, st, "injected failure");
  }
  return AJG_OK;
}

static inline uint32_t rd_u32(const void *base, size_t off) {
  uint32_t v;
  memcpy(&v, (const char *)base + off, 4);
  return v;
}
static inline f3 rd_f3(const void *base, size_t off) {
  float f[3];
  memcpy(f, (const char *)base + off, 12);
  return mk3(f[0], f[1], f[2]);
}

// PLOCAggregate::Build flatten (ploc_aggregate.cpp:255-286): triangles in
// primitive order, then MeshView order; the scene box is the Join of the
// primitives' AABBs, which for Quad, Cube and MeshGeo is the Min/Max fold
// over all positions of the MeshView (quad.cpp:92-104, cube.cpp:123-130,
// mesh.cpp:64-67).
int flatten_views(ajg_ctx *ctx, const ajg_mesh_view *views, uint32_t n_views,
                  std::vector<float> *tri, float *box) {
  box3 global = box_empty();
  size_t total = 0;
  for (uint32_t p = 0; p < n_views; ++p) {
    const ajg_mesh_view &mv = views[p];
    if (mv.index_size && (!mv.index_buffer || !mv.position_buffer))
      return fail(ctx, AJG_ERR_INVALID, "mesh view with null buffer");
    if (mv.position_stride < 12 && mv.position_size > 1)
      return fail(ctx, AJG_ERR_INVALID, "position stride below 12 bytes");
    total += mv.index_size;
  }
  if (total > 0x3FFFFFFFu) return fail(ctx, AJG_ERR_INVALID, "too many triangles");
  tri->resize(9 * total);
  size_t t = 0;
  for (uint32_t p = 0; p < n_views; ++p) {
    const ajg_mesh_view &mv = views[p];
    box3 pb = box_empty();
    for (size_t v = 0; v < mv.position_size; ++v) {
      f3 pos = rd_f3(mv.position_buffer, mv.position_offset + v * mv.position_stride);
      pb.mn = vmin(pb.mn, pos);
      pb.mx = vmax(pb.mx, pos);
    }
    global = box_join(global, pb);
    for (size_t i = 0; i < mv.index_size; ++i, ++t) {
      const size_t base = mv.index_offset + i * mv.index_stride;
      for (int k = 0; k < 3; ++k) {
        uint32_t idx = rd_u32(mv.index_buffer, base + 4 * k);
        if (idx >= mv.position_size)
          return fail(ctx, AJG_ERR_INVALID, "vertex index out of range in mesh view");
        f3 pos = rd_f3(mv.position_buffer, mv.position_offset + (size_t)idx * mv.position_stride);
        st3(tri->data() + 9 * t + 3 * k, pos);
      }
    }
  }
  st3(box, global.mn);
  st3(box + 3, global.mx);
  return AJG_OK;
}

int upload_and_build(ajg_ctx *ctx, const std::vector<float> &tri, const float *box,
                     uint32_t r, hipStream_t s, ajg_accel **out) {
  const uint32_t n = (uint32_t)(tri.size() / 9);
  if (n == 0) return fail(ctx, AJG_ERR_INVALID, "accel build: no triangles");
  float *d = nullptr;
  int st = dev_alloc(ctx, (void **)&d, sizeof(float) * tri.size(), "triangle soup");
  if (st) return st;
  hipError_t e = hipMemcpy(d, tri.data(), sizeof(float) * tri.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(ctx, e, "triangle upload");
  }
  st = build_accel(ctx, d, n, box, r, s, out);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return st;
}

}  // namespace ajg

using namespace ajg;

extern "C" {

int ajg_abi_version(void) { return AJG_ABI_VERSION; }

int ajg_ctx_create(int device, ajg_ctx **out) {
  if (!out) return AJG_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) return AJG_ERR_HIP;
  if (device < 0 || device >= count) return AJG_ERR_INVALID;
  e = hipSetDevice(device);
  if (e != hipSuccess) return AJG_ERR_HIP;
  ajg_ctx *c = new ajg_ctx();
  c->device = device;
  // the allocation-free training step's scratch address on this device
  if (guide_scratch_setup(c) != AJG_OK) {
    delete c;
    return AJG_ERR_HIP;
  }
  *out = c;
  return AJG_OK;
}

int ajg_ctx_destroy(ajg_ctx *ctx) {
  if (ctx && ctx->batch_free) ctx->batch_free(ctx->batch_cache);
  delete ctx;
  return AJG_OK;
}

const char *ajg_last_error(const ajg_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int ajg_ctx_set_cancel(ajg_ctx *ctx, int cancel) {
  if (!ctx) return AJG_ERR_INVALID;
  ctx->cancel.store(cancel ? 1 : 0, std::memory_order_relaxed);
  return AJG_OK;
}

int ajg_ctx_inject_failure(ajg_ctx *ctx, int countdown, int status) {
  if (!ctx) return AJG_ERR_INVALID;
  ctx->inject_countdown = countdown;
  ctx->inject_status = status;
  return AJG_OK;
}

int ajg_accel_build(ajg_ctx *ctx, const ajg_mesh_view *views, uint32_t n_views,
                    uint32_t search_radius, ajg_accel **out) {
  if (!ctx || !out || (n_views && !views)) return AJG_ERR_INVALID;
  *out = nullptr;
  if (int st = injected(ctx)) return st;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, AJG_ERR_HIP, "hipSetDevice");
  std::vector<float> tri;
  float box[6];
  if (int st = flatten_views(ctx, views, n_views, &tri, box)) return st;
  return upload_and_build(ctx, tri, box, search_radius, nullptr, out);
}

int ajg_accel_build_soup(ajg_ctx *ctx, const float *tri_dev, uint32_t n,
                         const float *scene_box, uint32_t search_radius, void *stream,
                         ajg_accel **out) {
  if (!ctx || !out || !tri_dev) return AJG_ERR_INVALID;
  *out = nullptr;
  if (int st = injected(ctx)) return st;
  hipStream_t s = (hipStream_t)stream;
  int st = build_accel(ctx, tri_dev, n, scene_box, search_radius, s, out);
  if (st == AJG_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      destroy_accel(*out);
      *out = nullptr;
      return hip_fail(ctx, e, "build synchronize");
    }
  }
  return st;
}

int ajg_accel_destroy(ajg_accel *accel) {
  destroy_accel(accel);
  return AJG_OK;
}

int ajg_accel_get_info(const ajg_accel *a, ajg_accel_info *out) {
  if (!a || !out) return AJG_ERR_INVALID;
  out->tri_count = a->n_tri;
  out->node_count = 2 * a->n_tri - 1;
  out->iterations = a->iterations;
  out->depth = a->depth;
  out->device_bytes = a->bytes;
  for (int i = 0; i < 6; ++i) out->scene_box[i] = a->scene_box[i];
  return AJG_OK;
}

int ajg_accel_download(const ajg_accel *a, ajg_bvh_node *nodes, uint32_t *perm) {
  if (!a) return AJG_ERR_INVALID;
  if (int st = injected(a->ctx)) return st;
  if (nodes) {
    hipError_t e = hipMemcpy(nodes, a->final_nodes, sizeof(PlocNode) * (2 * (size_t)a->n_tri - 1),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(a->ctx, e, "node download");
  }
  if (perm) {
    hipError_t e = hipMemcpy(perm, a->perm, sizeof(uint32_t) * a->n_tri, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(a->ctx, e, "perm download");
  }
  return AJG_OK;
}

int ajg_accel_trace_closest(ajg_accel *a, const ajg_ray *rays, uint32_t n, ajg_hit *hits,
                            void *stream) {
  if (!a || (n && (!rays || !hits))) return AJG_ERR_INVALID;
  if (int st = injected(a->ctx)) return st;
  return launch_trace_closest(a, rays, n, hits, (hipStream_t)stream);
}

int ajg_accel_trace_any(ajg_accel *a, const ajg_ray *rays, uint32_t n, uint8_t *occluded,
                        void *stream) {
  if (!a || (n && (!rays || !occluded))) return AJG_ERR_INVALID;
  if (int st = injected(a->ctx)) return st;
  return launch_trace_any(a, rays, n, occluded, (hipStream_t)stream);
}

int ajg_accel_trace_counts(ajg_accel *a, const ajg_ray *rays, uint32_t n,
                           uint32_t *nodes_fetched, uint32_t *tris_tested, void *stream) {
  if (!a || (n && (!rays || !nodes_fetched || !tris_tested))) return AJG_ERR_INVALID;
  if (int st = injected(a->ctx)) return st;
  return launch_trace_counts(a, rays, n, nodes_fetched, tris_tested, (hipStream_t)stream);
}

// Camera matrices in double precision, rounded to float:
// camera2world = from({right, realUp, backward}, eye) (matrix4.h:145-153),
// raster2camera = camera2screen^-1 * screen2raster^-1 (thin_lens.cpp:41-48)
// with perspectiveProjection(fov, aspect, near 1, far 1000).
int ajg_camera_setup(ajg_camera *cam, const float *pos, const float *look_at, const float *up,
                     float fov_deg, float res_x, float res_y, float lens_radius,
                     float focal_distance) {
  if (!cam || !pos || !look_at || !up || !(res_x > 0) || !(res_y > 0) || !(fov_deg > 0) ||
      !(fov_deg < 180))
    return AJG_ERR_INVALID;
  double b[3] = {(double)pos[0] - look_at[0], (double)pos[1] - look_at[1],
                 (double)pos[2] - look_at[2]};
  auto norm = [](double *v) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
  };
  auto cross = [](const double *x, const double *y, double *o) {
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
  };
  norm(b);
  double u[3] = {up[0], up[1], up[2]}, rt[3], ru[3];
  cross(u, b, rt);
  norm(rt);
  cross(b, rt, ru);
  norm(ru);
  const double c2w[16] = {rt[0], rt[1], rt[2], 0, ru[0], ru[1], ru[2], 0,
                          b[0],  b[1],  b[2],  0, pos[0], pos[1], pos[2], 1};
  // Closed form of raster2camera: raster (x, y, 1) -> screen
  // (2x/W - 1, 1 - 2y/H, 1) -> camera through the inverse projection.
  const double fov = (double)fov_deg * 3.14159265358979323846 / 180.0;
  const double t = std::tan(fov * 0.5), aspect = (double)res_x / (double)res_y;
  const double sx = 1.0 / (t * aspect), sy = -1.0 / t;
  const double near_ = 1.0, far_ = 1000.0;
  const double m22 = near_ / (far_ - near_), m32 = m22 * far_;
  // camera2screen P (column-major): c0=(sx,0,0,0) c1=(0,sy,0,0) c2=(0,0,m22,-1) c3=(0,0,m32,0)
  // P^-1: x = X/sx, y = Y/sy, z = -W, w = (Z + m22*W)/m32 ... expressed column-major:
  const double pinv[16] = {1.0 / sx, 0, 0, 0, 0, 1.0 / sy, 0, 0,
                           0,        0, 0, 1.0 / m32, 0, 0, -1, m22 / m32};
  // screen2raster S = scale(W/2, -H/2, 1) with translation (W/2, H/2, 0);
  // S^-1: screen = (2/W) x - 1, (-2/H) y + 1, z.
  const double sinv[16] = {2.0 / res_x, 0, 0, 0, 0, -2.0 / res_y, 0, 0,
                           0,           0, 1, 0, -1, 1, 0, 1};
  double r2c[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double acc = 0;
      for (int k = 0; k < 4; ++k) acc += pinv[k * 4 + r] * sinv[c * 4 + k];
      r2c[c * 4 + r] = acc;
    }
  for (int i = 0; i < 16; ++i) {
    cam->raster2camera[i] = (float)r2c[i];
    cam->camera2world[i] = (float)c2w[i];
  }
  cam->resolution[0] = res_x;
  cam->resolution[1] = res_y;
  cam->lens_radius = lens_radius;
  cam->focal_distance = focal_distance;
  return AJG_OK;
}

}  // extern "C"
