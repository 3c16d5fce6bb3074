/*
 * ajisai_gpu.h — C-ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * Plain C types only: opaque handles, pointers and sizes, int status codes.
 * No exception crosses this boundary. Entry points taking a `stream`
 * argument enqueue work on that hipStream_t (NULL = the null stream) and
 * return without synchronising unless stated otherwise.
 *
 * Each declaration names the reference interface it replaces
 * (file:line relative to the reference root, siyuanpan/ajisai_render).
 */
#ifndef AJISAI_GPU_H
#define AJISAI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AJG_ABI_VERSION 1

/* Status codes. The reference has no error returns (noexcept hot calls,
 * std::exit in loaders: factory/creator/geometry_creators.cpp:118-123);
 * every entry point here returns one of these instead. */
enum ajg_status {
  AJG_OK = 0,
  AJG_ERR_HIP = 1,       /* a HIP runtime call failed */
  AJG_ERR_OOM = 2,       /* device allocation failed */
  AJG_ERR_INVALID = 3,   /* bad argument, unsupported type, empty scene */
  AJG_ERR_CANCELLED = 4, /* the context's cancel flag was raised (ajg_ctx_set_cancel) */
  AJG_ERR_INTERNAL = 5   /* an internal limit or invariant was hit */
};

typedef struct ajg_ctx ajg_ctx;
typedef struct ajg_accel ajg_accel;
typedef struct ajg_scene ajg_scene;
typedef struct ajg_guide ajg_guide;

/* ---------------------------------------------------------------- context */
int ajg_abi_version(void);
/* One context per device; owns the default allocations and the error text. */
int ajg_ctx_create(int device, ajg_ctx **out);
int ajg_ctx_destroy(ajg_ctx *ctx);
/* Last error message of this context (valid until the next call). */
const char *ajg_last_error(const ajg_ctx *ctx);
/* Fault injection for host-side failure tests: the next `countdown`-th
 * entry point call on this context returns `status` (0 disables). */
int ajg_ctx_inject_failure(ajg_ctx *ctx, int countdown, int status);
/* Cancellation (no counterpart in the reference, whose Render runs to the
 * end): cancel = 1 raises the context's flag, 0 clears it. Renders on the
 * context (ajg_render_pt, _guided, _batch, _sppm, _vcm) poll it at their
 * host checkpoints (every few wavefront iterations; each SPPM/VCM
 * iteration as it is enqueued), wait for the work in flight and return
 * AJG_ERR_CANCELLED, leaving the film partial. The flag stays raised until
 * cleared. May be called from any thread. */
int ajg_ctx_set_cancel(ajg_ctx *ctx, int cancel);

/* ------------------------------------------------------------ accelerator */

/* MeshView (ajisai/core/geometry/geometry.h:33-42): byte-strided position
 * (>= 3 floats per vertex) and index (3 x u32 per triangle) buffers. */
typedef struct ajg_mesh_view {
  const void *position_buffer;
  size_t position_offset;
  size_t position_stride;
  size_t position_size;
  const void *index_buffer;
  size_t index_offset;
  size_t index_stride;
  size_t index_size;
} ajg_mesh_view;

/* Ray as stored after the Ray constructor (ajisai/core/ray.h:28-45):
 * origin already offset, direction normalised. 32 bytes. */
typedef struct ajg_ray {
  float o[3];
  float d[3];
  float tmin, tmax;
} ajg_ray;

/* Closest-hit record (the t / id / barycentrics that Triangle::Intersect
 * writes into PrimitiveIntersection, ploc_aggregate.cpp:184-216). tri_id is
 * the global flattened triangle index (ploc_aggregate.cpp:257-284);
 * 0xFFFFFFFF and t = +inf on a miss. 16 bytes. */
typedef struct ajg_hit {
  float t;
  uint32_t tri_id;
  float u, v;
} ajg_hit;

/* PLOC node as built (Node, ploc_aggregate.cpp:146-152; size_t fields
 * narrowed to u32). count = 1 for a leaf, 0 for an interior node; first is
 * the first child (interior) or the index into the Morton-sorted primitive
 * order (leaf). The root is node 0. 32 bytes. */
typedef struct ajg_bvh_node {
  float mn[3], mx[3];
  uint32_t count, first;
} ajg_bvh_node;

typedef struct ajg_accel_info {
  uint32_t tri_count;
  uint32_t node_count;     /* 2 * tri_count - 1 */
  uint32_t iterations;     /* PLOC Cluster iterations */
  uint32_t depth;          /* tree depth (root = 1) */
  uint64_t device_bytes;   /* bytes held on the device by the accelerator */
  float scene_box[6];      /* Join of primitive AABBs (ploc_aggregate.cpp:256-258) */
} ajg_accel_info;

/* PLOCAggregate::Build (ploc_aggregate.cpp:252-327) reached through
 * Aggregate::Build (ajisai/core/aggregate/aggregate.h:36) and the factory
 * CreatePLOCAggregate(size_t search_radius) (aggregate.h:45): flattens the
 * views on the host in primitive order, uploads, and builds the PLOC BVH on
 * the device (Morton codes, LDS radix sort, nearest-neighbour merge).
 * Synchronous. search_radius defaults to 14 in the reference factory
 * (factory/creator/aggregate_creators.cpp:43). */
int ajg_accel_build(ajg_ctx *ctx, const ajg_mesh_view *views, uint32_t n_views,
                    uint32_t search_radius, ajg_accel **out);

/* Same build from a device-resident flattened triangle soup (n x 9 floats,
 * a,b,c per triangle, global id = position) and the scene box (host,
 * 6 floats; NULL = Min/Max over all soup vertices, which equals the Join of
 * primitive boxes when every position is referenced). Enqueued on `stream`,
 * then synchronised. This is the path timed for build throughput. */
int ajg_accel_build_soup(ajg_ctx *ctx, const float *tri_dev, uint32_t n,
                         const float *scene_box, uint32_t search_radius,
                         void *stream, ajg_accel **out);
int ajg_accel_destroy(ajg_accel *accel);
int ajg_accel_get_info(const ajg_accel *accel, ajg_accel_info *out);
/* Copy the PLOC node array (2n-1 nodes, root 0) and the Morton-sorted
 * primitive order (primitive_indices_, n entries) to host memory. */
int ajg_accel_download(const ajg_accel *accel, ajg_bvh_node *nodes,
                       uint32_t *perm);

/* Aggregate::Intersect batched (aggregate.h:38-39 -> PLOCAggregate::Intersect,
 * ploc_aggregate.cpp:329-335, IntersectAux :367-391): left-first DFS order
 * with pop-time culling, so hits are bit-identical to the reference's
 * recursive traversal. rays/hits are device pointers. */
int ajg_accel_trace_closest(ajg_accel *accel, const ajg_ray *rays, uint32_t n,
                            ajg_hit *hits, void *stream);
/* Aggregate::Occlude batched (aggregate.h:41 -> ploc_aggregate.cpp:337-365):
 * occluded[i] = 1 if any triangle is hit in (tmin, tmax). Device pointers. */
int ajg_accel_trace_any(ajg_accel *accel, const ajg_ray *rays, uint32_t n,
                        uint8_t *occluded, void *stream);
/* Counting pass of trace_closest for the byte model: per ray, the number of
 * traversal nodes fetched (64-byte quantised nodes, each covering up to
 * two tree levels) and triangles tested. Device pointers; not timed. */
int ajg_accel_trace_counts(ajg_accel *accel, const ajg_ray *rays, uint32_t n,
                           uint32_t *nodes_fetched, uint32_t *tris_tested,
                           void *stream);

/* --------------------------------------------------------------- renderer */

enum ajg_geom_kind { AJG_GEOM_QUAD = 0, AJG_GEOM_CUBE = 1, AJG_GEOM_MESH = 2 };
enum ajg_material_type { AJG_MAT_DIFFUSE = 0, AJG_MAT_DISNEY = 1 };
enum ajg_light_type { AJG_LIGHT_DIRECTIONAL = 0, AJG_LIGHT_ENV = 1 };

/* Primitive (ajisai/core/primitive/primitive.h) reduced to what the hot path
 * reads: the geometry's MeshView, its kind (selects the PostIntersect rule:
 * quad.cpp:73-90, cube.cpp:96-108, mesh.cpp:183-204), the TwoSided wrapper
 * (two_sided.cpp:44-53), the material index and, for meshes, the per-vertex
 * normal_and_vs buffer (mesh.cpp:190-193; 4 floats per vertex). A non-black
 * emission makes the primitive an area light (GeometricPrimitive,
 * geometric.cpp:35-51; AreaLight, area_light.cpp:27-85), listed before the
 * scene's own lights (default_scene.cpp:34-48). */
typedef struct ajg_primitive {
  ajg_mesh_view view;
  uint32_t kind;
  uint32_t two_sided;
  uint32_t material;
  uint32_t reserved;
  const float *normal_and_vs;
  float emission[3];
  uint32_t reserved2;
} ajg_primitive;

/* Material: Diffuse (material/diffuse.cpp:35-48; base_color = albedo) or
 * Disney (material/disney.cpp:497-525; parameter order of CreateDisney,
 * material/material.h:40-54) with Constant2D textures. */
typedef struct ajg_material {
  uint32_t type;
  float base_color[3];
  float subsurface, metallic, specular, specular_tint, roughness, anisotropic,
      sheen, sheen_tint, clearcoat, clearcoat_gloss;
} ajg_material;

/* Light: DirectionalLight (light/directional_light.cpp:30-58; dir points
 * toward the light, angle in radians) or EnvLight with an HDR texture
 * (light/env_light.cpp:29-170; texels w x h RGB float, row-major, row 0 at
 * theta = 0; rot in degrees; scaler). */
typedef struct ajg_light {
  uint32_t type;
  float dir[3];
  float angle;
  float radiance[3];
  uint32_t env_w, env_h;
  const float *env_texels;
  float rot, scaler;
} ajg_light;

/* Thin-lens camera (camera/thin_lens.cpp:32-83): column-major raster-to-
 * camera and camera-to-world matrices, film resolution, lens parameters. */
typedef struct ajg_camera {
  float raster2camera[16];
  float camera2world[16];
  float resolution[2];
  float lens_radius, focal_distance;
} ajg_camera;

/* PTRendererArgs (renderer/renderer.h:37-45) plus the build's options. */
typedef struct ajg_pt_args {
  int32_t spp, tile_size, min_bounces, max_bounces;
  float cont_prob;
  int32_t use_mis;      /* 1: TraceMIS (pt.cpp:98-227); 0: Trace (pt.cpp:229-311), not guided */
  int32_t seed_mode;    /* 0: per-pixel stream x + y*W, samples in sequence
                           (pt.cpp:56-64); 1: one stream per sample */
  uint32_t seed_base;   /* seed_mode 1 only */
  int32_t env_cdf_fp64; /* env CDF prefix sums: 0 float (distribution.h:44-48),
                           1 double rounded to float */
  int32_t max_slots;    /* path slots in flight; 0 = automatic (1 << 20). With
                           seed_mode 1 each pixel's samples are split into work
                           items of equal sample chunks, about 8 items per slot */
  uint32_t flags;       /* AJG_RENDER_* bits */
} ajg_pt_args;

/* ajg_pt_args.flags */
#define AJG_RENDER_TIMING 1u   /* per-stage device times via HIP events */
#define AJG_RENDER_COUNT 2u    /* count node fetches / triangles per query (slower;
                                  for the byte model, not for timing). In
                                  instanced scenes an instance visit counts
                                  as one node fetch (its map and mesh record)
                                  and one triangle (its box) */
#define AJG_RENDER_FINITE_PDF 4u /* end a path whose continuation sample has
                                  pdf 0 (the reference continues it with an
                                  infinite throughput, pt.cpp:216-221; guided
                                  sampling always ends it) */
#define AJG_RENDER_INLINE 8u   /* no counterpart in the reference: trace each
                                  work item's paths to their ends in one
                                  thread, the whole render one kernel, instead
                                  of the wavefront's per-bounce launches. For
                                  renders with few paths in flight (per-pixel
                                  streams on small images), where each
                                  wavefront iteration costs its kernels'
                                  latency floors. Films and per-sample
                                  radiance are bitwise the wavefront's. Not
                                  with instanced scenes, guiding or
                                  AJG_RENDER_COUNT (AJG_ERR_INVALID) */

/* Film accumulators per traced pixel (traced area = whole tiles,
 * pt.cpp:43-49): Kahan-compensated float sums of L and L^2 per channel and
 * the sample count. Device pointers, sized Wp*Hp*3, Wp*Hp*3 and Wp*Hp. */
typedef struct ajg_film {
  float *sum;
  float *sum_sq;
  uint32_t *count;
  float *per_sample; /* optional: L of every sample, (pixel*spp + s)*3 */
} ajg_film;

typedef struct ajg_render_stats {
  uint64_t closest_queries; /* camera + continuation + MIS-BSDF rays */
  uint64_t any_queries;     /* shadow rays */
  uint64_t paths;           /* samples completed */
  uint32_t iterations;      /* wavefront iterations */
  uint32_t slots;
  float ms_total;           /* device time of the render (events) */
  float ms_trace_closest, ms_trace_any, ms_shade; /* AJG_RENDER_TIMING */
  uint32_t launches_closest, launches_any, launches_shade;
  uint32_t reserved;
  uint64_t nodes_fetched;   /* AJG_RENDER_COUNT: node fetches, closest */
  uint64_t tris_tested;     /* AJG_RENDER_COUNT: triangles tested, closest */
  uint64_t nodes_fetched_any, tris_tested_any; /* same for shadow rays */
} ajg_render_stats;

/* DefaultScene (scene/default_scene.cpp:29-103) with a GPU PLOC accelerator:
 * flattens and builds like ajg_accel_build, uploads PostIntersect data,
 * materials and lights, builds the env Distribution2D tables on the device
 * (EnvLight ctor, env_light.cpp:29-48) and processes lights against the
 * world bound. Synchronous. Host pointers are read during the call only. */
int ajg_scene_create(ajg_ctx *ctx, const ajg_primitive *prims,
                     uint32_t n_prims, const ajg_material *mats,
                     uint32_t n_mats, const ajg_light *lights,
                     uint32_t n_lights, uint32_t search_radius,
                     int32_t env_cdf_fp64, ajg_scene **out);
int ajg_scene_destroy(ajg_scene *scene);
/* Borrowed accelerator of the scene (owned by the scene). */
int ajg_scene_get_accel(ajg_scene *scene, ajg_accel **out);
/* Env Distribution2D tables of light li (an index into the `lights` given to
 * ajg_scene_create) as built on the device (host copies): cond_cdf[h][w+1],
 * cond_func_int[h], marg_cdf[h+1], marg_func_int. */
int ajg_scene_env_tables(ajg_scene *scene, uint32_t li, float *cond_cdf,
                         float *cond_func_int, float *marg_cdf,
                         float *marg_func_int);

/* ------------------------------------------------------------- instancing
 * No counterpart in the reference, which flattens every primitive into one
 * PLOC tree (ploc_aggregate.cpp:252-327). An instanced scene keeps one PLOC
 * tree per mesh in the mesh's own space (bottom level) and a PLOC tree over
 * the instances' world boxes (top level); a ray entering an instance is
 * mapped into the mesh's space with its direction left unnormalised, so the
 * ray parameter t is the same at both levels. */
#define AJG_INSTANCE_KEEP_MATERIAL 0xFFFFFFFFu

typedef struct ajg_instance {
  uint32_t mesh;           /* index into the scene's mesh list */
  uint32_t material;       /* material of the instance, or
                              AJG_INSTANCE_KEEP_MATERIAL for the mesh's own */
  float object2world[12];  /* row-major 3x4 affine map (non-singular) */
} ajg_instance;

/* Instanced DefaultScene: each of the n_meshes primitives (any kind; its
 * material and two-sided flag as in ajg_scene_create) becomes one
 * bottom-level tree; instances place meshes in the world. Lights are
 * processed against the union of the instance boxes. The scene renders
 * with ajg_render_pt / ajg_render_pt_guided (not in batches).
 * Synchronous. */
int ajg_scene_create_instanced(ajg_ctx *ctx, const ajg_primitive *meshes,
                               uint32_t n_meshes, const ajg_instance *instances,
                               uint32_t n_instances, const ajg_material *mats,
                               uint32_t n_mats, const ajg_light *lights,
                               uint32_t n_lights, uint32_t search_radius,
                               int32_t env_cdf_fp64, ajg_scene **out);
/* Host copies of the per-instance world->object maps (12 floats, row-major
 * 3x4, as used on the device) and padded world boxes (6 floats); either may
 * be NULL. */
int ajg_scene_instances(ajg_scene *scene, float *world2object, float *boxes);
/* Scene::Intersect / Occlude batched (default_scene.cpp:63-70) for one-level
 * and instanced scenes. hits[i].tri_id is the triangle index within the hit
 * mesh (one-level scenes: the flattened id, as ajg_accel_trace_closest);
 * instance[i] (optional) receives the instance (0 for one-level scenes,
 * 0xFFFFFFFF on a miss). Device pointers; enqueued on `stream`. */
int ajg_scene_trace_closest(ajg_scene *scene, const ajg_ray *rays, uint32_t n,
                            ajg_hit *hits, uint32_t *instance, void *stream);
int ajg_scene_trace_any(ajg_scene *scene, const ajg_ray *rays, uint32_t n,
                        uint8_t *occluded, void *stream);

/* Renderer::Render for PathTracing::TraceMIS (renderer.h:64-65,
 * pt.cpp:40-73 and 98-227) as a wavefront on the device: ray-gen,
 * closest/any-hit traversal, shade/sample, shadow and accumulate stages.
 * Fills `film` (device memory, zeroed by the call). Synchronous; stats may
 * be NULL. */
int ajg_render_pt(ajg_scene *scene, const ajg_camera *camera,
                  const ajg_pt_args *args, const ajg_film *film,
                  ajg_render_stats *stats, void *stream);

/* Film::MergeTile + WriteImage resolve (film.cpp:27-63, film.h:119-152,
 * image.h:38-48): W x H RGB image from the per-traced-pixel sums. Device
 * pointers. */
int ajg_film_resolve(ajg_ctx *ctx, uint32_t w, uint32_t h, uint32_t tile_size,
                     const float *sum, const uint32_t *count, float *rgb,
                     void *stream);

/* Film filters (Filter, core/filter/filter.h): BoxFilter (box_filter.cpp:28-42,
 * weight 1; the reference's default radius 0.5) and GaussianFilter
 * (gaussian_filter.cpp:28-50, max(0, exp(-alpha d^2) - exp(-alpha r^2)) per
 * axis, product; defaults radius 2, alpha 2, filter_creators.cpp:39-50). */
enum ajg_filter_type { AJG_FILTER_BOX = 0, AJG_FILTER_GAUSSIAN = 1 };
typedef struct ajg_filter {
  uint32_t type;
  float radius[2];  /* x, y; in (0, 7.5] */
  float alpha;      /* gaussian only */
} ajg_filter;

/* Tile::AddSample + Film::MergeTile + WriteImage's division (film.cpp:27-63,
 * film.h:119-147, image.h:38-48) for any of the filters above: each traced
 * pixel p adds its sums into [ceil(p - 0.5 - r), floor(p - 0.5 + r)] per axis
 * with the filter's weight at the target's offset, clamped to the film; the
 * image is radiance / weight. The filter weights are evaluated on the host in
 * float (as the reference does); accumulation is in double on the device.
 * Device pointers; enqueued. */
int ajg_film_resolve_filter(ajg_ctx *ctx, uint32_t w, uint32_t h, uint32_t tile_size,
                            const ajg_filter *filter, const float *sum,
                            const uint32_t *count, float *rgb, void *stream);

/* ImageWriter::Write's pixel path (image.cpp:32-62, color.h:54-64): linear
 * RGB (w*h*3 floats) -> 8-bit sRGB (w*h*3 bytes): 12.92 x below 0.0031308,
 * else 1.055 x^(1/2.4) - 0.055; clamp to [0, 1]; round(v * 255.5); flip != 0
 * stores rows bottom to top (stbi_flip_vertically_on_write). NaN -> 0.
 * Device pointers; enqueued. */
int ajg_image_srgb8(ajg_ctx *ctx, uint32_t w, uint32_t h, const float *rgb,
                    int32_t flip, uint8_t *out, void *stream);

/* ThinLensCamera matrices (thin_lens.cpp:32-55) from look-at parameters;
 * fov in degrees. Host only. */
int ajg_camera_setup(ajg_camera *cam, const float *pos, const float *look_at,
                     const float *up, float fov_deg, float res_x, float res_y,
                     float lens_radius, float focal_distance);

/* ---------------------------------------------------------- path guiding
 * No counterpart in the reference: its path tracer samples continuation
 * directions from the BSDF only (renderer/pt.cpp:215-222). A small MLP
 * (x(112) -> 64 -> 64 -> 64 -> 48, ReLU, BF16 weights and activations, FP32
 * accumulation) maps a shading point (position, normal) to an 8-lobe
 * von Mises-Fisher mixture; it is trained online on the renderer's own
 * path vertices with Adam. */

/* Number of FP32 parameters (masters): W1[112][64] b1 W2[64][64] b2
 * W3[64][64] b3 W4[64][48] b4, in that order. */
int ajg_guide_param_count(void);
/* Creates a guide with a seeded He-uniform initialisation. scene_box: host
 * pointer to (min xyz, max xyz), used to normalise positions. */
int ajg_guide_create(ajg_ctx *ctx, uint32_t seed, const float *scene_box,
                     ajg_guide **out);
int ajg_guide_destroy(ajg_guide *guide);
/* Host copies of the FP32 masters (ajg_guide_param_count() floats);
 * set_params also refreshes the BF16 copies. Synchronous. */
int ajg_guide_get_params(ajg_guide *guide, float *params);
int ajg_guide_set_params(ajg_guide *guide, const float *params);
/* Inference. pos, nrm: device, n x 3 floats. mixture (device, n x 40 or
 * NULL): lobe j = (mu xyz, kappa, weight) at 5j. raw (device, n x 48 or
 * NULL): the network outputs. */
int ajg_guide_forward(ajg_guide *guide, const float *pos, const float *nrm,
                      uint32_t n, float *mixture, float *raw, void *stream);
/* One training step on n samples (device, n x 10 floats: position, normal,
 * direction, weight). Loss = -sum w log q(dir) / sum w. lr > 0 applies an
 * Adam step (beta 0.9 / 0.999, eps 1e-8); lr = 0 only computes the
 * gradient. loss_out / grad_out (host, optional) synchronise the stream. */
int ajg_guide_train_step(ajg_guide *guide, const float *samples, uint32_t n,
                         float lr, float *loss_out, float *grad_out,
                         void *stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same
 * training step as ajg_guide_train_step, with bit-identical results, that
 * allocates nothing. Its work buffers (activations, gradient partials) live
 * in a static per-device scratch of 45,219,840 bytes that each GPU the
 * library runs on holds until the process ends; ajg_ctx_create records its
 * address. At most 32768 samples per step (AJG_ERR_INVALID above). Only
 * enqueues on `stream` (no loss/gradient copies, no synchronisation; the
 * updated parameters stay in the guide). Calls that overlap on one GPU
 * share the scratch and race: issue them on one stream, one at a time. */
int ajg_guide_train_step_noalloc(ajg_guide *guide, const float *samples,
                                 uint32_t n, float lr, void *stream);

/* Guided rendering options. */
typedef struct ajg_guide_args {
  float alpha;          /* probability of sampling the guide instead of the
                           BSDF (one-sample MIS: pdf = alpha q + (1-alpha) p) */
  int32_t train;        /* 1: train online on the render's own radiance
                           samples; 0: use the guide as it is */
  uint32_t batch;       /* training records per Adam step (2^18); records
                           collect in the guide across renders until a batch
                           is full */
  float lr;             /* Adam learning rate (1e-3) */
  uint32_t train_slots; /* path slots that record training data (0: all,
                           within a 256 MiB vertex-history budget) */
  uint32_t max_steps;   /* stop training after this many steps (0: none) */
  uint32_t steps_per_batch; /* Adam steps on each batch (0 or 1: one) */
  /* Warm-up: before the render, `warmup_passes` guided passes of
   * `warmup_spp` samples per pixel train the guide (steps_per_batch
   * `warmup_steps_per_batch`); their images are discarded. They draw from
   * random streams independent of the render's own (seed_mode 1 streams
   * from a seed base no render uses), so the trained guide is independent
   * of the final samples. */
  uint32_t warmup_passes, warmup_spp, warmup_steps_per_batch;
} ajg_guide_args;

typedef struct ajg_guide_stats {
  uint32_t train_steps;   /* Adam steps taken (warm-up included) */
  uint32_t warmup_steps;  /* of which during the warm-up passes */
  uint64_t records;       /* training records produced (warm-up included) */
  float ms_infer;         /* AJG_RENDER_TIMING: hit points + MLP inference */
  float ms_train;         /* AJG_RENDER_TIMING: training steps */
  float last_loss;        /* loss of the last training step (0 if none) */
  float ms_warmup;        /* device time of the warm-up passes */
} ajg_guide_stats;

/* ajg_render_pt with learned path guiding (no reference counterpart; the
 * BSDF-only sampling of renderer/pt.cpp:215-222 and renderer/helper.h:209-246 is
 * replaced by the one-sample mixture of the guide and the BSDF for the
 * continuation and the MIS-BSDF ray at Lambertian vertices, and env-light
 * NEE weights use the mixture pdf there). Disney vertices keep BSDF-only
 * sampling: the reference's Disney Pdf() is not the density of its sampler,
 * so mixing in another strategy would move the converged value away from
 * the plain render's. The expectation equals the plain render's wherever
 * the BSDF's sampled value equals its evaluation (flat-shaded geometry).
 * Synchronous. */
int ajg_render_pt_guided(ajg_scene *scene, const ajg_camera *camera,
                         const ajg_pt_args *args, ajg_guide *guide,
                         const ajg_guide_args *gargs, const ajg_film *film,
                         ajg_render_stats *stats, ajg_guide_stats *gstats,
                         void *stream);

/* ---------------------------------------------------------- batched jobs
 * No counterpart in the reference, which renders one (scene, camera) per
 * Renderer::Render call (renderer.h:64-65, pt.cpp:40-73). A batch runs many
 * jobs of uneven size through one wavefront: every job's work items (pixel
 * x chunk of samples) join one work list, the path slots take items from
 * any job, and every queued ray carries its job so traversal uses that
 * job's BVH. Each job's film ends as ajg_render_pt leaves it: for seed_mode 0
 * and for seed_mode 1 with the same samples per work item the sums are
 * bitwise those of the job rendered alone; per-sample radiance
 * (film.per_sample) is bitwise equal in every case. */
typedef struct ajg_job {
  ajg_scene *scene;     /* all jobs of a batch share one ajg_ctx */
  ajg_camera camera;
  ajg_pt_args args;     /* max_slots and flags are ignored (batch-wide) */
  ajg_film film;
} ajg_job;

typedef struct ajg_batch_args {
  int32_t max_slots;          /* path slots in flight (0: 1 << 20) */
  int32_t samples_per_item;   /* seed_mode 1 jobs: samples per work item
                                 (0: as ajg_render_pt chooses for max_slots) */
  int32_t trav_blocks_per_cu; /* persistent traversal grid (0: 6; at least 16 waves in all) */
  int32_t lds_stack;          /* LDS ring entries per lane: 4, 8 or 16 (0: 8) */
  int32_t iters_per_check;    /* iterations per host check, even (0: 4) */
  int32_t use_graph;          /* 1: replay the iterations from a hipGraph */
  uint32_t flags;             /* AJG_RENDER_COUNT is not supported */
  int32_t pixel_tile;         /* work-list pixel order: 0 scanline, or T x T
                                 tiles (4, 8, 16) where T divides the traced area */
} ajg_batch_args;

typedef struct ajg_batch_stats {
  float ms_total;            /* device time of the batch (events) */
  uint32_t iterations;
  uint32_t slots;
  uint32_t graph_launches;   /* hipGraph replays (0 without use_graph) */
  uint64_t paths;            /* sum over jobs of W*H*spp (traced area) */
  uint64_t closest_queries;
  uint64_t any_queries;
} ajg_batch_stats;

/* Renders n_jobs jobs in one wavefront. job_done_ms (optional, n_jobs
 * floats) receives each job's completion time in ms after the batch start
 * (device wall clock at the last work item's final accumulate). Films are
 * cleared and filled like ajg_render_pt. Synchronous. */
int ajg_render_batch(ajg_ctx *ctx, const ajg_job *jobs, uint32_t n_jobs,
                     const ajg_batch_args *bargs, ajg_batch_stats *stats,
                     float *job_done_ms, void *stream);

/* Autotuner: coordinate descent over max_slots, samples_per_item,
 * pixel_tile, lds_stack, trav_blocks_per_cu and iters_per_check, each
 * candidate timed
 * on the batch with every job's spp cut to trial_spp (the films receive
 * the trial renders). Starts from *inout and leaves the fastest setting
 * there; best_paths_per_s (optional) gets its rate. Synchronous. */
int ajg_batch_autotune(ajg_ctx *ctx, const ajg_job *jobs, uint32_t n_jobs,
                       uint32_t trial_spp, ajg_batch_args *inout,
                       float *best_paths_per_s);

/* ------------------------------------------------------------------ SPPM
 * Replaces CreateSPPMRenderer (renderer/renderer.h:82-97, 113) /
 * SPPMRenderer::Render (renderer/sppm.cpp:184-294) over the GPU scene
 * (one-level scenes). Fields and defaults of SPPMRendererArgs
 * (renderer_creators.cpp:63-71), plus photon_seed_base: photon i of
 * iteration k draws from its own PCG32 stream seeded with
 * mix(photon_seed_base, i, k), where the reference hands all photon threads
 * the renderer's one shared sampler (sppm.cpp:255-260), whose draws are not
 * reproducible. Visible points (TracerVP, sppm.cpp:35-106) use the
 * reference's per-pixel seeding x + y*W, which every iteration repeats, so
 * they are traced once and their direct light is accumulated once per
 * iteration in the reference's order of additions. The hash grid is
 * VisiblePointSearcher's (photon_mapping.h:70-151, 40960 entries, cells
 * [min, max) of each radius box). The film width and height must be
 * multiples of 64 (the reference's 64-pixel tiles are not clipped to the
 * film, sppm.cpp:213-234). */
typedef struct ajg_sppm_args {
  int32_t forward_max_depth;      /* 8 */
  float init_radius;              /* 0.1; < 0: world-box diagonal / 1000 */
  int32_t iteration_count;        /* 100 */
  int32_t photons_per_iteration;  /* 100000 */
  int32_t photon_min_depth;       /* 5 */
  int32_t photon_max_depth;       /* 10 */
  float photon_cont_prob;         /* 0.9 */
  float update_alpha;             /* 2/3 */
  uint32_t photon_seed_base;
  uint32_t flags;                 /* AJG_RENDER_TIMING */
} ajg_sppm_args;

/* Per-pixel state after the last iteration (test and debugging output):
 * the visible point (position, throughput as traced; valid = a visible
 * point was found), the accumulated direct light, tau, radius and N. */
typedef struct ajg_sppm_pixel {
  float pos[3], beta[3];
  float direct[3];
  float tau[3];
  float radius, n;
  uint32_t valid, pad;
} ajg_sppm_pixel;

typedef struct ajg_sppm_stats {
  uint64_t photons;          /* iteration_count x photons_per_iteration */
  uint64_t deposits;         /* photon-visible point contributions */
  uint64_t grid_entries;     /* registrations summed over iterations */
  uint32_t visible_points;   /* pixels with a visible point */
  uint32_t pad;
  float ms_total, ms_visible_points, ms_photons, ms_grid;  /* AJG_RENDER_TIMING */
} ajg_sppm_stats;

/* radiance: device float[W*H*3] receiving ComputePixelRadiance per pixel
 * (the value Film::AddSample then filters, sppm.cpp:282-293; resolve it with
 * ajg_film_resolve_filter, tile size 64, count 1 per pixel). state (device
 * ajg_sppm_pixel[W*H]) and stats may be NULL. Synchronous. */
int ajg_render_sppm(ajg_scene *scene, const ajg_camera *camera, const ajg_sppm_args *args,
                    float *radiance, ajg_sppm_pixel *state, ajg_sppm_stats *stats,
                    void *stream);

/* ------------------------------------------------------------------- VCM
 * Replaces CreateVCMRenderer (renderer/renderer.h:102-116) /
 * VCMRenderer::Render (renderer/vcm.cpp:135-147) over the GPU scene
 * (one-level scenes). Fields and defaults of VCMRendererArgs
 * (renderer_creators.cpp:111-118) plus seed_base: light sub-path i of
 * iteration k draws from PCG32 seeded with mix(seed_base, i, 2k), the camera
 * sub-path of pixel p from mix(seed_base, p, 2k+1) (the reference draws all
 * paths from one sampler in sequence). One light sub-path per pixel per
 * iteration, as the reference. color_sum (device float[W*H*3]) receives,
 * per pixel, the sum over iterations of its camera sub-path's color (the
 * film tile gets it with weight 1 per iteration: resolve it with
 * ajg_film_resolve_filter, tile size 1, count = iteration); splat (device
 * float[W*H*3]) receives the light-tracing splats (Film::AddSplat), summed
 * over iterations; the reference's image is resolve + splat
 * (Film::WriteImage). Synchronous. */
typedef struct ajg_vcm_args {
  int32_t iteration;        /* 1 */
  float base_radius;        /* 0.003 (x the world box's smallest extent) */
  float radius_alpha;       /* 0.75 */
  int32_t use_vm, use_vc;   /* 1, 1 */
  int32_t min_path_length;  /* 0 */
  int32_t max_path_length;  /* 10 */
  float cont_prob;          /* 0.9 */
  uint32_t seed_base;
  uint32_t flags;           /* AJG_RENDER_TIMING */
} ajg_vcm_args;

typedef struct ajg_vcm_stats {
  uint64_t paths;            /* camera sub-paths (W*H x iteration) */
  uint64_t light_vertices;   /* stored light vertices of the last iteration */
  float ms_total, ms_light, ms_grid, ms_camera;  /* AJG_RENDER_TIMING */
} ajg_vcm_stats;

int ajg_render_vcm(ajg_scene *scene, const ajg_camera *camera, const ajg_vcm_args *args,
                   float *color_sum, float *splat, ajg_vcm_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif
