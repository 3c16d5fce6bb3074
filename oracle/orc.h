/*
 * CPU restatement of the Ajisai path-tracing hot path — TEST INFRASTRUCTURE.
 *
 * This library is the correctness oracle. Only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark may load it, and only as a
 * checker or as the timed CPU baseline. The product (the HIP library under
 * ajisai_render_amd/) never links, loads or calls it.
 *
 * Every function restates one reference function; the citation next to it
 * names the reference file:line (paths relative to the reference root).
 * Build flags: -O3, no -march, -ffp-contract=off, no fast-math — the same
 * arithmetic the reference gets from `-O3 -DNDEBUG` on baseline x86-64.
 *
 * How each part is pinned (the reference ships no tests, fixtures or golden
 * vectors and cannot be built here, so no part is pinned by running it):
 *   PCG32                — published PCG32 XSH-RR step; numpy uint64 re-derivation.
 *   Morton / radix sort  — naive bit interleave for all 1024 values; numpy stable argsort.
 *   PLOC tree            — invariants (2N-1 nodes, root 0, each leaf once,
 *                          parent == Join(children) bit-exact, mutual-NN per
 *                          level) and brute-force NN on small inputs.
 *   Slab + MT traversal  — brute-force MT over all triangles (same t/id except
 *                          exact-t ties), hand-computed special cases.
 *   Two-level traversal  — identity instances equal the one-level traversal
 *                          bit for bit; brute force over world-space copies of
 *                          the instances (ids except near-ties, t to 1e-4 rel).
 *   Warps / frames       — closed forms (unit length, orthonormality, pdfs).
 *   Lambert              — closed forms (f = a/pi, furnace integral, pdf).
 *   Disney               — GGX D normalisation, the visible-normal identity
 *                          of Smith G1 with D, the visible-normal sampler's
 *                          histogram against its density, the conductor
 *                          Fresnel (closed forms and PBRT-2 FrCond), Eval in
 *                          closed form at normal incidence and in
 *                          retro-reflection, sample/pdf MC consistency. Values
 *                          off those configurations follow from the pinned
 *                          pieces; the world-wo coat pdf quirk is stated.
 *   Env Distribution2D   — pdf integrates to 1; sample histogram vs pdf;
 *                          1x1 map gives 1/(4 pi).
 *   Film resolve         — closed form of the 2x2 clamped box splat.
 *   TraceMIS             — closed-form scenes (white furnace, directional on
 *                          a Lambert plane with the reference MIS weight).
 */
#ifndef AJG_ORACLE_H
#define AJG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flattened PLOC node: Node (ploc_aggregate.cpp:146-152) with size_t fields
 * narrowed to u32. count = primitive_count (0 interior, 1 leaf);
 * first = first_child_or_primitive. */
typedef struct orc_node {
  float mn[3], mx[3];
  uint32_t count, first;
} orc_node;

/* Ray after the Ray constructor (ray.h:33-44): o already offset, d unit. */
typedef struct orc_ray {
  float o[3], d[3];
  float tmin, tmax;
} orc_ray;

/* Closest hit: t, global flattened triangle id (ploc_aggregate.cpp:257-284),
 * MT barycentrics. id = 0xFFFFFFFF and t = +inf on a miss. */
typedef struct orc_hit {
  float t;
  uint32_t id;
  float u, v;
} orc_hit;

/* MeshView (geometry/geometry.h:33-42). */
typedef struct orc_mesh_view {
  const void *position_buffer;
  size_t position_offset, position_stride, position_size;
  const void *index_buffer;
  size_t index_offset, index_stride, index_size;
} orc_mesh_view;

/* One bottom-level tree of an instanced scene: nodes, leaf order and the
 * mesh's triangles (object space). */
typedef struct orc_blas {
  const orc_node *nodes;
  const uint32_t *perm;
  const float *tri;
} orc_blas;

/* ---- PLOC build (ploc_aggregate.cpp) ---------------------------------- */
uint32_t orc_morton_split(uint32_t x);
uint32_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z);
/* Flatten MeshViews into triangles (a,b,c) x n, and the scene box as the Join
 * of per-view position boxes. Returns triangle count. tri may be NULL. */
uint32_t orc_flatten(const orc_mesh_view *views, uint32_t n_views, float *tri,
                     float *scene_box /* [6] */);
void orc_morton_codes(const float *tri, uint32_t n, const float *scene_box,
                      uint32_t *codes);
void orc_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t n,
                    uint32_t bit_count);
/* Full build: nodes[2n-1], perm[n] (primitive_indices_). Returns the number
 * of Cluster iterations, or -1 on a degenerate input (n == 0 or no merge). */
int orc_ploc_build(const float *tri, uint32_t n, const float *scene_box,
                   uint32_t search_radius, orc_node *nodes, uint32_t *perm);
/* One Cluster call on explicit arrays (for level-by-level checks). */
void orc_ploc_nearest(const orc_node *in, uint32_t begin, uint32_t end,
                      uint32_t radius, uint32_t *neighbor);

/* ---- traversal (ploc_aggregate.cpp:154-242, 329-391) ------------------ */
void orc_trace_closest(const orc_node *nodes, const uint32_t *perm,
                       const float *tri, const orc_ray *rays, uint32_t n,
                       orc_hit *hits);
void orc_trace_any(const orc_node *nodes, const uint32_t *perm,
                   const float *tri, const orc_ray *rays, uint32_t n,
                   uint8_t *occluded);
/* Brute force MT over all triangles in id order, strict '<' on t. */
void orc_brute_closest(const float *tri, uint32_t n_tri, const orc_ray *rays,
                       uint32_t n, orc_hit *hits);
/* Counting traversal: visited nodes and tested triangles per ray (closest). */
void orc_trace_closest_counts(const orc_node *nodes, const uint32_t *perm,
                              const float *tri, const orc_ray *rays,
                              uint32_t n, uint32_t *nodes_visited,
                              uint32_t *tris_tested);

/* ---- sampler (random.h:30-69, random_sampler.cpp:28-45) --------------- */
typedef struct orc_pcg32 {
  uint64_t state;
} orc_pcg32;
void orc_pcg32_seed(orc_pcg32 *r, uint64_t seed);
uint32_t orc_pcg32_next_u32(orc_pcg32 *r);
float orc_pcg32_next_float(orc_pcg32 *r);
void orc_pcg32_floats(uint64_t seed, uint32_t n, float *out);

/* ---- scene for the path tracer ----------------------------------------- */
enum { ORC_GEOM_QUAD = 0, ORC_GEOM_CUBE = 1, ORC_GEOM_MESH = 2 };
enum { ORC_MAT_DIFFUSE = 0, ORC_MAT_DISNEY = 1 };
enum { ORC_LIGHT_DIRECTIONAL = 0, ORC_LIGHT_ENV = 1 };

typedef struct orc_primitive {
  orc_mesh_view view;
  uint32_t kind;       /* ORC_GEOM_* */
  uint32_t two_sided;  /* TwoSided wrapper (two_sided.cpp:44-53) */
  uint32_t material;   /* index into materials */
  uint32_t pad;
  const float *normal_and_vs; /* MESH: per vertex (nx,ny,nz,v) */
  float emission[3];   /* non-black: an AreaLight (geometric.cpp:35-51) */
  uint32_t pad2;
} orc_primitive;

/* Material record: DIFFUSE uses base_color as albedo. DISNEY parameter order
 * follows CreateDisney (material/material.h:40-54). */
typedef struct orc_material {
  uint32_t type;
  float base_color[3];
  float subsurface, metallic, specular, specular_tint, roughness, anisotropic,
      sheen, sheen_tint, clearcoat, clearcoat_gloss;
} orc_material;

typedef struct orc_light {
  uint32_t type;
  float dir[3];       /* DIRECTIONAL: toward-light direction (un-normalised) */
  float angle;        /* DIRECTIONAL: cone half-angle, radians */
  float radiance[3];  /* DIRECTIONAL emission */
  /* ENV: texture (w x h RGB float, row-major, y down) + rot (deg) + scaler */
  uint32_t env_w, env_h;
  const float *env_texels;
  float rot, scaler;
} orc_light;

typedef struct orc_camera {
  float raster2camera[16]; /* column-major (Matrix4 columns) */
  float camera2world[16];
  float resolution[2];
  float lens_radius, focal_distance;
} orc_camera;

typedef struct orc_pt_args {
  int32_t spp, tile_size, min_bounces, max_bounces;
  float cont_prob;
  int32_t use_mis;
  /* 0: reference seeding (seed x + y*W per pixel, spp sequential in one
   *    stream, pt.cpp:56-64). 1: per-sample streams (seed = mix(pixel,
   *    sample, seed_base)). */
  int32_t seed_mode;
  uint32_t seed_base;
  /* env Distribution2D prefix sums: 0 = float sequential like
   * distribution.h:44-48; 1 = double sequential then rounded to float. */
  int32_t env_cdf_fp64;
  int32_t pad;
} orc_pt_args;

typedef struct orc_scene orc_scene;

/* Build a scene: flattens the primitives (PLOC order), builds the PLOC BVH
 * with search radius r, builds env distributions, processes lights against
 * the world bound (default_scene.cpp:31-61). Returns NULL on error. */
orc_scene *orc_scene_create(const orc_primitive *prims, uint32_t n_prims,
                            const orc_material *mats, uint32_t n_mats,
                            const orc_light *lights, uint32_t n_lights,
                            uint32_t search_radius, int32_t env_cdf_fp64);
void orc_scene_destroy(orc_scene *s);
uint32_t orc_scene_tri_count(const orc_scene *s);
const orc_node *orc_scene_nodes(const orc_scene *s);
const uint32_t *orc_scene_perm(const orc_scene *s);
const float *orc_scene_tris(const orc_scene *s);
void orc_scene_world_bound(const orc_scene *s, float *box6);

/* Env Distribution2D tables of light `li` (for parity of the GPU tables):
 * cond_cdf[h][w+1], cond_func_int[h], marg_cdf[h+1], marg_func_int. */
int orc_scene_env_tables(const orc_scene *s, uint32_t li, float *cond_cdf,
                         float *cond_func_int, float *marg_cdf,
                         float *marg_func_int);
/* EnvLight::Sample / Pdf / Radiance (env_light.cpp:50-170) on given numbers. */
void orc_env_sample(const orc_scene *s, uint32_t li, float u0, float u1,
                    float *dir3, float *pdf, float *rgb);
float orc_env_pdf(const orc_scene *s, uint32_t li, const float *dir3);
void orc_env_radiance(const orc_scene *s, uint32_t li, const float *dir3,
                      float *rgb);

/* BSDF entry points on an explicit shading point (ng, ns, wo). */
void orc_bsdf_sample_batch(const orc_material *m, const float *ng, const float *ns,
                           const float *wo, uint32_t n, const float *sam3, float *wi,
                           int *valid);
void orc_bsdf_sample(const orc_material *m, const float *ng, const float *ns,
                     const float *wo, const float *sam3, float *wi,
                     float *f, float *pdf, int *valid);
void orc_bsdf_eval(const orc_material *m, const float *ng, const float *ns,
                   const float *wi, const float *wo, float *f);
float orc_bsdf_pdf(const orc_material *m, const float *ng, const float *ns,
                   const float *wi, const float *wo);

/* Warps (warp.h) and frame (coordinate_system.h). */
void orc_square_to_disk(const float *s2, float *out2);
void orc_cosine_hemisphere(const float *s2, float *out3, float *pdf);
void orc_frame(const float *n, float *dx, float *dy);

/* Camera ray for raster (x, y) and lens sample (thin_lens.cpp:62-83),
 * returned after the Ray constructor. */
void orc_camera_ray(const orc_camera *cam, float rx, float ry,
                    const float *lens2, orc_ray *out);
/* Camera matrices from look-at parameters (thin_lens.cpp:32-55). */
void orc_camera_setup(orc_camera *cam, const float *pos, const float *look_at,
                      const float *up, float fov_deg, float res_x, float res_y,
                      float lens_radius, float focal_distance);

/* Render. Traced pixels cover whole tiles (pt.cpp:43-49):
 * Wp = ceil(W/ts)*ts, Hp = ceil(H/ts)*ts. Outputs per traced pixel (Wp*Hp):
 * sum[3] (double), sum_sq[3] (double, per channel L^2), count.
 * If per_sample != NULL it receives L (3 floats) for sample s of traced
 * pixel p at index (p * spp + s). Pixels with index outside
 * [pix_begin, pix_end) are skipped (for bounded baseline samples).
 * Returns the number of closest-hit + any-hit queries issued. */
uint64_t orc_render(const orc_scene *s, const orc_camera *cam,
                    const orc_pt_args *args, uint32_t pix_begin,
                    uint32_t pix_end, double *sum, double *sum_sq,
                    uint32_t *count, float *per_sample);
/* Same on n_threads std threads over rows of traced pixels (CPU baseline:
 * mirrors the tile pool of parallel.cpp:42-130). */
uint64_t orc_render_mt(const orc_scene *s, const orc_camera *cam,
                       const orc_pt_args *args, uint32_t pix_begin,
                       uint32_t pix_end, double *sum, double *sum_sq,
                       uint32_t *count, int n_threads);
/* Film resolve (film.cpp:27-63, film.h:119-129, image.h:38-48) from
 * per-traced-pixel sums to the W x H image (RGB float). */
/* Filtered film resolve: type 0 box, 1 gaussian (radius rx, ry; alpha). */
void orc_film_resolve_filter(uint32_t w, uint32_t h, uint32_t tile_size, int type, float rx,
                             float ry, float alpha, const double *sum, const uint32_t *count,
                             float *rgb);
/* Linear RGB (w*h*3) -> 8-bit sRGB as the image writer stores it. */
void orc_image_srgb8(uint32_t w, uint32_t h, const float *rgb, int flip, uint8_t *out);
void orc_film_resolve(uint32_t w, uint32_t h, uint32_t tile_size,
                      const double *sum, const uint32_t *count, float *rgb);

/* Disney microfacet pieces (test hooks): per i, D(h_i), Smith G1(v_i, h_i)
 * and the visible-normal pdf of h_i given v_i; the visible-normal sampler
 * (u: 2 numbers per sample); the conductor Fresnel. Local shading frame. */
void orc_ggx_parts(uint32_t n, const float *v, const float *h, float alpha, float *d, float *g1,
                   float *pdf_vn);
void orc_ggx_sample_vn_batch(uint32_t n, const float *v, const float *u2, float alpha, float *h,
                             float *pdf);
float orc_fresnel_conductor(float cosi, float eta, float k);

/* Two-level closest / any hit (instancing, no reference counterpart): the
 * top-level tree's leaf primitives are instances; w2o holds 12 floats
 * (row-major 3x4) per instance, inst_mesh its bottom-level tree. hits[i].id
 * is the triangle index within the mesh, inst_out[i] the instance
 * (0xFFFFFFFF on a miss). */
void orc_trace_closest_2l(const orc_node *tlas, const uint32_t *tlas_perm, const float *w2o,
                          const uint32_t *inst_mesh, const orc_blas *blas, const orc_ray *rays,
                          uint32_t n, orc_hit *hits, uint32_t *inst_out);
void orc_trace_any_2l(const orc_node *tlas, const uint32_t *tlas_perm, const float *w2o,
                      const uint32_t *inst_mesh, const orc_blas *blas, const orc_ray *rays,
                      uint32_t n, uint8_t *occluded);

/* SPPM (SPPMRenderer, sppm.cpp:35-303; VisiblePointSearcher,
 * photon_mapping.h:29-153). Fields and defaults of SPPMRendererArgs
 * (renderer.h:82-97, renderer_creators.cpp:63-71) plus the photon streams'
 * seed: photon i of iteration k draws from PCG32 seeded with
 * mix(photon_seed_base, i, k) (the reference hands every photon thread the
 * one shared sampler, whose draws are not reproducible). */
typedef struct orc_sppm_args {
  int32_t forward_max_depth;
  float init_radius;
  int32_t iteration_count;
  int32_t photons_per_iteration;
  int32_t photon_min_depth;
  int32_t photon_max_depth;
  float photon_cont_prob;
  float update_alpha;
  uint32_t photon_seed_base;
  uint32_t pad;
} orc_sppm_args;

/* Per-pixel SPPM state after the last iteration run. */
typedef struct orc_sppm_pixel {
  float pos[3], beta[3];   /* visible point (beta: throughput) */
  float direct[3];         /* accumulated direct illumination */
  float tau[3];
  float radius, n;
  uint32_t valid, pad;
} orc_sppm_pixel;

/* Runs `iterations` of args->iteration_count iterations (all if 0) and
 * writes the radiance ComputePixelRadiance gives after them (w*h*3, the
 * per-pixel value the film then filters) and optionally the per-pixel state.
 * Returns the number of photon deposits, or -1 if the film size is not a
 * multiple of 64. */
int64_t orc_render_sppm(const orc_scene *s, const orc_camera *cam, const orc_sppm_args *a,
                        int32_t iterations, float *radiance, orc_sppm_pixel *state);
/* The same on n_threads threads (visible points and photons split across
 * them; with more than one, photon flux sums depend on arrival order). */
int64_t orc_render_sppm_mt(const orc_scene *s, const orc_camera *cam, const orc_sppm_args *a,
                           int32_t iterations, float *radiance, orc_sppm_pixel *state,
                           int n_threads);
/* Test hook: n draws of Light::SampleEmit (area_light.cpp:60-73,
 * directional_light.cpp:60-79, env_light.cpp:105-147) from light li (the
 * emissive primitives' lights first, then the scene's), draw i from its own
 * seeded stream; pos/dir/normal/radiance are n x 3. */
void orc_light_emit_batch(const orc_scene *s, uint32_t li, uint32_t seed, uint32_t n,
                          float *pos, float *dir, float *normal, float *radiance,
                          float *pdf_pos, float *pdf_dir);
/* Test hook: UpdatePixel (sppm.cpp:145-167) on one pixel over `iters`
 * iterations of m[k] photons with flux phi (iters x 3) at throughput beta;
 * out (iters x 5) holds radius, N, tau after each. */
void orc_sppm_update_probe(float alpha, float r0, const float *beta, uint32_t iters,
                           const uint32_t *m, const float *phi, float *out);
/* Test hook: the visible-point grid alone (VisiblePointSearcher AddVp and
 * AddPhoton's lookup and distance test, no BSDF): per visible point, the
 * number of the photon positions matched to it. */
void orc_sppm_grid_probe(const float *bmin, float side, uint32_t n_vp, const float *vp_pos,
                         const float *vp_r, uint32_t n_ph, const float *ph_pos, uint32_t *counts);

/* VCM (VCMRenderer, vcm.cpp:130-751). Fields and defaults of VCMRendererArgs
 * (renderer.h:102-114, renderer_creators.cpp:111-118) plus seed_base: light
 * path i of iteration k draws from PCG32 seeded with mix(seed_base, i, 2k),
 * the camera path of pixel p from mix(seed_base, p, 2k+1). */
typedef struct orc_vcm_args {
  int32_t iteration;
  float base_radius, radius_alpha;
  int32_t use_vm, use_vc;
  int32_t min_path_length, max_path_length;
  float cont_prob;
  uint32_t seed_base;
  uint32_t variant; /* 0: the reference as written; else ORC_VCM_* bits (test-only) */
} orc_vcm_args;

/* Test-only estimator variants that each replace one formula of the
 * reference's VCM by its consistent form (Georgiev et al. 2012, "Light
 * transport simulation with vertex connection and merging", and its
 * published SmallVCM code), so that tests can show the rest converges:
 *   PIXEL_UNITS  the camera's solid-angle pdf and ConnectToCamera's
 *                image-to-surface factor per unit pixel area instead of
 *                per unit image-plane area at the focal distance, and the
 *                camera's dVCM = light path count / camera pdf
 *                (vcm.cpp:556-600, 713-729 use image-plane units and
 *                dVCM = 1 / camera pdf);
 *   REVERSE_PDF  the reverse pdf of a scattering event is the BSDF's pdf of
 *                the incoming direction given the new one (vcm.cpp:416-423
 *                evaluates PdfAll(-new_dir, new_dir));
 *   LT_ONLY      light tracing alone: every light vertex connects to the
 *                camera with weight 1, nothing is stored, and camera paths
 *                only collect what their first ray hits or escapes to. */
enum { ORC_VCM_PIXEL_UNITS = 1, ORC_VCM_REVERSE_PDF = 2, ORC_VCM_LT_ONLY = 4 };

/* color_sum (w*h*3): per pixel, the sum over iterations of its camera path's
 * color (the value the film's tile receives with weight 1 per iteration);
 * splat (w*h*3): the light-tracing splats summed over iterations. Returns
 * the number of light vertices of the last iteration. */
int64_t orc_render_vcm(const orc_scene *s, const orc_camera *cam, const orc_vcm_args *a,
                       float *color_sum, float *splat);
/* The same on n_threads threads (sub-paths split across them; with more
 * than one, splat sums depend on arrival order). */
int64_t orc_render_vcm_mt(const orc_scene *s, const orc_camera *cam, const orc_vcm_args *a,
                          float *color_sum, float *splat, int n_threads);
/* Test hooks: ThinLensCamera::ToRaster (returns 0 when the point does not
 * project into the film) and ImageAreaToSolidAngle. */
int orc_vcm_to_raster(const orc_camera *cam, const float *u2, const float *pos, float *dir_out,
                      float *raster_out);
float orc_vcm_image_to_solid(const orc_camera *cam, const float *dir);

#ifdef __cplusplus
}
#endif
#endif
