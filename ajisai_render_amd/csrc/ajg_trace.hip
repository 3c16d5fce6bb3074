// ajg_trace.hip — batch closest/any-hit kernels behind
// ajg_accel_trace_closest / ajg_accel_trace_any / ajg_accel_trace_counts.
// Closest-hit batches run the persistent lane-refill traversal of
// ajg_traverse.h; any-hit and counting kernels run one ray per thread.
#include "ajg_internal.h"
#include "ajg_traverse.h"

namespace ajg {

static DevBVH dev_bvh(const ajg_accel *a) {
  DevBVH b;
  b.wide = a->wide;
  b.leaf_tris = a->leaf_tris;
  b.root_ref = a->root_ref;
  for (int i = 0; i < 3; ++i) {
    b.root_mn[i] = a->root_box[i];
    b.root_mx[i] = a->root_box[3 + i];
  }
  return b;
}

#define AJG_LDS_STACK                            \
  __shared__ int s_node[kLdsStack * kTravThreads]; \
  __shared__ float s_entry[kLdsStack * kTravThreads]

__device__ __forceinline__ TRay batch_ray(const ajg_ray *rays, uint32_t i) {
  const ajg_ray ray = rays[i];
  return make_tray(ld3(ray.o), ld3(ray.d), ray.tmin, ray.tmax);
}

// Persistent lane-refill kernels over a batch; cursor cleared by k_cursor_clear.
template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_trace_closest(DevBVH bvh, const ajg_ray *rays, uint32_t n, ajg_hit *hits,
                const uint32_t *perm, uint32_t *cursor) {
  AJG_LDS_STACK;
  uint32_t nn = 0, nt = 0;
  persistent_closest<SPILL, false>(
      bvh, n, cursor, s_node + threadIdx.x, s_entry + threadIdx.x,
      [&](uint32_t i, DevBVH &) { return batch_ray(rays, i); },
      [&](uint32_t i, const THit &h) {
        ajg_hit o;
        o.t = h.t;
        o.tri_id = h.slot >= 0 ? perm[h.slot] : kNone;
        o.u = h.u;
        o.v = h.v;
        hits[i] = o;
      },
      &nn, &nt);
}

template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_trace_counts(DevBVH bvh, const ajg_ray *rays, uint32_t n, uint32_t *nodes,
               uint32_t *tris) {
  AJG_LDS_STACK;
  const uint32_t i = blockIdx.x * kTravThreads + threadIdx.x;
  if (i >= n) return;
  const TRay r = batch_ray(rays, i);
  uint32_t nn = 0, nt = 0;
  (void)trace_closest<SPILL, true>(bvh, r, s_node + threadIdx.x, s_entry + threadIdx.x, &nn, &nt);
  nodes[i] = nn;
  tris[i] = nt;
}

template <int SPILL>
__global__ void __launch_bounds__(kTravThreads)
k_trace_any(DevBVH bvh, const ajg_ray *rays, uint32_t n, uint8_t *occ) {
  AJG_LDS_STACK;
  const uint32_t i = blockIdx.x * kTravThreads + threadIdx.x;
  if (i >= n) return;
  const TRay r = batch_ray(rays, i);
  occ[i] = trace_any<SPILL>(bvh, r, s_node + threadIdx.x, s_entry + threadIdx.x) ? 1 : 0;
}

// A fill kernel rather than a memset, so that callers may capture the
// trace calls in a HIP graph.
__global__ void k_cursor_clear(uint32_t *cursor) {
  if (threadIdx.x == 0) *cursor = 0;
}

static uint32_t persistent_grid(const ajg_accel *a, uint32_t n) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = v;
  }
  (void)a;
  const uint32_t full = (uint32_t)cus * 8;            // 8 blocks of 4 waves per CU
  const uint32_t need = (n + kTravThreads - 1) / kTravThreads;
  return need < full ? (need ? need : 1) : full;
}

static inline uint32_t grid(uint32_t n) { return (n + 255) / 256; }

int launch_trace_closest(ajg_accel *a, const ajg_ray *rays, uint32_t n, ajg_hit *hits,
                         hipStream_t s) {
  if (n == 0) return AJG_OK;
  const DevBVH b = dev_bvh(a);
  const uint32_t g = persistent_grid(a, n);
  hipLaunchKernelGGL(k_cursor_clear, dim3(1), dim3(64), 0, s, a->cursor);
  if (spill_for_depth(a->depth) == kSpillS)
    hipLaunchKernelGGL(k_trace_closest<kSpillS>, dim3(g), dim3(256), 0, s, b, rays, n, hits,
                       (const uint32_t *)a->perm, a->cursor);
  else if (spill_for_depth(a->depth) == kSpillL)
    hipLaunchKernelGGL(k_trace_closest<kSpillL>, dim3(g), dim3(256), 0, s, b, rays, n, hits,
                       (const uint32_t *)a->perm, a->cursor);
  else
    return fail(a->ctx, AJG_ERR_INTERNAL, "trace: tree too deep for the traversal stack");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(a->ctx, e, "k_trace_closest");
}

int launch_trace_counts(ajg_accel *a, const ajg_ray *rays, uint32_t n, uint32_t *nodes,
                        uint32_t *tris, hipStream_t s) {
  if (n == 0) return AJG_OK;
  const DevBVH b = dev_bvh(a);
  if (spill_for_depth(a->depth) == kSpillS)
    hipLaunchKernelGGL(k_trace_counts<kSpillS>, dim3(grid(n)), dim3(256), 0, s, b, rays, n, nodes, tris);
  else if (spill_for_depth(a->depth) == kSpillL)
    hipLaunchKernelGGL(k_trace_counts<kSpillL>, dim3(grid(n)), dim3(256), 0, s, b, rays, n, nodes, tris);
  else
    return fail(a->ctx, AJG_ERR_INTERNAL, "trace: tree too deep for the traversal stack");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(a->ctx, e, "k_trace_counts");
}

int launch_trace_any(ajg_accel *a, const ajg_ray *rays, uint32_t n, uint8_t *occ,
                     hipStream_t s) {
  if (n == 0) return AJG_OK;
  const DevBVH b = dev_bvh(a);
  if (spill_for_depth(a->depth) == kSpillS)
    hipLaunchKernelGGL(k_trace_any<kSpillS>, dim3(grid(n)), dim3(256), 0, s, b, rays, n, occ);
  else if (spill_for_depth(a->depth) == kSpillL)
    hipLaunchKernelGGL(k_trace_any<kSpillL>, dim3(grid(n)), dim3(256), 0, s, b, rays, n, occ);
  else
    return fail(a->ctx, AJG_ERR_INTERNAL, "trace: tree too deep for the traversal stack");
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AJG_OK : hip_fail(a->ctx, e, "k_trace_any");
}

}  // namespace ajg
