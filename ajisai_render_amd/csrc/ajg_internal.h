// ajg_internal.h — host-side handle definitions shared by the translation
// units of libajisai_gpu.so (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/ajisai_gpu.h"
#include "ajg_common.h"

struct ajg_ctx {
  int device = 0;
  std::atomic<int> cancel{0};              // raised by ajg_ctx_set_cancel, polled by renders
  std::string last_error;
  int inject_countdown = 0;
  int inject_status = 0;
  void *batch_cache = nullptr;             // buffers of batched renders
  void (*batch_free)(void *) = nullptr;    // frees batch_cache
};

struct ajg_accel {
  ajg_ctx *ctx = nullptr;
  uint32_t n_tri = 0;
  uint32_t iterations = 0;
  uint32_t depth = 0;
  float scene_box[6] = {0, 0, 0, 0, 0, 0};
  ajg::PlocNode *final_nodes = nullptr;  // 2n-1, root 0 (download / parity)
  uint32_t *perm = nullptr;              // n: leaf slot -> global triangle id
  float4 *leaf_tris = nullptr;           // 3n: a.xyz|id, b.xyz, c.xyz per leaf slot
  ajg::FatNode *fat = nullptr;           // n-1 interior nodes (root at 0)
  ajg::WideNode *wide = nullptr;         // n-1 traversal nodes (fat index)
  int root_ref = 0;                      // >= 0 fat index, < 0 ~leaf slot
  float root_box[6] = {0, 0, 0, 0, 0, 0};
  uint64_t bytes = 0;
  uint32_t *cursor = nullptr;            // batch traversal fetch cursor (device)
};

namespace ajg {

// A render polls this at each host checkpoint (every few wavefront
// iterations, every SPPM/VCM iteration) and stops with AJG_ERR_CANCELLED.
inline bool cancelled(const ajg_ctx *ctx) { return ctx->cancel.load(std::memory_order_relaxed) != 0; }

// Error helpers: record a message on the context and return a status.
int fail(ajg_ctx *ctx, int status, const std::string &msg);
int hip_fail(ajg_ctx *ctx, hipError_t e, const char *what);
// Host fault injection: the status to report from the next entry point.
int injected(ajg_ctx *ctx);
// Device allocation that maps hipErrorOutOfMemory to AJG_ERR_OOM.
int dev_alloc(ajg_ctx *ctx, void **p, size_t bytes, const char *what);

// Builds the accelerator from a device-resident triangle soup (n x 9
// floats). scene_box: host pointer (6 floats) or nullptr.
int build_accel(ajg_ctx *ctx, const float *tri_dev, uint32_t n,
                const float *scene_box, uint32_t search_radius,
                hipStream_t stream, ajg_accel **out);
void destroy_accel(ajg_accel *a);

// Device traversal launches (ajg_trace.hip).
int launch_trace_closest(ajg_accel *a, const ajg_ray *rays, uint32_t n,
                         ajg_hit *hits, hipStream_t s);
int launch_trace_any(ajg_accel *a, const ajg_ray *rays, uint32_t n,
                     uint8_t *occ, hipStream_t s);
int launch_trace_counts(ajg_accel *a, const ajg_ray *rays, uint32_t n,
                        uint32_t *nodes, uint32_t *tris, hipStream_t s);

// Path-guide network (ajg_guide.hip). guide_forward_launch writes the
// decoded mixture (40 values per point: FP32 to mix, FP16 to mixh, each
// optional) for min(n, *n_dev) points.
int guide_forward_launch(ajg_guide *gd, const float *pos, const float *nrm, uint32_t n,
                         const uint32_t *n_dev, uint32_t stride, float *mix, _Float16 *mixh,
                         float *raw, hipStream_t s);
int guide_train_launch(ajg_guide *gd, const float *samples, uint32_t n, float lr, hipStream_t s);
// Records the device address of the allocation-free training step's scratch
// for ctx's device (called once per context, with that device current).
int guide_scratch_setup(ajg_ctx *ctx);
// The guide's training-record pool (grown to at least cap rows of 10 floats;
// growing empties it); ctr[0] = rows held, ctr[2..3] = u64 running total.
int guide_records(ajg_guide *gd, uint32_t cap, float **rec, uint32_t **ctr, uint32_t *cap_out);
int guide_steps(const ajg_guide *gd);
const float *guide_loss_dev(const ajg_guide *gd);
const float *guide_box(const ajg_guide *gd);

// u32 scan on the device (ajg_scan.hip). tmp must hold scan_tmp_words(n).
size_t scan_tmp_words(uint32_t n);
void scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, bool inclusive,
              uint32_t *tmp, hipStream_t s);

}  // namespace ajg

#define AJG_HIP(ctx, call)                                   \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return ajg::hip_fail(ctx, e_, #call); \
  } while (0)
