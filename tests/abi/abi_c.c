/* Compiles the C-ABI header as C99 and calls the entry points that need no
 * GPU: the ABI version, and the error paths of a context that could not be
 * created (no device) or was never given. Built and run by
 * tests/test_capi_cpu.py. */
#include <stdio.h>
#include <string.h>

#include "ajisai_gpu.h"

int main(void) {
  if (ajg_abi_version() != AJG_ABI_VERSION) return 1;
  ajg_ctx *ctx = NULL;
  int st = ajg_ctx_create(0, &ctx);
  if (st != AJG_OK) {
    /* no usable device here: a status, no context, no exception */
    if (ctx != NULL) return 2;
    if (st != AJG_ERR_HIP && st != AJG_ERR_INVALID) return 3;
  } else {
    ajg_ctx_destroy(ctx);
  }
  if (ajg_ctx_inject_failure(NULL, 1, AJG_ERR_OOM) != AJG_ERR_INVALID) return 4;
  if (ajg_ctx_set_cancel(NULL, 1) != AJG_ERR_INVALID) return 5;
  if (strcmp(ajg_last_error(NULL), "null context") != 0) return 6;
  if (ajg_accel_destroy(NULL) != AJG_OK) return 7;
  ajg_camera cam;
  const float pos[3] = {0.f, 0.f, 1.f}, at[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
  if (ajg_camera_setup(&cam, pos, at, up, 40.f, 64, 48, 0.f, 1.f) != AJG_OK) return 8;
  printf("c abi ok\n");
  return 0;
}
