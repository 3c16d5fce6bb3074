"""Writes the 65,536-spp plain path-traced reference of the occluded room
(scenes.occluded_room(), 512 x 512, max 16 bounces) used by the guided
equal-time comparisons: tests/golden/occluded_room_512_plain_64k.npz.

Rendered on the GPU by this package's plain path tracer (per-sample
streams, seed_base 7001 + chunk, AJG_RENDER_FINITE_PDF so that a zero-pdf
continuation ends the path instead of carrying an infinite throughput),
in 64 chunks of 1024 spp. The plain renderer's per-sample radiance is pinned
against the oracle by tests/test_gpu_render.py; this file only accumulates
it. Stored: per-pixel mean radiance and the variance of that mean, both as
float16 (relative rounding 2^-11, far below the image's own noise).

Run: python tests/golden/make_occluded_room_reference.py
"""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(ROOT, "tests", "golden", "occluded_room_512_plain_64k.npz")
SPP, CHUNK, SEED0 = 65536, 1024, 7001


def main():
    import numpy as np
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import Context, GpuScene, PathTracer

    ctx = Context(0)
    sc = scenes.occluded_room()
    g = GpuScene(ctx, sc)
    s = s2 = cnt = None
    t0 = time.time()
    for k in range(SPP // CHUNK):
        o = PathTracer(g, spp=CHUNK, seed_mode=1, seed_base=SEED0 + k,
                       flags=capi.RENDER_FINITE_PDF).Render()
        a, b, c = (o["sum"].astype(np.float64), o["sum_sq"].astype(np.float64),
                   o["count"].astype(np.float64))
        s, s2, cnt = (a, b, c) if s is None else (s + a, s2 + b, cnt + c)
        print(f"chunk {k + 1}/{SPP // CHUNK}: {time.time() - t0:.1f} s", flush=True)
    torch.cuda.synchronize()
    w, h = sc.camera.res
    mean = s / cnt[:, None]
    var_mean = np.maximum(s2 / cnt[:, None] - mean * mean, 0.0) / cnt[:, None]
    np.savez_compressed(OUT, mean=mean.astype(np.float16).reshape(h, w, 3),
                        var_mean=var_mean.astype(np.float16).reshape(h, w, 3),
                        spp=np.int64(SPP), seed_base=np.int64(SEED0), chunk=np.int64(CHUNK),
                        seconds=np.float64(time.time() - t0))
    print(f"wrote {OUT}: mean {mean.mean():.6f}, {time.time() - t0:.1f} s", flush=True)
    g.close()


if __name__ == "__main__":
    main()
