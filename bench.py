"""Benchmark of the path-tracing hot path on one MI355X.

Workload (one "step"): a full render of the heavy-mesh configuration — a
displaced icosphere of 1,310,720 triangles on a 2-triangle ground quad (PLOC
BVH of 1,310,722 triangles), Disney materials, a 1024x512 FP32 HDR
environment map importance-sampled through its 2-D CDF, 1024x1024 pixels,
64 samples per pixel, at most 16 bounces (Russian roulette after bounce 5,
p = 0.9), MIS path tracing with one closest-hit continuation ray, one
closest-hit MIS-BSDF ray and one shadow ray per bounce. Inputs (scene, BVH,
env tables) are resident in HBM before the timed region.

Prints one JSON line: paths/s over the timed steps (whole job), Mrays/s, PLOC
build Mtris/s, the roofline of the dominant kernel (closest-hit traversal)
and the CPU baseline (the oracle's all-core CPU restatement of the reference
path tracer on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FAT_NODE_BYTES = 64      # traversal node: two child boxes + two child refs
LEAF_TRI_BYTES = 48      # leaf triangle: 3 x float4
RAY_BYTES = 32           # o, d, tmin, tmax
HIT_BYTES = 16           # t, leaf, u, v
OCC_BYTES = 1


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--seed-mode", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    return ap.parse_args()


def cpu_baseline(scene, target_s):
    """Oracle (CPU restatement, oracle/liborc.so) on the host cores over a
    bounded band of image rows through the image centre."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = max(1, min(16, os.cpu_count() or 1))
    o = orc.OracleScene(scene)
    cam = o.camera()
    args = o.args()
    wp, hp = o.traced_dims()
    # render 8-row bands alternating outward from the image centre until the
    # time budget is spent
    order = []
    mid = hp // 2
    for k in range(hp):
        r = mid + (k // 2 if k % 2 == 0 else -(k // 2 + 1))
        if 0 <= r < hp:
            order.append(r)
    rows, queries, dt = 0, 0, 0.0
    for i in range(0, len(order), 8):
        band = sorted(order[i:i + 8])
        for lo in band:
            t = time.time()
            res = o.render(cam=cam, args=args, pix_begin=lo * wp, pix_end=(lo + 1) * wp,
                           threads=threads)
            dt += time.time() - t
            queries += res["queries"]
            rows += 1
        if dt >= target_s:
            break
    n_paths = rows * wp * args.spp
    o.close()
    return {"value": n_paths / dt, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"{rows} image rows x {wp} px x {args.spp} spp around the image centre "
                      f"({n_paths} paths, {dt:.1f} s, {queries / dt / 1e6:.2f} Mrays/s)"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); "
              "this build measures a single MI355X", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import Context, GpuScene, PathTracer, PLOCAggregate

    scene = scenes.heavy_mesh(res=(a.res, a.res), spp=a.spp)
    ctx = Context(0)

    # --- PLOC build throughput from a device-resident soup
    tri = torch.from_numpy(scene.soup()).cuda()
    box = scene.scene_box()
    build_ms = []
    for i in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        agg = PLOCAggregate(ctx, 14).build_soup(tri, box)
        torch.cuda.synchronize()
        build_ms.append((time.perf_counter() - t) * 1e3)
    n_tri = agg.info().tri_count
    build_iters = agg.info().iterations
    del agg, tri

    gs = GpuScene(ctx, scene)
    pt = PathTracer(gs, seed_mode=a.seed_mode, flags=capi.RENDER_TIMING)
    film, _ = pt.render_device()          # allocates the film once
    for _ in range(a.warmup):
        pt.render_device(film=film)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(a.steps):
        _, st = pt.render_device(film=film)
        stats.append(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    w, h = scene.camera.res
    paths = a.steps * w * h * a.spp
    queries = sum(s.closest_queries + s.any_queries for s in stats)
    ms_closest = sum(s.ms_trace_closest for s in stats) / a.steps
    launches_closest = stats[0].launches_closest
    closest_q = stats[0].closest_queries

    roof = None
    if not a.no_count:
        cpt = PathTracer(gs, seed_mode=a.seed_mode, flags=capi.RENDER_COUNT)
        _, cs = cpt.render_device(film=film)
        assert cs.closest_queries == closest_q, "counting pass diverged from the timed pass"
        alg_bytes = (closest_q * (RAY_BYTES + HIT_BYTES) + cs.nodes_fetched * FAT_NODE_BYTES +
                     cs.tris_tested * LEAF_TRI_BYTES)
        achieved = alg_bytes / (ms_closest * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_q_closest (closest-hit traversal)",
                "launches_per_step": launches_closest,
                "avg_launch_ms": round(ms_closest / max(1, launches_closest), 4),
                "bytes_per_step": int(alg_bytes),
                "nodes_per_query": round(cs.nodes_fetched / max(1, closest_q), 2),
                "tris_per_query": round(cs.tris_tested / max(1, closest_q), 2)}

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene, a.cpu_seconds)

    value = paths / elapsed
    out = {
        "metric": "Mrays/s & paths/s on 1 GPU; equal-time relMSE guided vs plain PT",
        "value": round(value, 1),
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "heavy_mesh: PLOC BVH over 1,310,722 triangles (displaced "
                               "icosphere + ground), Disney BSDFs, 1024x512 HDR env NEE, "
                               f"{w}x{h}, {a.spp} spp, max 16 bounces",
                   "resolution": [w, h], "spp": a.spp, "max_bounces": 16, "triangles": n_tri,
                   "seed_mode": a.seed_mode, "slots": stats[0].slots},
        "mrays_per_s": round(queries / elapsed / 1e6, 2),
        "rays_per_path": round(queries / paths, 3),
        "build_mtris_per_s": round(n_tri / (min(build_ms) * 1e-3) / 1e6, 2),
        "build_ms": round(min(build_ms), 3),
        "build_iterations": build_iters,
        "stage_ms_per_step": {"trace_closest": round(ms_closest, 3),
                              "trace_any": round(sum(s.ms_trace_any for s in stats) / a.steps, 3),
                              "shade": round(sum(s.ms_shade for s in stats) / a.steps, 3)},
        "wavefront_iterations": stats[0].iterations,
        "relmse_guided_vs_plain": None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu_paths"] = round(value / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
