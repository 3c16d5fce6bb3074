"""Benchmark of the path-tracing hot path on one MI355X.

Workload (one "step"): a full render of the heavy-mesh configuration — a
displaced icosphere of 1,310,720 triangles on a 2-triangle ground quad (PLOC
BVH of 1,310,722 triangles), Disney materials, a 1024x512 FP32 HDR
environment map importance-sampled through its 2-D CDF, 1024x1024 pixels,
64 samples per pixel, at most 16 bounces (Russian roulette after bounce 5,
p = 0.9), MIS path tracing with one closest-hit continuation ray, one
closest-hit MIS-BSDF ray and one shadow ray per bounce. Inputs (scene, BVH,
env tables) are resident in HBM before the timed region.

Random numbers: by default one counter-based PCG32 stream per (pixel,
sample) (seed_mode 1), so the 64 samples of a pixel run in parallel and 16M
paths are in flight; the reference's per-pixel sequential stream (seed_mode
0: a pixel's samples one after another, 1M paths in flight at 1024x1024) is
timed as well (paths_per_s_seed_mode0). Both trace the same paths per pixel
with bit-exact hits; only which random numbers a sample draws differs.

Prints one JSON line: paths/s over the timed steps (whole job), Mrays/s, PLOC
build Mtris/s, the roofline of the dominant kernel (the stage with the most
device time: shading, closest-hit or shadow-ray traversal) with every
stage's roofline beside it, and the CPU baseline (the oracle's all-core CPU
restatement of the reference path tracer on a bounded sample of the same
workload).

The "guided" object is the equal-time comparison of learned path guiding
against plain path tracing on the occluded room (4 x 3 x 4 room lit only
through a 0.2 x 1.0 slit by the same HDR map, 204,866 triangles, 512 x 512,
max 16 bounces): each method renders for the same wall-clock budget (the
guided one's warm-up training included), and relMSE = mean over pixels and
channels of (I - R)^2 / (R^2 + 1e-2) against the committed 65,536-spp
plain path-tracing reference R (tests/golden; --guided-ref guided or plain
renders one instead).
"mlp" times the guide network's kernels alone (MFMA utilisation).

The "batched" object is the batched uneven workload: 64 independent scenes
(displaced-icosphere sets, triangle totals log-uniform in [1e3, 2e6], 1-8
Disney materials, resolution among 128^2 ... 1024^2, 4/16/64 spp, max 12
bounces, the same HDR env map) rendered as one wavefront whose iterations
are replayed from a captured hipGraph, with the autotuner's setting; it
reports aggregate paths/s, each job's completion time, and the same jobs
rendered one after another for comparison.

The "instanced" object is the large instanced scene: 8 displaced meshes
(81,920 or 327,680 triangles each) placed by 2000 instances (random
rotations, scale log-uniform in [0.2, 5]) on a 200 x 200 ground, two-level
BVH (PLOC per mesh + PLOC over instance boxes), Disney materials, a
2048 x 1024 HDR env map (sigma 2.5) plus a directional light, 1920 x 1080,
256 spp, max 24 bounces: paths/s and rays/s of one full render.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 64          # traversal node fetch: grid origin and steps, four quantised slot boxes, four refs
LEAF_TRI_BYTES = 48      # leaf triangle: 3 x float4
RAY_BYTES = 32           # o, d, tmin, tmax
HIT_BYTES = 16           # t, leaf, u, v
OCC_BYTES = 1


def shade_vertex_bytes(n_lights):
    """Algorithmic bytes of the shade stage per shaded vertex (a closest-hit
    result turned into the next rays), one environment or directional light
    per entry of n_lights, no area lights: the records the shade stage
    (k_shade_pend, then k_shade) reads and writes for a live slot.
    Reads: the previous bounce's pending record (flags 4, per light shadow
    index 4 + occlusion byte 1 + contribution 12, MIS index 4, MIS ray 32,
    f 12, pdf 4, shading normal 12, throughput 12, radiance 12, work / sample
    / bank 12), the extension ray 32 and hit 16, slot state (index 4,
    bounce 4, radiance 12, throughput 12, work / sample / end 12, rng 8),
    the triangle's shade record with its three vertex normals 64 and its
    vertices 48, the material 48.
    Writes: the new pending record (radiance 12, throughput 12, work /
    sample / bank 12, flags 4, per light contribution 12 + shadow index 4 +
    shadow ray 32, MIS ray 32 + f / pdf / normal 28 + index 4), the
    continuation ray 32, slot index 4, bounce 4, throughput 12, rng 8."""
    reads = 4 + 17 * n_lights + 4 + 32 + 12 + 4 + 12 + 12 + 12 + 12 + \
        32 + 16 + 4 + 4 + 12 + 12 + 12 + 8 + 64 + 48 + 48
    writes = 12 + 12 + 12 + 4 + 48 * n_lights + 32 + 28 + 4 + 32 + 4 + 4 + 12 + 8
    return reads + writes


def stage_roofline(name, alg_bytes, ms, launches, extra=None):
    achieved = alg_bytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "ms_per_step": round(ms, 3), "launches_per_step": launches,
           "avg_launch_ms": round(ms / max(1, launches), 4), "bytes_per_step": int(alg_bytes)}
    if extra:
        out.update(extra)
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--seed-mode", type=int, default=1,
                    help="1: one PCG32 stream per (pixel, sample), samples of a pixel in parallel; "
                         "0: the reference's per-pixel sequential stream (one path per pixel in flight)")
    ap.add_argument("--slots", type=int, default=1 << 24, help="paths in flight (seed_mode 1)")
    ap.add_argument("--no-seed0", action="store_true",
                    help="skip the per-pixel-stream (seed_mode 0) comparison render")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cornell", dest="cornell", action="store_false",
                    help="skip the Cornell-box (256x256, 16 spp) section")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--guided-budget", type=float, default=10.0,
                    help="equal-time budget (s) of the guided vs plain comparison; 0 skips it")
    ap.add_argument("--guided-res", type=int, default=512)
    ap.add_argument("--guided-ref", choices=["committed", "guided", "plain"], default="committed",
                    help="committed: the 65,536-spp plain reference in tests/golden (512 x 512 only)")
    ap.add_argument("--guided-ref-spp", type=int, default=4096)
    ap.add_argument("--batch-jobs", type=int, default=64,
                    help="jobs of the batched uneven workload (0: skip)")
    ap.add_argument("--batch-trial-spp", type=int, default=4)
    ap.add_argument("--batch-fixed", type=str, default="",
                    help="skip the autotuner and the one-by-one renders; use this setting "
                         "(e.g. 'max_slots=2097152,pixel_tile=8')")
    ap.add_argument("--inst-spp", type=int, default=256,
                    help="spp of the instanced-scene render (0: skip)")
    ap.add_argument("--inst-guided-budget", type=float, default=3.0,
                    help="equal-time budget (s) of guided vs plain on the instanced scene (0: skip)")
    ap.add_argument("--inst-guided-res", type=int, default=480)
    ap.add_argument("--inst-guided-runs", type=int, default=5,
                    help="independent images per method in the instanced guided comparisons "
                         "(medians compared)")
    ap.add_argument("--inst-guided-ref-spp", type=int, default=4096,
                    help="spp of the plain-PT reference of the instanced guided comparison")
    ap.add_argument("--area-spp", type=int, default=64,
                    help="spp of the area-lit Cornell box render (0: skip)")
    ap.add_argument("--area-cpu-seconds", type=float, default=4.0)
    ap.add_argument("--sppm-iters", type=int, default=100,
                    help="SPPM iterations (x 100k photons) on the area-lit Cornell box (0: skip)")
    ap.add_argument("--no-sppm-cpu", action="store_true")
    ap.add_argument("--vcm-iters", type=int, default=8,
                    help="VCM iterations on the area-lit Cornell box at 1024x1024 (0: skip)")
    ap.add_argument("--no-vcm-cpu", action="store_true")
    ap.add_argument("--only", choices=["batched", "instanced", "instanced_guided", "mlp", "guided",
                                       "area", "sppm", "vcm", "cornell"],
                    help="run one section alone and print its JSON (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def calibrate_copy():
    import torch
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "copies": 3}))


BF16_PEAK_TFLOPS = 2500.0
MLP_FLOPS_PER_POINT = 2 * (112 * 64 + 64 * 64 * 2 + 64 * 48)  # forward multiply-adds x 2


def mlp_bench(ctx):
    """Guide network kernels alone: inference over 2^20 points and one
    training step (forward + loss + backward + Adam) on 2^18 records."""
    import numpy as np
    import torch
    from ajisai_render_amd.core import PathGuide
    rng = np.random.default_rng(1)
    box = np.array([0, 0, 0, 4, 3, 4], np.float32)
    g = PathGuide(ctx, box, seed=3)
    n = 1 << 20
    pos = torch.from_numpy(rng.uniform(0, 3, (n, 3)).astype(np.float32)).cuda()
    nrm = torch.nn.functional.normalize(torch.randn(n, 3), dim=1).cuda()
    nb = 1 << 18
    smp = torch.cat([pos[:nb], nrm[:nb], torch.nn.functional.normalize(torch.randn(nb, 3), dim=1).cuda(),
                     torch.rand(nb, 1).cuda()], dim=1).contiguous()
    for _ in range(3):
        g.forward(pos, nrm, want_raw=False)
        g.train_step(smp, lr=1e-3)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        g.forward(pos, nrm, want_raw=False)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t) / 20
    t = time.perf_counter()
    for _ in range(20):
        g.train_step(smp, lr=1e-3)
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t) / 20
    g.close()
    fwd_tf = MLP_FLOPS_PER_POINT * n / t_fwd / 1e12
    train_tf = 3 * MLP_FLOPS_PER_POINT * nb / t_train / 1e12
    return {"inference_points": n, "inference_ms": round(t_fwd * 1e3, 3),
            "inference_tflops": round(fwd_tf, 1),
            "train_batch": nb, "train_step_ms": round(t_train * 1e3, 3),
            "train_tflops": round(train_tf, 1), "peak_bf16_tflops": BF16_PEAK_TFLOPS,
            "inference_frac": round(fwd_tf / BF16_PEAK_TFLOPS, 4),
            "train_frac": round(train_tf / BF16_PEAK_TFLOPS, 4)}


def _img(out):
    import numpy as np
    n = out["count"].astype(np.float64)
    return out["sum"].astype(np.float64) / np.maximum(n, 1)[:, None]


RELMSE_EPS = 1e-2


def relmse(img, ref):
    """relMSE = mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2),
    over the pixels finite in both images (the reference's estimator can
    overflow to inf/NaN on rare heavy-tailed Disney paths; such pixels are
    counted separately by nonfinite())."""
    import numpy as np
    ok = np.isfinite(img).all(axis=1) & np.isfinite(ref).all(axis=1)
    return float(np.mean((img[ok] - ref[ok]) ** 2 / (ref[ok] ** 2 + RELMSE_EPS)))


def nonfinite(img):
    import numpy as np
    return int((~np.isfinite(img).all(axis=1)).sum())


ROOM_REFERENCE = os.path.join(ROOT, "tests", "golden", "occluded_room_512_plain_64k.npz")


def guided_equal_time(ctx, a):
    """Equal-time relMSE of guided vs plain path tracing on the occluded room,
    against the committed 65,536-spp plain-PT reference by default."""
    from ajisai_render_amd import scenes
    sc = scenes.occluded_room(res=(a.guided_res, a.guided_res), spp=16)
    label = (f"occluded_room: 4x3x4 room, 0.2x1.0 slit, HDR env 1024x512 through the slit, "
             f"{sc.n_tri} triangles, {a.guided_res}x{a.guided_res}, max 16 bounces")
    kind = a.guided_ref
    if kind == "committed" and (a.guided_res != 512 or not os.path.exists(ROOM_REFERENCE)):
        kind = "plain"
    return guided_compare(ctx, sc, a.guided_budget, kind, a.guided_ref_spp, label)


def guided_instanced(ctx, a):
    """The same comparison on the large instanced scene at reduced resolution
    (the guide acts at the Lambert ground; Disney vertices sample the BSDF),
    and on a variant whose materials are all Lambertian (albedo = the Disney
    base colour), where the estimator has finite variance: the reference's
    Disney sampler reports a pdf that is not its sampling density, so single
    Disney bounces reach weights of 1e5 and relMSE at equal time is decided
    by the largest sample either run happens to draw."""
    import dataclasses
    from ajisai_render_amd import scenes
    w = a.inst_guided_res
    h = max(16, (w * 9) // 16)
    sc = scenes.instanced_field(res=(w, h), spp=16)
    label = (f"instanced_field: 2000 instances of 8 meshes ({sc.n_tri} instanced triangles), "
             f"HDR env 2048x1024 + directional, {w}x{h}, max 24 bounces")
    out = {"disney": guided_compare(ctx, sc, a.inst_guided_budget, "plain", a.inst_guided_ref_spp,
                                    label, runs=a.inst_guided_runs)}
    lam = dataclasses.replace(sc, materials=[
        scenes.Material(scenes.MAT_DIFFUSE, tuple(float(c) for c in m.base_color))
        for m in sc.materials])
    out["lambert_variant"] = guided_compare(ctx, lam, a.inst_guided_budget, "plain",
                                            a.inst_guided_ref_spp, label + ", all materials Lambertian",
                                            runs=a.inst_guided_runs)
    out["note"] = ("disney: relMSE of both runs is set by a few samples of weight >= 1e5 from "
                   "the Disney sampler (pdf not its sampling density); the ratio is not a "
                   "measure of the guide. lambert_variant: finite variance, but the smooth "
                   "shading normals' |ns.wi|/|ng.wi| factor leaves a few pixels carrying most "
                   "of a single image's relMSE (10 pixels: 50-95%), so the comparison takes "
                   "the median over independent runs.")
    return out


def guided_compare(ctx, sc, budget, ref_kind, ref_spp, label, runs=1):
    """Equal-time relMSE of guided vs plain path tracing on one scene; with
    runs > 1, each method renders `runs` independent images and the medians
    of relMSE x seconds are compared."""
    import numpy as np
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd.core import GpuScene, PathGuide, PathTracer, guide_defaults
    g = GpuScene(ctx, sc)
    box = sc.scene_box()
    ga = guide_defaults(warmup_passes=8, warmup_spp=16, warmup_steps_per_batch=64, train=False)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = fn()
        torch.cuda.synchronize()
        return o, time.perf_counter() - t0

    # reference
    if ref_kind == "committed":
        import numpy as np
        z = np.load(ROOM_REFERENCE)
        R = z["mean"].astype(np.float64).reshape(-1, 3)
        VR = z["var_mean"].astype(np.float64).reshape(-1, 3)
        ref, t_ref, ref_spp = None, float(z["seconds"]), int(z["spp"])
    elif ref_kind == "guided":
        gr = PathGuide(ctx, box, seed=99)
        ref, t_ref = timed(lambda: PathTracer(g, spp=ref_spp, seed_mode=1, seed_base=901)
                           .Render(guide=gr, guide_args=ga))
        gr.close()
    else:
        acc, t_ref = None, 0.0
        chunk = min(1024, ref_spp)
        for k in range(max(1, ref_spp // chunk)):
            o, dt = timed(lambda: PathTracer(g, spp=chunk, seed_mode=1, seed_base=901 + k,
                                             flags=capi.RENDER_FINITE_PDF).Render())
            t_ref += dt
            acc = o if acc is None else {"sum": acc["sum"] + o["sum"], "sum_sq": acc["sum_sq"] + o["sum_sq"],
                                         "count": acc["count"] + o["count"]}
            print(f"reference chunk {k}: {dt:.1f} s", file=sys.stderr, flush=True)
        ref = acc
    if ref is not None:
        R = _img(ref)
        n = ref["count"].astype(np.float64)
        with np.errstate(invalid="ignore", over="ignore"):
            VR = np.maximum(ref["sum_sq"].astype(np.float64) / n[:, None] - R * R, 0) / n[:, None]
    var_mean = VR
    ok = np.isfinite(var_mean).all(axis=1) & np.isfinite(R).all(axis=1)
    ref_self = float(np.mean(var_mean[ok] / (R[ok] ** 2 + RELMSE_EPS)))
    # plain: calibrate, then render for the budget. Plain paths end on a
    # zero-pdf continuation sample as guided ones do (AJG_RENDER_FINITE_PDF),
    # so the two estimators differ only where the guide is mixed in.
    fin = capi.RENDER_FINITE_PDF
    _, t32 = timed(lambda: PathTracer(g, spp=32, seed_mode=1, seed_base=2, flags=fin).Render())
    sp = max(1, int(budget / (t32 / 32)))
    # guided: calibrate warm-up time and per-spp time with a throw-away guide
    gd = PathGuide(ctx, box, seed=5)
    oc, tc = timed(lambda: PathTracer(g, spp=32, seed_mode=1, seed_base=4).Render(guide=gd, guide_args=ga))
    gd.close()
    per_spp = oc["stats"].ms_total * 1e-3 / 32
    t_warm = tc - oc["stats"].ms_total * 1e-3
    sg = max(1, int((budget - t_warm) / per_spp))
    prun, grun = [], []
    for k in range(runs):
        op, tp = timed(lambda: PathTracer(g, spp=sp, seed_mode=1, seed_base=3 + 100 * k,
                                          flags=fin).Render())
        prun.append((relmse(_img(op), R), tp, nonfinite(_img(op))))
        gd = PathGuide(ctx, box, seed=6 + k)
        og, tg = timed(lambda: PathTracer(g, spp=sg, seed_mode=1, seed_base=5 + 100 * k)
                       .Render(guide=gd, guide_args=ga))
        gst = og["guide_stats"]
        gd.close()
        grun.append((relmse(_img(og), R), tg, nonfinite(_img(og))))
    med = lambda xs: float(np.median(xs))  # noqa: E731
    rp, tp = med([r for r, _, _ in prun]), med([t for _, t, _ in prun])
    rg, tg = med([r for r, _, _ in grun]), med([t for _, t, _ in grun])
    rp_c = rp - ref_self
    rg_c = rg - ref_self
    g.close()
    return {
        "workload": label,
        "budget_s": budget,
        "reference": {"kind": ref_kind, "spp": ref_spp, "seconds": round(t_ref, 2),
                      "file": os.path.relpath(ROOM_REFERENCE, ROOT) if ref_kind == "committed" else None,
                      "mean": float(np.mean(R[ok])), "relmse_self_estimate": round(ref_self, 5),
                      "nonfinite_pixels": nonfinite(R)},
        "runs": runs,
        "plain": {"spp": sp, "seconds": round(tp, 3), "relmse": round(rp, 5),
                  "relmse_less_reference_variance": round(rp_c, 5),
                  "relmse_runs": [round(r, 5) for r, _, _ in prun],
                  "flags": "AJG_RENDER_FINITE_PDF",
                  "nonfinite_pixels": max(n for _, _, n in prun)},
        "guided": {"spp": sg, "seconds": round(tg, 3), "relmse": round(rg, 5),
                   "relmse_less_reference_variance": round(rg_c, 5),
                   "relmse_runs": [round(r, 5) for r, _, _ in grun],
                   "nonfinite_pixels": max(n for _, _, n in grun),
                   "warmup_seconds": round(gst.ms_warmup * 1e-3, 3),
                   "warmup": "8 passes x 16 spp, 64 Adam steps per 2^18-record batch, lr 1e-3",
                   "train_steps": gst.train_steps, "records": gst.records,
                   "last_loss": round(gst.last_loss, 4), "alpha": 0.5},
        "equal_time_relmse_ratio": round((rp * tp) / (rg * tg), 3),
        "equal_time_relmse_ratio_less_reference_variance":
            round((rp_c * tp) / (rg_c * tg), 3) if rg_c > 0 else None,
        "relmse": "mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2); the "
                  "'less_reference_variance' figures subtract the reference's own share "
                  "mean(Var(R) / (R^2 + 1e-2)), since E(I - R)^2 = E(I - mu)^2 + Var(R)",
    }


PROFILE_TRAFFIC = "profiles/traversal_shade_heavy_mesh_1024_64spp.txt"


def batched_bench(ctx, a):
    """Batched uneven workload: autotune, then one timed batched render, then
    the same jobs rendered alone one after another."""
    import numpy as np
    import torch
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import BatchRenderer, GpuScene, PathTracer, batch_defaults

    t = time.perf_counter()
    specs = scenes.batched_jobs(a.batch_jobs)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    gss = [GpuScene(ctx, sc) for sc in specs]
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t
    trs = [PathTracer(g, seed_mode=0) for g in gss]
    br = BatchRenderer(ctx, trs)
    films = br.films()
    # untuned batch (library defaults), also the warm-up of the buffers
    _, st0, _ = br.render_device(batch_defaults(), films=films)
    t = time.perf_counter()
    if a.batch_fixed:
        kw = {k: int(v) for k, v in (x.split("=") for x in a.batch_fixed.split(","))}
        tuned, tuned_rate = batch_defaults(**kw), 0.0
    else:
        tuned, tuned_rate = br.autotune(trial_spp=a.batch_trial_spp, films=films)
    t_tune = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st, done = br.render_device(tuned, films=films)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    # the same jobs alone, one after another (each job's own film)
    torch.cuda.synchronize()
    t = time.perf_counter()
    if not a.batch_fixed:
        for pt, f in zip(trs, films):
            pt.render_device(film=f)
    torch.cuda.synchronize()
    seq_wall = time.perf_counter() - t if not a.batch_fixed else float("nan")
    paths = int(st.paths)
    order = np.argsort(done)
    res = [int(sc.camera.res[0]) for sc in specs]
    out = {
        "jobs": len(specs),
        "triangles": int(sum(sc.n_tri for sc in specs)),
        "paths": paths,
        "seconds": round(wall, 3),
        "paths_per_s": round(paths / wall, 1),
        "device_ms": round(st.ms_total, 2),
        "mrays_per_s": round((st.closest_queries + st.any_queries) / wall / 1e6, 2),
        "iterations": st.iterations,
        "graph_launches": st.graph_launches,
        "slots": st.slots,
        "untuned_paths_per_s": round(paths / (st0.ms_total * 1e-3), 1),
        "autotuned": {k: getattr(tuned, k) for k, _ in tuned._fields_},
        "autotune_trial_spp": a.batch_trial_spp,
        "autotune_seconds": round(t_tune, 2),
        "autotune_best_trial_paths_per_s": round(tuned_rate, 1),
        "sequential_seconds": round(seq_wall, 3),
        "sequential_paths_per_s": round(paths / seq_wall, 1),
        "speedup_vs_sequential": round(seq_wall / wall, 3),
        "job_done_ms": [round(float(x), 2) for x in done],
        "job_done_ms_summary": {"min": round(float(done.min()), 2),
                                "median": round(float(np.median(done)), 2),
                                "max": round(float(done.max()), 2)},
        "job_paths": [int(sc.camera.res[0] * sc.camera.res[1] * sc.pt.spp) for sc in specs],
        "job_triangles": [int(sc.n_tri) for sc in specs],
        "job_res": res,
        "job_spp": [int(sc.pt.spp) for sc in specs],
        "first_done_job": int(order[0]),
        "scene_generation_s": round(t_gen, 2),
        "scene_build_s": round(t_build, 2),
    }
    for g in gss:
        g.close()
    return out


def instanced_bench(ctx, a):
    """One full render of the large instanced scene with per-sample streams
    (seed_mode 1, `--slots` paths in flight), and one with the reference's
    per-pixel sequential streams (seed_mode 0, one slot per traced pixel)."""
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import GpuScene, PathTracer

    t = time.perf_counter()
    sc = scenes.instanced_field(spp=a.inst_spp)
    t_gen = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    g = GpuScene(ctx, sc)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t
    w, h = sc.camera.res
    paths = w * h * sc.pt.spp
    warm = PathTracer(g, spp=1, seed_mode=0)
    warm.render_device()

    def timed(pt):
        film, _ = pt.render_device()
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, st = pt.render_device(film=film)
        torch.cuda.synchronize()
        return time.perf_counter() - t, st

    wall, st = timed(PathTracer(g, seed_mode=1, max_slots=a.slots, flags=capi.RENDER_TIMING))
    wall0, st0 = timed(PathTracer(g, seed_mode=0, flags=capi.RENDER_TIMING))
    stages = None
    if not a.no_count:
        # algorithmic bytes from an untimed counting pass of the same render
        _, cs = PathTracer(g, seed_mode=1, max_slots=a.slots,
                           flags=capi.RENDER_COUNT).render_device()
        cq, aq = st.closest_queries, st.any_queries
        b_closest = (cq * (RAY_BYTES + HIT_BYTES + 4) + cs.nodes_fetched * NODE_BYTES +
                     cs.tris_tested * LEAF_TRI_BYTES)
        b_any = (aq * (RAY_BYTES + OCC_BYTES) + cs.nodes_fetched_any * NODE_BYTES +
                 cs.tris_tested_any * LEAF_TRI_BYTES)
        stages = {
            "trace_closest": stage_roofline(
                "k_pq_closest2", b_closest, st.ms_trace_closest, st.launches_closest,
                {"nodes_per_query": round(cs.nodes_fetched / max(1, cq), 3),
                 "tris_per_query": round(cs.tris_tested / max(1, cq), 3)}),
            "trace_any": stage_roofline(
                "k_pq_any2", b_any, st.ms_trace_any, st.launches_any,
                {"nodes_per_query": round(cs.nodes_fetched_any / max(1, aq), 3),
                 "tris_per_query": round(cs.tris_tested_any / max(1, aq), 3)}),
            "byte_model": ("per query: 32 B ray + 16 B hit + 4 B instance (any-hit: 1 B) + 64 B "
                           "per node fetch + 48 B per leaf triangle; an instance visit counts "
                           "as one node (its 64 B map and mesh record) and one triangle (its "
                           "48 B box); counts from an untimed AJG_RENDER_COUNT pass"),
        }
    wp, hp = PathTracer(g, seed_mode=0).traced_dims()
    q = st.closest_queries + st.any_queries
    out = {
        "instances": len(sc.instances),
        "meshes": len(sc.primitives),
        "instanced_triangles": int(sc.n_tri),
        "unique_triangles": int(sum(p.n_tri for p in sc.primitives)),
        "resolution": [w, h],
        "traced_pixels": wp * hp,
        "spp": sc.pt.spp,
        "max_bounces": sc.pt.max_bounces,
        "seed_mode": 1,
        "slots": a.slots,
        "seconds": round(wall, 3),
        "paths_per_s": round(paths / wall, 1),
        "mrays_per_s": round(q / wall / 1e6, 2),
        "rays_per_path": round(q / paths, 3),
        "stage_ms": {"trace_closest": round(st.ms_trace_closest, 2),
                     "trace_any": round(st.ms_trace_any, 2), "shade": round(st.ms_shade, 2)},
        "iterations": st.iterations,
        "stage_roofline": stages,
        "seed0": {"seconds": round(wall0, 3), "paths_per_s": round(paths / wall0, 1),
                  "stage_ms": {"trace_closest": round(st0.ms_trace_closest, 2),
                               "trace_any": round(st0.ms_trace_any, 2),
                               "shade": round(st0.ms_shade, 2)},
                  "iterations": st0.iterations},
        "scene_generation_s": round(t_gen, 2),
        "scene_build_s": round(t_build, 3),
    }
    g.close()
    return out


def area_bench(ctx, a):
    """Cornell box lit by four emissive primitives (one-sided quad lamp,
    two-sided quad, mesh of three triangles, cube), 1024x1024: one full
    render, with the oracle timed on a bounded sample of rows beside it."""
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import GpuScene, PathTracer

    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(1024, 1024),
                             spp=a.area_spp)
    g = GpuScene(ctx, sc)
    pt = PathTracer(g, seed_mode=1, max_slots=a.slots, flags=capi.RENDER_TIMING)
    PathTracer(g, spp=1, seed_mode=1, max_slots=a.slots).render_device()
    film, _ = pt.render_device()
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st = pt.render_device(film=film)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    w, h = sc.camera.res
    paths = w * h * sc.pt.spp
    q = st.closest_queries + st.any_queries
    out = {
        "workload": "cornell_area: Cornell box, 4 area lights (quad, two-sided quad, "
                    "3-triangle mesh, cube), Lambert, max 8 bounces",
        "resolution": [w, h], "spp": sc.pt.spp, "lights": 4,
        "seconds": round(wall, 3),
        "paths_per_s": round(paths / wall, 1),
        "mrays_per_s": round(q / wall / 1e6, 2),
        "rays_per_path": round(q / paths, 3),
        "stage_ms": {"trace_closest": round(st.ms_trace_closest, 2),
                     "trace_any": round(st.ms_trace_any, 2), "shade": round(st.ms_shade, 2)},
    }
    g.close()
    if not a.no_cpu_baseline and a.area_cpu_seconds > 0:
        cpu = cpu_baseline(scenes.cornell_area(("quad", "twosided", "mesh", "cube"),
                                               res=(1024, 1024), spp=a.area_spp),
                           a.area_cpu_seconds)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu_paths"] = round(out["paths_per_s"] / cpu["value"], 1)
    return out


def sppm_bench(ctx, a):
    """SPPM with the reference's default arguments (100 iterations x 100,000
    photons, forward depth 8, photon depth 5..10, alpha 2/3) except the initial
    radius, which takes the reference's scene-scaled option (init_radius < 0:
    world-box diagonal / 1000; the default 0.1 is a tenth of this unit box),
    on the area-lit Cornell box at 1024x1024. The oracle is timed on the
    host's cores on one iteration of the same work (a visible point per
    pixel plus one iteration's photons), which every iteration of the
    reference repeats."""
    import dataclasses
    import torch
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import GpuScene, SPPMRenderer

    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(1024, 1024))
    spec = dataclasses.replace(scenes.SPPMSpec(), iteration_count=a.sppm_iters, photon_seed_base=1,
                               init_radius=-1.0)
    g = GpuScene(ctx, sc)
    SPPMRenderer(g, dataclasses.replace(spec, iteration_count=1,
                                        photons_per_iteration=1000)).render_device()
    r = SPPMRenderer(g, spec, timing=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, _, st = r.render_device()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    w, h = sc.camera.res
    out = {
        "workload": "sppm: Cornell box, 4 area lights, 1024x1024, SPPMRendererArgs defaults "
                    f"({spec.iteration_count} iterations x {spec.photons_per_iteration} photons, "
                    "forward depth 8, photon depth 5..10, alpha 2/3) with init_radius -1 "
                    "(world diagonal / 1000)",
        "seconds": round(wall, 3),
        "photons_per_s": round(st.photons / wall, 1),
        "iterations_per_s": round(spec.iteration_count / wall, 2),
        "deposits": int(st.deposits),
        "visible_points": int(st.visible_points),
        "ms_visible_points": round(st.ms_visible_points, 3),
        "ms_photons_total": round(st.ms_photons, 3),
        "ms_grid_total": round(st.ms_grid, 3),
    }
    g.close()
    if not a.no_cpu_baseline and not a.no_sppm_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        o = orc.OracleScene(sc)
        threads = cpu_threads()
        t = time.perf_counter()
        o.render_sppm(dataclasses.replace(spec, iteration_count=1), threads=threads)
        dt = time.perf_counter() - t
        o.close()
        est = dt * spec.iteration_count
        out["cpu_baseline"] = {"value": round(1.0 / dt, 4), "unit": "iterations/s",
                               "cores": threads, "kind": "port", "host": host_cpu(),
                               "sample": f"1 of {spec.iteration_count} iterations ({w * h} visible "
                                         f"points + {spec.photons_per_iteration} photons, "
                                         f"{dt:.1f} s); whole render est. {est:.0f} s"}
        out["gpu_over_cpu"] = round(est / wall, 1)
    return out


def vcm_bench(ctx, a):
    """VCM with the reference's default arguments (radius 0.003 x the world
    box's smallest extent, alpha 0.75, connections and merging, path length
    0..10, continuation 0.9) on the area-lit Cornell box at 1024x1024, for
    `--vcm-iters` iterations (one light and one camera sub-path per pixel
    each); the oracle is timed on the host's cores on one iteration of the
    same work."""
    import dataclasses
    import torch
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import GpuScene, VCMRenderer

    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(1024, 1024))
    spec = dataclasses.replace(scenes.VCMSpec(), iteration=a.vcm_iters, seed_base=1)
    g = GpuScene(ctx, sc)
    VCMRenderer(g, dataclasses.replace(spec, iteration=1)).render_device()
    r = VCMRenderer(g, spec, timing=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, _, st = r.render_device()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    w, h = sc.camera.res
    paths = 2 * w * h * spec.iteration  # light + camera sub-paths
    out = {
        "workload": "vcm: Cornell box, 4 area lights, 1024x1024, VCMRendererArgs defaults "
                    f"(radius 0.003, alpha 0.75, VC + VM, path length <= 10), "
                    f"{spec.iteration} iterations",
        "seconds": round(wall, 3),
        "subpaths_per_s": round(paths / wall, 1),
        "iterations_per_s": round(spec.iteration / wall, 2),
        "light_vertices_last_iteration": int(st.light_vertices),
        "ms_light": round(st.ms_light, 2), "ms_grid": round(st.ms_grid, 2),
        "ms_camera": round(st.ms_camera, 2),
    }
    g.close()
    if not a.no_cpu_baseline and not a.no_vcm_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        o = orc.OracleScene(sc)
        threads = cpu_threads()
        t = time.perf_counter()
        o.render_vcm(dataclasses.replace(spec, iteration=1), threads=threads)
        dt = time.perf_counter() - t
        o.close()
        out["cpu_baseline"] = {"value": round(2 * w * h / dt, 1), "unit": "subpaths/s",
                               "cores": threads, "kind": "port", "host": host_cpu(),
                               "sample": f"1 of {spec.iteration} iterations ({2 * w * h} "
                                         f"sub-paths, {dt:.1f} s)"}
        out["gpu_over_cpu"] = round(out["subpaths_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def cornell_bench(ctx, a):
    """The Cornell box of the reference's CPU-runnable configuration (5 quads
    + two rotated boxes = 34 triangles, Lambert walls, directional light of
    radiance 5 plus a constant 0.1 environment, pinhole camera), 256 x 256,
    16 spp, max 8 bounces, Russian roulette after bounce 3: the GPU render
    with the reference's per-pixel sequential streams (seed_mode 0) and with
    per-sample streams (seed_mode 1), each the mean of 20 timed renders, both
    as the wavefront and as the inline kernel (AJG_RENDER_INLINE: each pixel's
    paths traced to their ends by one thread, bitwise the same films), and
    the oracle's full render of the same image on the host (median of 3)."""
    import numpy as np
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import GpuScene, PathTracer

    sc = scenes.cornell()
    w, h = sc.camera.res
    paths = w * h * sc.pt.spp
    g = GpuScene(ctx, sc)
    out = {"workload": "cornell: 34 triangles, Lambert, directional + constant env, "
                       f"{w}x{h}, {sc.pt.spp} spp, max {sc.pt.max_bounces} bounces",
           "paths": paths}
    for mode, inline in ((0, False), (1, False), (0, True), (1, True)):
        pt = PathTracer(g, seed_mode=mode, flags=capi.RENDER_INLINE if inline else 0)
        film, _ = pt.render_device()
        for _ in range(3):
            pt.render_device(film=film)
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 20
        q = 0
        for _ in range(reps):
            _, st = pt.render_device(film=film)
            q += st.closest_queries + st.any_queries
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        out[f"seed_mode{mode}" + ("_inline" if inline else "")] = {
            "ms": round(dt * 1e3, 3), "paths_per_s": round(paths / dt, 1),
            "mrays_per_s": round(q / reps / dt / 1e6, 2),
            "slots": int(st.slots) if not inline else w * h}
    g.close()
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = cpu_threads()
        o = orc.OracleScene(sc)
        cam, args = o.camera(), o.args()
        times = []
        for _ in range(3):
            t = time.perf_counter()
            res = o.render(cam=cam, args=args, threads=threads)
            times.append(time.perf_counter() - t)
        o.close()
        dt = float(np.median(times))
        out["cpu_baseline"] = {"value": round(paths / dt, 1), "unit": "paths/s", "cores": threads,
                               "kind": "port", "host": host_cpu(),
                               "mrays_per_s": round(res["queries"] / dt / 1e6, 3),
                               "sample": f"the full image ({paths} paths), median of 3 renders "
                                         f"({', '.join(f'{x:.2f}' for x in times)} s)"}
        out["gpu_over_cpu_paths_seed_mode0"] = round(out["seed_mode0"]["paths_per_s"] /
                                                     out["cpu_baseline"]["value"], 1)
        out["gpu_over_cpu_paths_seed_mode1"] = round(out["seed_mode1"]["paths_per_s"] /
                                                     out["cpu_baseline"]["value"], 1)
        for key in ("seed_mode0_inline", "seed_mode1_inline"):
            out["gpu_over_cpu_paths_" + key] = round(out[key]["paths_per_s"] /
                                                     out["cpu_baseline"]["value"], 1)
    return out


def ploc_cpu_baseline(tri_host, box):
    """The oracle's PLOC build (one thread) of the same triangle soup."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    t = time.perf_counter()
    orc.ploc_build(tri_host, box)
    dt = time.perf_counter() - t
    return {"value": round(tri_host.shape[0] / dt / 1e6, 3), "unit": "Mtris/s", "cores": 1,
            "kind": "port", "sample": f"one build of the {tri_host.shape[0]} triangles ({dt:.2f} s)"}


def film_bench(pt, film):
    """Film resolve (box, gaussian) and 8-bit sRGB encode of the heavy-mesh
    film: device time per call and algorithmic bytes (sums 28 B per traced
    pixel read, 12 B per output pixel written; encode 12 B read, 3 B written)."""
    import torch
    w, h = pt.g.scene.camera.res
    wp, hp = pt.traced_dims()
    out = {}
    for name, fn in [("resolve_box", lambda: pt.resolve_filtered(film, "box")),
                     ("resolve_gaussian", lambda: pt.resolve_filtered(film, "gaussian"))]:
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            rgb = fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        b = 28 * wp * hp + 12 * w * h
        out[name] = {"ms": round(ms, 4), "gb_per_s": round(b / (ms * 1e-3) / 1e9, 1)}
    pt.write_srgb8(rgb)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        pt.write_srgb8(rgb)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20  # includes the 3 MB copy to the host
    out["srgb8_with_host_copy"] = {"ms": round(ms, 4)}
    out["pixels"] = w * h
    return out


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` as measured by the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the heavy-mesh render
    (calibrated against a plain copy); None when the summary is absent."""
    key = f"traffic_bytes_per_launch {kernel}:"
    try:
        with open(os.path.join(ROOT, PROFILE_TRAFFIC)) as fh:
            for line in fh:
                if line.startswith(key):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def sum_traffic(*kernels):
    """Summed profiled HBM bytes per launch of kernels that run back to back
    as one stage; None unless every one of them is in the summary."""
    t = [profiled_traffic(k) for k in kernels]
    return None if any(x is None for x in t) else sum(t)


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may
    run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "allowed": allowed}


# The CPU baselines use every CPU this process may run on, up to the host
# share of one GPU on a shared box (16 threads).
CPU_SHARE = 16


def cpu_threads():
    return max(1, min(CPU_SHARE, host_cpu()["allowed"]))


def cpu_baseline(scene, target_s):
    """Oracle (CPU restatement, oracle/liborc.so) on the host cores over a
    bounded sample of image rows spread evenly over the image (bit-reversed
    row order, so any prefix samples every part of the image alike)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = cpu_threads()
    o = orc.OracleScene(scene)
    cam = o.camera()
    args = o.args()
    wp, hp = o.traced_dims()
    # rows in bit-reversed order until the time budget is spent
    bits = max(1, (hp - 1).bit_length())
    order = [r for r in (int(format(k, f"0{bits}b")[::-1], 2) for k in range(1 << bits)) if r < hp]
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
            "host": host_cpu(),
            "mrays_per_s": round(queries / dt / 1e6, 3),
            "sample": f"{rows} image rows (bit-reversed order) x {wp} px x {args.spp} spp "
                      f"({n_paths} paths, {dt:.1f} s, {queries / dt / 1e6:.2f} Mrays/s)"}


_T0 = time.time()


def progress(msg):
    """One progress line per finished section on stderr (the JSON result is
    the only line on stdout)."""
    print(f"[bench {time.time() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); "
              "this build measures a single MI355X", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd import scenes
    from ajisai_render_amd.core import Context, GpuScene, PathTracer, PLOCAggregate

    ctx = Context(0)
    if a.only:
        fn = {"batched": batched_bench, "instanced": instanced_bench,
              "instanced_guided": guided_instanced, "mlp": lambda c, _: mlp_bench(c),
              "guided": guided_equal_time, "area": area_bench, "sppm": sppm_bench,
              "vcm": vcm_bench, "cornell": cornell_bench}[a.only]
        print(json.dumps({a.only: fn(ctx, a)}))
        return
    scene = scenes.heavy_mesh(res=(a.res, a.res), spp=a.spp)

    # --- PLOC build throughput from a device-resident soup
    tri_host = scene.soup()
    tri = torch.from_numpy(tri_host).cuda()
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
    pt = PathTracer(gs, seed_mode=a.seed_mode, max_slots=a.slots, flags=capi.RENDER_TIMING)
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

    # the reference's seeding: per-pixel sequential streams (one path per pixel)
    pps0, slots0 = None, None
    if not a.no_seed0:
        p0 = PathTracer(gs, seed_mode=0, flags=capi.RENDER_TIMING)
        p0.render_device(film=film)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, st0 = p0.render_device(film=film)
        torch.cuda.synchronize()
        pps0 = round(w * h * a.spp / (time.perf_counter() - t), 1)
        slots0 = st0.slots

    roof, stages = None, None
    if not a.no_count:
        cpt = PathTracer(gs, seed_mode=a.seed_mode, max_slots=a.slots, flags=capi.RENDER_COUNT)
        _, cs = cpt.render_device(film=film)
        assert cs.closest_queries == closest_q, "counting pass diverged from the timed pass"
        any_q = stats[0].any_queries
        ms_any = sum(s.ms_trace_any for s in stats) / a.steps
        ms_shade = sum(s.ms_shade for s in stats) / a.steps
        b_closest = (closest_q * (RAY_BYTES + HIT_BYTES) + cs.nodes_fetched * NODE_BYTES +
                     cs.tris_tested * LEAF_TRI_BYTES)
        b_any = (any_q * (RAY_BYTES + OCC_BYTES) + cs.nodes_fetched_any * NODE_BYTES +
                 cs.tris_tested_any * LEAF_TRI_BYTES)
        b_shade = closest_q * shade_vertex_bytes(1)
        stages = {
            "trace_closest": stage_roofline(
                "k_pq_closest", b_closest, ms_closest, launches_closest,
                {"nodes_per_query": round(cs.nodes_fetched / max(1, closest_q), 3),
                 "tris_per_query": round(cs.tris_tested / max(1, closest_q), 3),
                 "traffic": profiled_traffic("k_pq_closest")}),
            "trace_any": stage_roofline(
                "k_pq_any", b_any, ms_any, stats[0].launches_any,
                {"nodes_per_query": round(cs.nodes_fetched_any / max(1, any_q), 3),
                 "tris_per_query": round(cs.tris_tested_any / max(1, any_q), 3),
                 "traffic": profiled_traffic("k_pq_any")}),
            # two kernels per launch of the stage: k_shade_pend (the previous
            # bounce's direct light) then k_shade (the new hit), one stream
            "shade": stage_roofline("k_shade_pend + k_shade", b_shade, ms_shade,
                                    stats[0].launches_shade,
                                    {"bytes_per_vertex": shade_vertex_bytes(1),
                                     "traffic": sum_traffic("k_shade_pend", "k_shade")}),
        }
        # the dominant kernel: the stage with the most device time per step
        dom = max(stages, key=lambda k: stages[k]["ms_per_step"])
        d = stages[dom]
        roof = {"bound": "hbm", "achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["frac"], "traffic": d["traffic"],
                "traffic_source": PROFILE_TRAFFIC if d["traffic"] else None,
                "kernel": d["kernel"], "launches_per_step": d["launches_per_step"],
                "avg_launch_ms": d["avg_launch_ms"], "bytes_per_step": d["bytes_per_step"],
                "byte_model": ("per query: 32 B ray + 16 B hit (any-hit: 1 B) + 64 B per node "
                               "fetch + 48 B per leaf triangle; counts from an untimed "
                               "AJG_RENDER_COUNT pass of the same render")}

    progress("heavy-mesh path tracing timed")
    film_perf = film_bench(pt, film)
    mlp = mlp_bench(ctx)
    progress("film and guide network timed")
    guided = guided_equal_time(ctx, a) if a.guided_budget > 0 else None
    progress("guided equal-time comparison done")
    batched = batched_bench(ctx, a) if a.batch_jobs > 0 else None
    progress("batched jobs done")
    instanced = instanced_bench(ctx, a) if a.inst_spp > 0 else None
    progress("instanced scene done")
    area = area_bench(ctx, a) if a.area_spp > 0 else None
    progress("area lights done")
    sppm = sppm_bench(ctx, a) if a.sppm_iters > 0 else None
    progress("sppm done")
    vcm = vcm_bench(ctx, a) if a.vcm_iters > 0 else None
    progress("vcm done")
    if instanced is not None and a.inst_guided_budget > 0:
        instanced["guided"] = guided_instanced(ctx, a)
        progress("instanced guided comparison done")

    cornell = cornell_bench(ctx, a) if a.cornell else None
    progress("cornell box done")
    cpu, ploc_cpu = None, None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene, a.cpu_seconds)
        ploc_cpu = ploc_cpu_baseline(tri_host, box)
        progress("cpu baselines done")

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
        "paths_per_s_seed_mode0": pps0,
        "slots_seed_mode0": slots0,
        "mrays_per_s": round(queries / elapsed / 1e6, 2),
        "rays_per_path": round(queries / paths, 3),
        "build_mtris_per_s": round(n_tri / (min(build_ms) * 1e-3) / 1e6, 2),
        "build_ms": round(min(build_ms), 3),
        "build_iterations": build_iters,
        "build_cpu_baseline": ploc_cpu,
        "cornell": cornell,
        "stage_ms_per_step": {"trace_closest": round(ms_closest, 3),
                              "trace_any": round(sum(s.ms_trace_any for s in stats) / a.steps, 3),
                              "shade": round(sum(s.ms_shade for s in stats) / a.steps, 3)},
        "stage_roofline": stages,
        "wavefront_iterations": stats[0].iterations,
        # relMSE(guided) / relMSE(plain) at equal time (time-normalised)
        "relmse_guided_vs_plain": (round(1.0 / guided["equal_time_relmse_ratio"], 4)
                                   if guided else None),
        "guided": guided,
        "batched": batched,
        "instanced": instanced,
        "area_lights": area,
        "sppm": sppm,
        "vcm": vcm,
        "mlp": mlp,
        "film": film_perf,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu_paths"] = round(value / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
