"""Instanced scenes on the GPU (two-level BVH): trees and hits bit-exact
against the oracle's two-level traversal, and renders that converge to
those of the world-space copy of the scene."""
import dataclasses

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes
from ajisai_render_amd._capi import HIT_DTYPE
from ajisai_render_amd.scenes import LIGHT_DIRECTIONAL, LIGHT_ENV, Light
from instancing_util import (coincident_scene, mesh_tree, oracle_trace, small_instanced,
                             tlas_tree, world_soup)
from raygen import random_rays

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).cuda()


def _hits(t):
    return t.cpu().numpy().view(HIT_DTYPE)


@pytest.mark.parametrize("seed,n_inst", [(5, 40), (11, 300)])
def test_instanced_trace_bitexact(gpu_ctx, seed, n_inst):
    from ajisai_render_amd.core import GpuScene
    print(f"case: instanced trace seed {seed} instances {n_inst}", flush=True)
    sc = small_instanced(seed=seed, n_inst=n_inst)
    g = GpuScene(gpu_ctx, sc)
    w2o, boxes = g.instances()
    # maps: the inverse of object->world; boxes hold every mapped vertex
    for k, ins in enumerate(sc.instances):
        m = np.asarray(ins.o2w, np.float64)
        full = np.vstack([m, [0, 0, 0, 1]])
        inv = np.vstack([w2o[k].reshape(3, 4).astype(np.float64), [0, 0, 0, 1]])
        assert np.allclose(inv @ full, np.eye(4), atol=1e-5)
        pos = sc.primitives[ins.mesh].positions[:, :3].astype(np.float64) @ m[:, :3].T + m[:, 3]
        assert np.all(pos >= boxes[k, :3]) and np.all(pos <= boxes[k, 3:])
    # top-level tree: the oracle's PLOC over the same boxes, bit for bit
    tn, tp, wbox = tlas_tree(boxes)
    nodes, perm = g.accel.download()
    assert nodes.tobytes() == tn.tobytes() and np.array_equal(perm, tp)
    rays = np.concatenate([random_rays(wbox, 20000, seed=seed),
                           random_rays(wbox * 0.3, 10000, seed=seed + 1)])
    hd, idev = g.Intersect(_dev(rays))
    h = _hits(hd)
    inst = idev.cpu().numpy().view(np.uint32)
    ref, rinst = oracle_trace(sc, w2o, boxes, rays)
    assert (ref["id"] != 0xFFFFFFFF).sum() > 1000
    assert h.tobytes() == ref.tobytes()
    assert np.array_equal(inst, rinst)
    occ = g.Occlude(_dev(rays)).cpu().numpy()
    assert np.array_equal(occ, oracle_trace(sc, w2o, boxes, rays, closest=False))
    # and against brute force over the world-space copy (rounding aside)
    if n_inst <= 40:
        soup, owner = world_soup(sc)
        bf = orc.brute_closest(soup, rays[:5000])
        hit = bf["id"] != 0xFFFFFFFF
        o = owner[bf["id"][hit]]
        same = (inst[:5000][hit] == o[:, 0]) & (h["tri_id"][:5000][hit] == o[:, 1])
        assert same.mean() > 0.99
    g.close()


@pytest.mark.parametrize("front", [False, True])
def test_instanced_ties_bitexact(gpu_ctx, front):
    """Coincident instances tie at every hit on them: the device visits the
    top level nearest entry first and the oracle left first, and both return
    the least (t, instance number) bit for bit."""
    from ajisai_render_amd.core import GpuScene
    print(f"case: coincident instances front={front}", flush=True)
    sc, orig = coincident_scene(front=front)
    g = GpuScene(gpu_ctx, sc)
    w2o, boxes = g.instances()
    wbox = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])
    rays = np.concatenate([random_rays(wbox, 20000, seed=21),
                           random_rays(wbox * 0.3, 10000, seed=22)])
    hd, idev = g.Intersect(_dev(rays))
    h = _hits(hd)
    inst = idev.cpu().numpy().view(np.uint32)
    ref, rinst = oracle_trace(sc, w2o, boxes, rays)
    hit = ref["id"] != 0xFFFFFFFF
    tied = hit & (np.bincount(orig, minlength=len(orig))[orig[np.minimum(rinst, len(orig) - 1)]] > 1)
    assert tied.sum() > 200, tied.sum()
    assert h.tobytes() == ref.tobytes()
    assert np.array_equal(inst, rinst)
    g.close()


def test_one_level_scene_trace_equals_accel(gpu_ctx):
    from ajisai_render_amd.core import GpuScene
    sc = scenes.cornell(res=(32, 32))
    g = GpuScene(gpu_ctx, sc)
    rays = random_rays(sc.scene_box(), 5000, seed=3)
    hd, idev = g.Intersect(_dev(rays))
    ref = g.accel.Intersect(_dev(rays))
    assert hd.cpu().numpy().tobytes() == ref.cpu().numpy().tobytes()
    h = _hits(hd)
    inst = idev.cpu().numpy().view(np.uint32)
    assert np.all(inst[h["tri_id"] == 0xFFFFFFFF] == 0xFFFFFFFF)
    assert np.all(inst[h["tri_id"] != 0xFFFFFFFF] == 0)
    g.close()


def _lit(sc, sun_gain=1e4):
    return dataclasses.replace(sc, lights=[
        Light(LIGHT_ENV, env=scenes.hdr_env(64, 32, 3, sun_gain=sun_gain)),
        Light(LIGHT_DIRECTIONAL, dir=(0.3, 0.8, 0.5), radiance=(2.0, 2.0, 2.0))])


def _stats(out):
    n = out["count"].astype(np.float64)
    m = out["sum"].astype(np.float64) / n[:, None]
    var = np.maximum(out["sum_sq"].astype(np.float64) / n[:, None] - m * m, 0.0) / n[:, None]
    return m, var


def test_instanced_render_matches_world_copy(gpu_ctx):
    """Same seeds, same pixels: per-sample radiance of the instanced scene
    equals the world-space copy's up to rounding of the ray map (samples
    whose path takes another branch aside); images agree statistically."""
    from ajisai_render_amd.core import GpuScene, PathTracer
    print("case: instanced vs world copy render 64x48", flush=True)
    sc = _lit(small_instanced(seed=5, n_inst=40))
    flat = scenes.flatten_instances(sc)
    gi, gf = GpuScene(gpu_ctx, sc), GpuScene(gpu_ctx, flat)
    oi = PathTracer(gi, spp=16, seed_mode=0, max_bounces=6).Render(per_sample=True)
    of = PathTracer(gf, spp=16, seed_mode=0, max_bounces=6).Render(per_sample=True)
    a, b = oi["per_sample"], of["per_sample"]
    close = np.all(np.abs(a - b) <= 1e-3 * (1 + np.abs(b)), axis=1)
    assert close.mean() > 0.97, close.mean()
    mi = PathTracer(gi, spp=256, seed_mode=1, seed_base=1, max_bounces=6).Render()
    mf = PathTracer(gf, spp=256, seed_mode=1, seed_base=2, max_bounces=6).Render()
    (m1, v1), (m2, v2) = _stats(mi), _stats(mf)
    se = np.sqrt(v1.sum() + v2.sum()) / m1.size
    assert abs(m1.mean() - m2.mean()) < 5 * se + 1e-6
    z = np.abs(m1 - m2) / np.sqrt(v1 + v2 + 1e-30)
    assert np.mean(z > 4) < 0.01
    gi.close()
    gf.close()


def test_instanced_render_counts_without_changing_the_image(gpu_ctx):
    """AJG_RENDER_COUNT on an instanced scene: the same film bit for bit,
    and counts of node fetches and triangles (instance visits included) for
    both ray kinds."""
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd.core import GpuScene, PathTracer
    print("case: instanced render with counting", flush=True)
    sc = _lit(small_instanced(seed=5, n_inst=40))
    g = GpuScene(gpu_ctx, sc)
    plain = PathTracer(g, spp=4, seed_mode=1, seed_base=3).Render()
    cnt = PathTracer(g, spp=4, seed_mode=1, seed_base=3, flags=capi.RENDER_COUNT).Render()
    assert plain["sum"].tobytes() == cnt["sum"].tobytes()
    st = cnt["stats"]
    assert st.nodes_fetched > st.closest_queries > 0 and st.tris_tested > 0
    assert st.nodes_fetched_any > 0 and st.tris_tested_any > 0
    g.close()


def test_instanced_guided_render_runs_unbiased(gpu_ctx):
    """Guided and plain renders of an instanced scene under the full
    lighting of the large instanced configuration (HDR env with a 1e4 sun
    patch plus a directional light) agree: image means within 5 combined
    standard errors, and per pixel |z| > 4 in under 1% of the pixels. The
    meshes carry interpolated vertex normals (shading normal != geometric
    normal), where the env light's two estimates measure different
    integrands; the guided render keeps the plain render's MIS weight
    functions there, so it converges to the same image."""
    from ajisai_render_amd.core import GpuScene, PathGuide, PathTracer, guide_defaults
    print("case: instanced guided vs plain 64x48, 1024 spp, sun 1e4 + directional", flush=True)
    sc = _lit(small_instanced(seed=5, n_inst=40))
    g = GpuScene(gpu_ctx, sc)
    gd = PathGuide(gpu_ctx, sc.scene_box(), seed=3)
    op = PathTracer(g, spp=1024, seed_mode=1, seed_base=7, max_bounces=6).Render()
    og = PathTracer(g, spp=1024, seed_mode=1, seed_base=8, max_bounces=6).Render(
        guide=gd, guide_args=guide_defaults(warmup_passes=2, warmup_spp=16, batch=1 << 12))
    assert og["guide_stats"].train_steps > 0
    (m1, v1), (m2, v2) = _stats(op), _stats(og)
    se = np.sqrt(v1.sum() + v2.sum()) / m1.size
    assert abs(m1.mean() - m2.mean()) < 5 * se + 1e-6, (m1.mean(), m2.mean(), se)
    z = np.abs(m1 - m2) / np.sqrt(v1 + v2 + 1e-30)
    print(f"means {m1.mean():.5f} {m2.mean():.5f} se {se:.2e}; |z|>4 fraction {np.mean(z > 4):.4f}",
          flush=True)
    assert np.mean(z > 4) < 0.01, np.mean(z > 4)
    gd.close()
    g.close()


def test_instanced_scene_rejected_in_batch_and_bad_maps(gpu_ctx):
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd.core import BatchRenderer, GpuScene, PathTracer
    sc = _lit(small_instanced(seed=5, n_inst=4))
    g = GpuScene(gpu_ctx, sc)
    with pytest.raises(capi.AjgError):
        BatchRenderer(gpu_ctx, [PathTracer(g)]).render_device()
    g.close()
    bad = dataclasses.replace(sc, instances=[scenes.Instance(0, np.zeros((3, 4)))])
    with pytest.raises(capi.AjgError):
        GpuScene(gpu_ctx, bad)
    bad = dataclasses.replace(sc, instances=[scenes.Instance(99, np.eye(3, 4))])
    with pytest.raises(capi.AjgError):
        GpuScene(gpu_ctx, bad)
