"""Area lights (emissive primitives) on the GPU against the oracle.

Per-sample parity with the oracle as for the other lights (same random
streams; tolerance |dL| <= 1e-4 (1 + |L|) for 99% of samples, per-pixel sums
to 1e-3): NEE over quad, two-sided quad, mesh and cube emitters (the
reference's Sample and Pdf quirks included), MIS-BSDF rays that hit an
emitter (traced as closest-hit queries), first-hit emission. A batched
render of area-lit and plain jobs equals the lone renders bitwise; a guided
render converges to the plain render's value.
"""
import dataclasses

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes
from ajisai_render_amd.core import GpuScene, PathTracer
from test_gpu_render import compare_samples

pytestmark = pytest.mark.gpu

CASES = [(("quad",), 0.1), (("twosided",), 0.1), (("mesh",), 0.1), (("cube",), 0.1),
         (("quad", "twosided", "mesh", "cube"), 0.0)]


@pytest.mark.parametrize("seed_mode", [0, 1])
@pytest.mark.parametrize("emitters,env", CASES)
def test_area_light_per_sample_parity(gpu_ctx, emitters, env, seed_mode):
    print(f"case: cornell_area {emitters} env {env} seed_mode {seed_mode}", flush=True)
    sc = scenes.cornell_area(emitters, res=(48, 40), spp=4, env=env)
    gs = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    cam = o.camera()
    g = PathTracer(gs, spp=4, seed_mode=seed_mode, seed_base=21).Render(cam_override=cam,
                                                                       per_sample=True)
    r = o.render(cam=cam, args=o.args(spp=4, seed_mode=seed_mode, seed_base=21), per_sample=True)
    compare_samples(g, r)
    assert np.isfinite(g["per_sample"]).all()
    # the lamps are visible: some first hits carry emission
    assert g["per_sample"].max() >= min(max(p.emission) for p in sc.primitives if max(p.emission) > 0)
    o.close()
    gs.close()


def test_area_light_first_bounce_only(gpu_ctx):
    """max_bounces 1: camera rays that hit a lamp return its radiance plus
    the lamp's own (black-free) direct light, both as the oracle."""
    sc = scenes.cornell_area(("quad", "cube"), res=(32, 32), spp=2)
    sc.pt = dataclasses.replace(sc.pt, max_bounces=1)
    gs = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    g = PathTracer(gs, seed_mode=0).Render(per_sample=True)
    r = o.render(args=o.args(seed_mode=0), per_sample=True)
    compare_samples(g, r)
    o.close()
    gs.close()


def test_batch_mixes_area_lit_and_plain_jobs(gpu_ctx):
    """Jobs with and without area lights share one wavefront (MIS-BSDF rays
    of the area-lit jobs go to the closest-hit queue, the others' to the
    any-hit queue): every film bitwise equals the job rendered alone."""
    from ajisai_render_amd.core import BatchRenderer, batch_defaults
    print("case: batch of area-lit and plain cornell jobs", flush=True)
    scs = [scenes.cornell_area(("quad", "mesh"), res=(32, 32), spp=3, env=0.1),
           scenes.cornell(res=(48, 32), spp=2),
           scenes.cornell_area(("twosided", "cube"), res=(16, 48), spp=5)]
    gss = [GpuScene(gpu_ctx, sc) for sc in scs]
    trs = [PathTracer(g, seed_mode=0) for g in gss]
    ref = []
    for pt in trs:
        f = pt.render_device()[0]
        ref.append({k: v.cpu().numpy() for k, v in f.items()})
    for ba in (batch_defaults(use_graph=1), batch_defaults(use_graph=0, max_slots=700)):
        films, _, _ = BatchRenderer(gpu_ctx, trs).render_device(ba)
        for f, r in zip(films, ref):
            for k in ("sum", "sum_sq", "count"):
                assert f[k].cpu().numpy().tobytes() == r[k].tobytes(), k
    for g in gss:
        g.close()


def _film_mean(out):
    n = out["count"].astype(np.float64)
    m = out["sum"].astype(np.float64) / n[:, None]
    var = np.maximum(out["sum_sq"].astype(np.float64) / n[:, None] - m * m, 0.0)
    return m.mean(), np.sqrt(var.sum() / n[0]) / m.shape[0]


@pytest.mark.parametrize("train", [False, True])
def test_guided_area_lit_matches_plain(gpu_ctx, train):
    """All-Lambert Cornell box lit by quad, mesh and cube emitters and an
    environment: the guided render (mixture sampling, area-light MIS weights
    on the BSDF's own pdf) converges to the plain render's image mean."""
    from ajisai_render_amd.core import PathGuide, PathTracer, guide_defaults
    print(f"case: guided cornell_area train={train}", flush=True)
    sc = scenes.cornell_area(("quad", "mesh", "cube"), res=(32, 32), spp=512, env=0.1)
    g = GpuScene(gpu_ctx, sc)
    gd = PathGuide(gpu_ctx, sc.scene_box(), seed=3)
    op = PathTracer(g, spp=512, seed_mode=1, seed_base=1).Render()
    og = PathTracer(g, spp=512, seed_mode=1, seed_base=2).Render(
        guide=gd, guide_args=guide_defaults(alpha=0.5, train=train))
    mp, sp = _film_mean(op)
    mg, sg = _film_mean(og)
    assert abs(mp - mg) < 5 * np.hypot(sp, sg), (mp, mg, sp, sg)
    gd.close()
    g.close()


def test_area_light_limits_are_errors(gpu_ctx):
    from ajisai_render_amd import _capi as capi
    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), env=0.1)
    with pytest.raises(capi.AjgError) as e:       # 5 lights
        GpuScene(gpu_ctx, sc)
    assert e.value.status == capi.AJG_ERR_INVALID
    from instancing_util import small_instanced
    inst = small_instanced(n_inst=4)
    inst.primitives[0].emission = (1.0, 1.0, 1.0)
    with pytest.raises(capi.AjgError) as e:
        GpuScene(gpu_ctx, inst)
    assert e.value.status == capi.AJG_ERR_INVALID


def test_yaml_area_light_scene_renders_like_the_oracle(gpu_ctx, tmp_path):
    """A scene file with an emissive quad ("emission" on a geometry
    primitive) loads and renders with per-sample parity."""
    from ajisai_render_amd import loader
    from test_loader import _cornell_yaml
    ls = loader.load(_cornell_yaml(str(tmp_path), lamp=True))
    sc = ls.scene
    assert max(sc.primitives[-1].emission) > 0
    gs = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    cam = o.camera()
    g = PathTracer(gs, seed_mode=0).Render(cam_override=cam, per_sample=True)
    r = o.render(cam=cam, args=o.args(spp=sc.pt.spp, seed_mode=0), per_sample=True)
    compare_samples(g, r)
    o.close()
    gs.close()
