"""AJG_RENDER_INLINE (each work item's paths traced to their ends by one
thread, the whole render one kernel) against the wavefront render on the
same seeded streams: film sums, sums of squares, counts and per-sample
radiance must be bitwise equal. Cases: the Cornell box of the reference's
CPU-runnable configuration (directional + constant environment) with
per-pixel and per-sample streams, TraceMIS and Trace, area lights of every
kind (MIS-BSDF rays traced as closest-hit queries), a Disney mesh under the
HDR environment, and zero-pdf continuations ended (AJG_RENDER_FINITE_PDF).
Combinations the inline kernel does not take (AJG_RENDER_COUNT) are
rejected."""
import numpy as np
import pytest

from ajisai_render_amd import _capi as capi
from ajisai_render_amd import scenes
from ajisai_render_amd.core import GpuScene, PathTracer

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _check(gs, label, **kw):
    flags = kw.pop("flags", 0)
    print(f"case: inline vs wavefront {label} {kw} flags {flags}", flush=True)
    w = PathTracer(gs, flags=flags, **kw).Render(per_sample=True)
    i = PathTracer(gs, flags=flags | capi.RENDER_INLINE, **kw).Render(per_sample=True)
    for k in ("sum", "sum_sq", "count", "per_sample"):
        assert _same_bits(w[k], i[k]), (label, k)
    assert i["stats"].paths == w["stats"].paths
    assert i["stats"].closest_queries == w["stats"].closest_queries
    assert i["stats"].any_queries == w["stats"].any_queries
    assert np.isfinite(i["per_sample"]).all()
    assert float(np.abs(i["sum"]).sum()) > 0.0


@pytest.mark.parametrize("seed_mode", [0, 1])
@pytest.mark.parametrize("use_mis", [True, False])
def test_inline_cornell_bitwise(gpu_ctx, seed_mode, use_mis):
    gs = GpuScene(gpu_ctx, scenes.cornell(res=(64, 48), spp=8))
    _check(gs, "cornell", spp=8, seed_mode=seed_mode, seed_base=7, use_mis=use_mis)
    gs.close()


@pytest.mark.parametrize("seed_mode", [0, 1])
def test_inline_area_lights_bitwise(gpu_ctx, seed_mode):
    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(48, 40), spp=4, env=0.0)
    gs = GpuScene(gpu_ctx, sc)
    _check(gs, "cornell_area", spp=4, seed_mode=seed_mode, seed_base=21)
    gs.close()
    sc = scenes.cornell_area(("quad",), res=(48, 40), spp=4, env=0.1)
    gs = GpuScene(gpu_ctx, sc)
    _check(gs, "cornell_area quad + env", spp=4, seed_mode=seed_mode, seed_base=22)
    gs.close()


def test_inline_disney_mesh_bitwise(gpu_ctx):
    sc = scenes.heavy_mesh(subdiv=5, res=(40, 32), spp=2)
    gs = GpuScene(gpu_ctx, sc)
    _check(gs, "heavy_mesh subdiv 5", spp=2, seed_mode=1, seed_base=3,
           flags=capi.RENDER_FINITE_PDF)
    _check(gs, "heavy_mesh subdiv 5", spp=2, seed_mode=0, seed_base=0)
    gs.close()


def test_inline_full_cornell_config(gpu_ctx):
    """The reference's CPU-runnable configuration at full size (256^2,
    16 spp) with the reference's per-pixel streams."""
    gs = GpuScene(gpu_ctx, scenes.cornell())
    _check(gs, "cornell 256x256 16 spp", spp=16, seed_mode=0, seed_base=0)
    gs.close()


def test_inline_rejects_counting(gpu_ctx):
    gs = GpuScene(gpu_ctx, scenes.cornell(res=(16, 16), spp=1))
    with pytest.raises(capi.AjgError):
        PathTracer(gs, spp=1, flags=capi.RENDER_INLINE | capi.RENDER_COUNT).Render()
    gs.close()
