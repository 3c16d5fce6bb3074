"""Batched uneven jobs in one wavefront (ajg_render_batch, ajg_batch_autotune).

The judge of a batched render is the same job rendered alone
(ajg_render_pt): every job's film must be bitwise that render's film when
the work-item chunking is the same (seed_mode 0 always; seed_mode 1 with
the same samples per work item), and every sample's radiance must be
bitwise equal whatever the chunking, queue size, traversal stack, grid,
pixel order or graph replay.
"""
import dataclasses

import numpy as np
import pytest

from ajisai_render_amd import scenes

pytestmark = pytest.mark.gpu


def _env():
    return scenes.hdr_env(128, 64, 7)


def _jobs_spec(seed=3):
    """Four scenes of uneven size: (triangle target, (W, H), spp)."""
    rng = np.random.default_rng(seed)
    env = _env()
    spec = [(1e3, (32, 32), 4), (3e4, (48, 32), 8), (4e3, (16, 16), 2), (9e4, (64, 48), 3)]
    out = []
    for t, res, spp in spec:
        sc = scenes.batch_scene(rng, env, t, res[0], spp)
        sc.camera = dataclasses.replace(sc.camera, res=res)
        out.append(sc)
    return out


@pytest.fixture(scope="module")
def gscenes(gpu_ctx):
    from ajisai_render_amd.core import GpuScene
    gs = [GpuScene(gpu_ctx, sc) for sc in _jobs_spec()]
    yield gs
    for g in gs:
        g.close()


def _host(f):
    out = {k: v.cpu().numpy() for k, v in f.items()}
    out["count"] = out["count"].view(np.uint32)
    return out


def _alone(tracers, per_sample=False):
    return [_host(pt.render_device(per_sample=per_sample)[0]) for pt in tracers]


def _same(a, b, keys=("sum", "sum_sq", "count")):
    for k in keys:
        assert a[k].tobytes() == b[k].tobytes(), k


def test_batch_seed_mode0_bitwise_equals_alone(gpu_ctx, gscenes):
    """Per-pixel PCG32 streams (the reference's seeding): each job's film sums
    are bitwise those of the job rendered alone, for graph and direct launches,
    tiled pixel order and every LDS ring size."""
    from ajisai_render_amd.core import BatchRenderer, PathTracer, batch_defaults
    print("case: 4 jobs seed_mode 0", flush=True)
    trs = [PathTracer(g, seed_mode=0) for g in gscenes]
    ref = _alone(trs)
    br = BatchRenderer(gpu_ctx, trs)
    for ba in (batch_defaults(use_graph=1), batch_defaults(use_graph=0),
               batch_defaults(pixel_tile=8, lds_stack=4, iters_per_check=2),
               batch_defaults(lds_stack=16, max_slots=1000, trav_blocks_per_cu=3)):
        films, st, done = br.render_device(ba)
        assert st.paths == sum(pt.traced_dims()[0] * pt.traced_dims()[1] * pt.args.spp
                               for pt in trs)
        assert (st.graph_launches > 0) == bool(ba.use_graph)
        for f, r in zip(films, ref):
            _same(_host(f), r)
        assert np.all(done > 0) and np.all(done <= st.ms_total * 1.05 + 1.0), (done, st.ms_total)


def test_batch_seed_mode1_same_chunks_bitwise(gpu_ctx, gscenes):
    """Per-sample streams with the chunking a lone render picks for the same
    slot target: film sums bitwise equal."""
    from ajisai_render_amd.core import BatchRenderer, PathTracer, batch_defaults
    trs = [PathTracer(g, seed_mode=1, seed_base=11 + i, max_slots=1500)
           for i, g in enumerate(gscenes)]
    ref = _alone(trs)
    films, st, _ = BatchRenderer(gpu_ctx, trs).render_device(batch_defaults(max_slots=1500))
    for f, r in zip(films, ref):
        _same(_host(f), r)


@pytest.mark.parametrize("spi,tile", [(1, 0), (2, 16), (3, 4)])
def test_batch_per_sample_bitwise_any_chunking(gpu_ctx, gscenes, spi, tile):
    """Every sample's radiance is bitwise the lone render's whatever the
    samples per work item and pixel order; film counts equal spp and sums
    agree to float rounding of the chunk order."""
    from ajisai_render_amd.core import BatchRenderer, PathTracer, batch_defaults
    trs = [PathTracer(g, seed_mode=1, seed_base=5 + i) for i, g in enumerate(gscenes)]
    ref = _alone(trs, per_sample=True)
    films, _, _ = BatchRenderer(gpu_ctx, trs).render_device(
        batch_defaults(samples_per_item=spi, pixel_tile=tile, max_slots=2048), per_sample=True)
    for f, r, pt in zip(films, ref, trs):
        h = _host(f)
        assert h["per_sample"].tobytes() == r["per_sample"].tobytes()
        assert np.all(h["count"] == pt.args.spp)
        np.testing.assert_allclose(h["sum"], r["sum"], rtol=1e-5, atol=1e-5)


def test_batch_many_small_jobs(gpu_ctx, gscenes):
    """70 jobs (job table lookup, per-job completion counters): each film
    bitwise equals its lone render; completion times are positive."""
    from ajisai_render_amd.core import BatchRenderer, PathTracer, batch_defaults
    print("case: 70 jobs", flush=True)
    trs = [PathTracer(gscenes[i % 4], spp=1 + i % 3, seed_mode=0) for i in range(70)]
    ref = _alone(trs)
    films, st, done = BatchRenderer(gpu_ctx, trs).render_device(batch_defaults(max_slots=4096))
    for f, r in zip(films, ref):
        _same(_host(f), r)
    assert done.shape == (70,) and np.all(done > 0)


def test_autotune_returns_valid_setting(gpu_ctx, gscenes):
    """The autotuner leaves a valid setting with a positive rate; rendering
    with it (seed_mode 0) still gives the lone renders bitwise."""
    from ajisai_render_amd.core import BatchRenderer, PathTracer
    trs = [PathTracer(g, seed_mode=0) for g in gscenes]
    br = BatchRenderer(gpu_ctx, trs)
    ba, rate = br.autotune(trial_spp=1)
    assert rate > 0
    assert ba.lds_stack in (4, 8, 16) and ba.pixel_tile in (0, 4, 8, 16)
    assert ba.max_slots > 0 and ba.trav_blocks_per_cu > 0 and ba.iters_per_check % 2 == 0
    ref = _alone(trs)
    films, _, _ = br.render_device(ba)
    for f, r in zip(films, ref):
        _same(_host(f), r)


def test_batch_rejects_bad_arguments(gpu_ctx, gscenes):
    from ajisai_render_amd import _capi as capi
    from ajisai_render_amd.core import BatchRenderer, Context, GpuScene, PathTracer, batch_defaults
    br = BatchRenderer(gpu_ctx, [PathTracer(gscenes[0])])
    for ba in (batch_defaults(lds_stack=5), batch_defaults(pixel_tile=3),
               batch_defaults(trav_blocks_per_cu=100)):
        with pytest.raises(capi.AjgError) as e:
            br.render_device(ba)
        assert e.value.status == capi.AJG_ERR_INVALID
    other = Context(0)
    try:
        g2 = GpuScene(other, gscenes[0].scene)
        with pytest.raises(capi.AjgError):
            BatchRenderer(gpu_ctx, [PathTracer(gscenes[0]), PathTracer(g2)]).render_device()
        g2.close()
    finally:
        other.close()
