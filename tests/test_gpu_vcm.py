"""VCM on the GPU (ajg_render_vcm) against the oracle.

Both use the same per-path streams, so every light and camera sub-path is
the same. Connections and direct lighting are per-pixel computations
(compared like the path tracer: |d| <= 1e-3 (1 + |x|) for at least 99% of
pixels, the rest being last-bit flips of a discrete decision); merged
contributions are summed in a different order (the grid's cells are filled
with atomics) and light-tracing splats are added atomically, so those sums
agree to float rounding: per-pixel 1e-3 for 99% of pixels and 1e-3 on the
image totals. Light-vertex counts agree to 0.5%.
"""
import dataclasses

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes
from ajisai_render_amd.core import GpuScene, VCMRenderer

pytestmark = pytest.mark.gpu


def _close(a, b, rel=1e-3):
    a = np.asarray(a, np.float64).reshape(-1, 3)
    b = np.asarray(b, np.float64).reshape(-1, 3)
    return np.abs(a - b).max(axis=1) <= rel * (1 + np.abs(b).max(axis=1))


def _metal_cornell(res=(64, 64)):
    sc = scenes.cornell_area(("quad", "cube"), res=res, spp=1, env=0.1)
    sc.materials.append(scenes.Material(1, (0.9, 0.8, 0.6), metallic=1.0, specular=0.5,
                                        roughness=0.25))
    sc.primitives[5].material = len(sc.materials) - 1
    return sc


def _lens_cornell():
    sc = scenes.cornell_area(("quad",), res=(64, 48), spp=1, env=0.2)
    sc.camera = dataclasses.replace(sc.camera, lens_radius=0.03, focal_distance=2.0)
    return sc


SCENES = {"area": lambda: scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(64, 64)),
          "metal_env": _metal_cornell,
          "directional_env": lambda: scenes.cornell(res=(64, 48), spp=1),
          "thin_lens": _lens_cornell}


@pytest.mark.parametrize("name", list(SCENES))
def test_vcm_matches_oracle(gpu_ctx, name):
    print(f"case: vcm {name}, 2 iterations", flush=True)
    sc = SCENES[name]()
    spec = scenes.VCMSpec(iteration=2, seed_base=7)
    g = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    try:
        cam = o.camera()
        out = VCMRenderer(g, spec).Render(cam_override=cam)
        col_o, spl_o, nv_o = o.render_vcm(spec, cam=cam)
    finally:
        o.close()
    col_g, spl_g, st = out["color_sum"], out["splat"], out["stats"]
    assert abs(int(st.light_vertices) - nv_o) <= max(5, nv_o // 200), (st.light_vertices, nv_o)
    # non-finite pixels where the reference's weights are (Disney's pdf of
    # (-d, d), the reverse pdf of SampleScattering, is NaN); elsewhere close
    for a, b in ((col_g, col_o), (spl_g, spl_o)):
        fa, fb = np.isfinite(a).all(axis=-1), np.isfinite(b).all(axis=-1)
        assert (fa == fb).mean() >= 0.99
        m = (fa & fb).reshape(-1)
        assert _close(a.reshape(-1, 3)[m], b.reshape(-1, 3)[m]).mean() >= 0.99
        tot = np.abs(b.reshape(-1, 3)[m]).sum()
        assert abs(a.reshape(-1, 3)[m].sum() - b.reshape(-1, 3)[m].sum()) <= 1e-3 * tot + 1e-9
    if name != "metal_env":
        assert np.isfinite(col_g).all() and np.isfinite(spl_g).all()
    g.close()


def test_vcm_image_and_modes(gpu_ctx):
    """Film::WriteImage: filtered colors (weight 1 per iteration) plus splats,
    as the oracle's resolve gives them; connection-only and merging-only
    renders match the oracle too."""
    sc = scenes.cornell_area(("quad", "cube"), res=(64, 64), env=0.1)
    g = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    try:
        cam = o.camera()
        for vc, vm in ((True, True), (True, False), (False, True)):
            spec = scenes.VCMSpec(iteration=2, use_vc=vc, use_vm=vm, seed_base=2)
            r = VCMRenderer(g, spec)
            col, spl, _ = r.render_device(cam_override=cam)
            col_o, spl_o, _ = o.render_vcm(spec, cam=cam)
            assert _close(col.cpu().numpy(), col_o).mean() >= 0.99, (vc, vm)
            if vc and vm:
                img = r.image(col, spl).cpu().numpy()
                want = orc.film_resolve_filter(64, 64, 1, col_o.reshape(-1, 3).astype(np.float64),
                                               np.full(64 * 64, 2, np.uint32), 0, (0.5, 0.5), 2.0)
                want = want + spl_o
                assert _close(img, want).mean() >= 0.99
    finally:
        o.close()
    g.close()


def test_vcm_rejects_instanced(gpu_ctx):
    from ajisai_render_amd import _capi as capi
    from instancing_util import small_instanced
    gi = GpuScene(gpu_ctx, small_instanced(n_inst=4))
    with pytest.raises(capi.AjgError) as e:
        VCMRenderer(gi, scenes.VCMSpec()).render_device()
    assert e.value.status == capi.AJG_ERR_INVALID
    gi.close()
