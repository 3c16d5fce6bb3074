"""SPPM on the GPU (ajg_render_sppm) against the oracle.

Visible points use the path tracer's per-pixel streams, so each pixel's
visible point and direct light are compared directly (tolerance as for the
path tracer: |d| <= 1e-4 (1 + |x|) for at least 99% of pixels, the rest being
last-bit flips of a discrete decision). Photons use per-photon streams shared
by both, so every photon path and its deposits are the same; the GPU adds
deposits with atomics in another order, so the photon sums agree to float
rounding: deposit counts equal, radius and N equal (they depend on the
counts only), tau and the radiance to 1e-3 relative.
"""
import dataclasses

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes
from ajisai_render_amd.core import GpuScene, SPPMRenderer

pytestmark = pytest.mark.gpu


def _spec(**kw):
    return dataclasses.replace(scenes.SPPMSpec(), **kw)


def _close(a, b, rel=1e-4):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1)
    tol = rel * (1 + np.abs(b).reshape(b.shape[0], -1).max(axis=1))
    return d <= tol


def _metal_cornell(res=(64, 64)):
    """Area-lit Cornell box whose tall box is a Disney metal (metallic 1):
    visible points continue through it (HasDiffuseComponent false)."""
    sc = scenes.cornell_area(("quad", "cube"), res=res, spp=1, env=0.1)
    sc.materials.append(scenes.Material(1, (0.9, 0.8, 0.6), metallic=1.0, specular=0.5,
                                        roughness=0.25))
    sc.primitives[5].material = len(sc.materials) - 1
    return sc


SCENES = {"area": lambda: scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(64, 64)),
          "metal_env": _metal_cornell,
          "directional_env": lambda: scenes.cornell(res=(64, 64), spp=1)}


@pytest.mark.parametrize("name", list(SCENES))
def test_sppm_matches_oracle(gpu_ctx, name):
    print(f"case: sppm {name} 64x64, 4 iterations x 20000 photons", flush=True)
    sc = SCENES[name]()
    spec = _spec(iteration_count=4, photons_per_iteration=20000, photon_seed_base=3)
    g = GpuScene(gpu_ctx, sc)
    o = orc.OracleScene(sc)
    try:
        cam = o.camera()
        out = SPPMRenderer(g, spec).Render(cam_override=cam, state=True)
        rad_o, st_o, dep_o = o.render_sppm(spec, cam=cam)
    finally:
        o.close()
    st_g = out["state"]
    n = st_o.size
    vo, vg = st_o["valid"].reshape(-1), st_g["valid"].reshape(-1)
    assert (vo == vg).mean() >= 0.995
    both = (vo == vg) & (vo == 1)
    pos_ok = _close(st_g["pos"].reshape(n, 3)[both], st_o["pos"].reshape(n, 3)[both])
    beta_ok = _close(st_g["beta"].reshape(n, 3)[both], st_o["beta"].reshape(n, 3)[both])
    assert pos_ok.mean() >= 0.99 and beta_ok.mean() >= 0.99
    dir_ok = _close(st_g["direct"].reshape(n, 3), st_o["direct"].reshape(n, 3))
    assert dir_ok.mean() >= 0.99
    # photon deposits: the same photons reach the same visible points
    st = out["stats"]
    assert st.photons == 4 * 20000
    assert abs(int(st.deposits) - dep_o) <= max(10, dep_o // 200), (st.deposits, dep_o)
    assert dep_o > 0 or name == "directional_env"
    ok = dir_ok & (vo == vg)
    assert (st_g["radius"].reshape(-1)[ok] == st_o["radius"].reshape(-1)[ok]).mean() >= 0.99
    assert (st_g["n"].reshape(-1)[ok] == st_o["n"].reshape(-1)[ok]).mean() >= 0.99
    tau_ok = _close(st_g["tau"].reshape(n, 3)[ok], st_o["tau"].reshape(n, 3)[ok], rel=1e-3)
    assert tau_ok.mean() >= 0.99
    rg, ro = out["radiance"].reshape(n, 3), rad_o.reshape(n, 3)
    assert np.isfinite(rg).all()
    assert _close(rg, ro, rel=1e-3).mean() >= 0.99
    rel = np.abs(rg - ro).sum() / max(1e-12, np.abs(ro).sum())
    assert rel < 1e-3, rel
    g.close()


def test_sppm_film_resolve_and_repeat(gpu_ctx):
    """The per-pixel radiance resolves through the film's box filter like the
    oracle's resolve; a second render with the same arguments (buffers
    reused) gives the same visible points and direct light bitwise."""
    sc = scenes.cornell_area(("quad",), res=(128, 64))
    spec = _spec(iteration_count=2, photons_per_iteration=5000)
    g = GpuScene(gpu_ctx, sc)
    r = SPPMRenderer(g, spec)
    rad, st1, _ = r.render_device(state=True)
    rgb = r.resolve_filtered(rad).cpu().numpy()
    h, w = 64, 128
    want = orc.film_resolve_filter(w, h, 64, rad.cpu().numpy().reshape(-1, 3).astype(np.float64),
                                   np.ones(w * h, np.uint32), 0, (0.5, 0.5), 2.0)
    assert np.allclose(rgb, want, rtol=2e-5, atol=1e-7)
    _, st2, _ = r.render_device(state=True)
    a = st1.cpu().numpy().view(orc.SPPM_PIXEL_DTYPE)
    b = st2.cpu().numpy().view(orc.SPPM_PIXEL_DTYPE)
    for k in ("pos", "beta", "direct", "valid"):
        assert a[k].tobytes() == b[k].tobytes(), k
    g.close()


def test_sppm_rejects_bad_inputs(gpu_ctx):
    from ajisai_render_amd import _capi as capi
    from instancing_util import small_instanced
    g = GpuScene(gpu_ctx, scenes.cornell_area(("quad",), res=(48, 64)))
    with pytest.raises(capi.AjgError) as e:          # width not a multiple of 64
        SPPMRenderer(g, _spec(iteration_count=1)).render_device()
    assert e.value.status == capi.AJG_ERR_INVALID
    g.close()
    gi = GpuScene(gpu_ctx, small_instanced(n_inst=4))
    gi.scene.camera = dataclasses.replace(gi.scene.camera, res=(64, 64))
    with pytest.raises(capi.AjgError) as e:
        SPPMRenderer(gi, _spec(iteration_count=1)).render_device()
    assert e.value.status == capi.AJG_ERR_INVALID
    gi.close()


def test_sppm_scene_file(gpu_ctx, tmp_path):
    """A scene file with an "sppm" renderer node renders like the oracle."""
    import yaml
    from ajisai_render_amd import loader
    from test_loader import _cornell_yaml
    cfg = yaml.safe_load(open(_cornell_yaml(str(tmp_path), lamp=True)))
    cfg["camera"]["resolution"] = [64, 64]
    cfg["renderer"] = {"type": "sppm", "iteration_count": 3, "photons_per_iteration": 10000}
    ls = loader.load(cfg, base_dir=str(tmp_path))
    g = GpuScene(gpu_ctx, ls.scene)
    o = orc.OracleScene(ls.scene)
    try:
        cam = o.camera()
        out = SPPMRenderer(g, ls.sppm).Render(cam_override=cam)
        rad_o, _, _ = o.render_sppm(ls.sppm, cam=cam)
    finally:
        o.close()
    assert _close(out["radiance"].reshape(-1, 3), rad_o.reshape(-1, 3), rel=1e-3).mean() >= 0.99
    g.close()
