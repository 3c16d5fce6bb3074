"""Pins the oracle's VCM (VCMRenderer, vcm.cpp:130-751) where the mathematics
fixes it:

- ThinLensCamera::ToRaster with the matrices VCM derives inverts
  GenerateRay (the oracle's camera is pinned in test_oracle_pt.py): a point on
  the ray through raster (x, y) projects back to (x, y), and the direction to
  the camera points at the ray's origin. With a lens, the reference's rays
  leave the lens away from the view direction (ft = focal / d.z with d.z < 0,
  thin_lens.cpp:73-77), and ToRaster maps the points on the other side of
  the lens point, along the same line, back to (x, y).
- ImageAreaToSolidAngle is f^2 / cos^3 of the angle to the view axis (the
  image plane at the focal distance).
- With connections and merging off and paths of length 1, a camera path
  only collects emission at its first hit: each pixel's sum is exactly 0, the
  lamp's radiance, or the environment's, times the iteration count; nothing
  is splatted.
"""
import dataclasses

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes


def _ray(cam, x, y, lens=(0.5, 0.5)):
    r = orc.camera_ray(cam, x, y, lens=lens)
    return np.asarray(r["o"], np.float64), np.asarray(r["d"], np.float64)


def _cam(lens=0.0, focal=1.0, res=(64, 48)):
    cs = scenes.CameraSpec(pos=(0.3, 0.6, 2.4), look_at=(0.5, 0.4, 0.0), up=(0, 1, 0), fov=45.0,
                           res=res, lens_radius=lens, focal_distance=focal)
    sc = scenes.cornell(res=res, spp=1)
    sc.camera = cs
    o = orc.OracleScene(sc)
    cam = o.camera()
    o.close()
    return cam


@pytest.mark.parametrize("lens", [0.0, 0.05])
def test_to_raster_inverts_camera_rays(lens):
    cam = _cam(lens=lens, focal=2.2)
    rng = np.random.default_rng(3)
    for _ in range(300):
        x, y = rng.uniform(0.5, 63.5), rng.uniform(0.5, 47.5)
        u2 = rng.uniform(0.05, 0.95, 2).astype(np.float32)
        o, d = _ray(cam, x, y, lens=tuple(u2))
        t = rng.uniform(0.5, 4.0)
        pos = (o + t * d if lens == 0 else o - t * d).astype(np.float32)
        ok, dir_to_cam, r = orc.vcm_to_raster(cam, u2, pos)
        assert ok
        assert abs(r[0] - x) < 2e-3 and abs(r[1] - y) < 2e-3, (x, y, r)
        # the direction to the camera ends at the ray's lens point
        back = (pos + dir_to_cam).astype(np.float64)
        lens_pt = o - 1e-4 * d  # undo the ray constructor's offset
        assert np.allclose(back, lens_pt, atol=2e-5)
    # a point behind the camera does not project
    o, d = _ray(cam, 32.0, 24.0)
    behind = o - 2.0 * d if lens == 0 else o + 2.0 * d
    ok, _, _ = orc.vcm_to_raster(cam, np.array([0.5, 0.5], np.float32), behind.astype(np.float32))
    assert not ok


def test_image_to_solid_angle_closed_form():
    f = 1.7
    cam = _cam(focal=f)
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y = rng.uniform(0, 64), rng.uniform(0, 48)
        _, d = _ray(cam, x, y)
        d = np.asarray(d, np.float64)
        # the view axis: -z of camera space, i.e. camera2world applied to (0, 0, -1)
        c2w = np.asarray(cam.camera2world, np.float64).reshape(4, 4).T
        axis = -c2w[:3, 2] / np.linalg.norm(c2w[:3, 2])
        cos = float(d @ axis / np.linalg.norm(d))
        got = orc.vcm_image_to_solid(cam, d.astype(np.float32))
        assert got == pytest.approx(f * f / cos ** 3, rel=2e-5)


def test_without_connections_or_merging_only_first_hit_emission():
    sc = scenes.cornell_area(("quad",), res=(64, 64), spp=1, env=0.1)
    o = orc.OracleScene(sc)
    try:
        col, spl, nv = o.render_vcm(scenes.VCMSpec(iteration=3, use_vc=False, use_vm=False,
                                                   max_path_length=1))
    finally:
        o.close()
    assert not spl.any() and nv == 0
    # each iteration adds the lamp, the environment or nothing (camera jitter
    # differs between iterations, so edge pixels mix them)
    lamp = np.array([12.0, 10.0, 8.0])
    px = col.reshape(-1, 3).astype(np.float64)
    ok = np.zeros(px.shape[0], bool)
    for k_lamp in range(4):
        for k_env in range(4 - k_lamp):
            want = k_lamp * lamp + k_env * 0.1
            ok |= np.all(np.abs(px - want) <= 1e-6 * (1 + want), axis=1)
    assert ok.all()
    assert np.any(np.all(px == 3 * lamp, axis=1)) and np.any(np.all(px == 0, axis=1))


def test_vcm_render_deterministic_finite():
    sc = scenes.cornell_area(("quad", "cube"), res=(64, 64), spp=1, env=0.1)
    spec = scenes.VCMSpec(iteration=2, seed_base=4)
    o = orc.OracleScene(sc)
    try:
        c1, s1, n1 = o.render_vcm(spec)
        c2, s2, n2 = o.render_vcm(spec)
        c3, _, _ = o.render_vcm(dataclasses.replace(spec, seed_base=5))
    finally:
        o.close()
    assert n1 == n2 > 0 and np.array_equal(c1, c2) and np.array_equal(s1, s2)
    assert np.isfinite(c1).all() and np.isfinite(s1).all()
    assert s1.sum() > 0 and not np.array_equal(c1, c3)
