"""Filtered film resolve and 8-bit sRGB output on the GPU against the
oracle (film.cpp:27-63 with box/gaussian filters; image.cpp:32-62)."""
import ctypes as C

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes

pytestmark = pytest.mark.gpu


def _resolve_dev(ctx, w, h, ts, s, n, ftype, radius, alpha=2.0):
    import torch
    from ajisai_render_amd import _capi as capi
    sd = torch.from_numpy(np.ascontiguousarray(s, np.float32).ravel()).cuda()
    nd = torch.from_numpy(np.ascontiguousarray(n, np.uint32).view(np.int32)).cuda()
    out = torch.empty(w * h * 3, dtype=torch.float32, device="cuda")
    f = capi.Filter(ftype, (C.c_float * 2)(*radius), alpha)
    ctx.check(ctx.lib.ajg_film_resolve_filter(ctx.h, w, h, ts, C.byref(f), sd.data_ptr(),
                                              nd.data_ptr(), out.data_ptr(), None))
    return out.cpu().numpy().reshape(h, w, 3)


@pytest.mark.parametrize("ftype,radius", [(0, (0.5, 0.5)), (1, (2.0, 2.0)), (1, (1.5, 2.5)),
                                          (0, (1.5, 1.0))])
@pytest.mark.parametrize("w,h", [(45, 37), (16, 16), (1, 5), (130, 3)])
def test_filtered_resolve_matches_oracle(gpu_ctx, ftype, radius, w, h):
    rng = np.random.default_rng(w * 100 + h)
    ts = 16
    wp, hp = (w + ts - 1) // ts * ts, (h + ts - 1) // ts * ts
    s = rng.random((wp * hp, 3)) * 10
    n = rng.integers(0, 5, wp * hp).astype(np.uint32)
    got = _resolve_dev(gpu_ctx, w, h, ts, s.astype(np.float32), n, ftype, radius)
    want = orc.film_resolve_filter(w, h, ts, s.astype(np.float32).astype(np.float64), n, ftype,
                                   radius)
    assert np.allclose(got, want, rtol=2e-5, atol=1e-6)


def test_srgb8_matches_oracle(gpu_ctx):
    import torch
    rng = np.random.default_rng(5)
    h, w = 97, 131
    img = (rng.random((h, w, 3)) ** 3 * 1.4 - 0.1).astype(np.float32)
    img[0, :4, 0] = [np.nan, np.inf, -np.inf, 0.0031308]
    d = torch.from_numpy(img.ravel()).cuda()
    for flip in (0, 1):
        out = torch.empty(h * w * 3, dtype=torch.uint8, device="cuda")
        gpu_ctx.check(gpu_ctx.lib.ajg_image_srgb8(gpu_ctx.h, w, h, d.data_ptr(), flip,
                                                  out.data_ptr(), None))
        got = out.cpu().numpy().reshape(h, w, 3).astype(int)
        want = orc.image_srgb8(img, flip=bool(flip)).astype(int)
        # device and host powf may differ in the last ulp, which moves a value
        # across a rounding boundary only rarely
        assert np.abs(got - want).max() <= 1
        assert np.mean(got != want) < 1e-3


def test_render_to_png_bytes_pipeline(gpu_ctx):
    """Cornell render -> gaussian resolve -> sRGB bytes on the GPU equals the
    oracle's resolve and encode of the same film (the ImageWriter path)."""
    from ajisai_render_amd.core import GpuScene, PathTracer
    sc = scenes.cornell(res=(40, 30), spp=8)
    g = GpuScene(gpu_ctx, sc)
    pt = PathTracer(g, seed_mode=0)
    film, _ = pt.render_device()
    rgb = pt.resolve_filtered(film, "gaussian")
    png = pt.write_srgb8(rgb)
    s = film["sum"].cpu().numpy().reshape(-1, 3).astype(np.float64)
    n = film["count"].cpu().numpy().view(np.uint32)
    want_rgb = orc.film_resolve_filter(40, 30, 16, s, n, 1, (2.0, 2.0), 2.0)
    assert np.allclose(rgb.cpu().numpy(), want_rgb, rtol=2e-5, atol=1e-6)
    want = orc.image_srgb8(want_rgb, flip=True).astype(int)
    assert np.abs(png.astype(int) - want).max() <= 1 and np.mean(png != want) < 1e-3
    g.close()
