"""Pins the oracle's filtered film resolve and 8-bit sRGB image path
(film.cpp:27-63, box_filter.cpp, gaussian_filter.cpp, image.cpp:32-62,
color.h:54-64) by closed forms: a constant image stays constant under any
normalised splat; a single lit pixel reproduces the filter's own weights;
the box filter equals the 2x2 closed form already pinned; sRGB known
answers at the branch threshold, the ends and NaN; the flip reverses rows."""
import numpy as np
import pytest

import orc


def _g(d, r=2.0, a=2.0):
    return max(0.0, float(np.float32(np.exp(np.float32(-a * d * d)))) - float(np.float32(np.exp(np.float32(-a * r * r)))))


@pytest.mark.parametrize("ftype,radius", [(0, (0.5, 0.5)), (1, (2.0, 2.0)), (1, (1.5, 2.5)), (0, (1.5, 1.0))])
def test_constant_image_stays_constant(ftype, radius):
    rng = np.random.default_rng(1)
    w, h, ts = 37, 29, 16
    wp, hp = 48, 32
    n = rng.integers(1, 9, wp * hp).astype(np.uint32)
    c = np.array([0.3, 1.7, 0.05])
    img = orc.film_resolve_filter(w, h, ts, n[:, None] * c[None], n, ftype, radius)
    assert np.allclose(img, c[None, None], rtol=1e-5)


def test_box_filter_equals_closed_form_resolve():
    rng = np.random.default_rng(2)
    w, h, ts = 45, 37, 16
    wp, hp = 48, 48
    s = rng.random((wp * hp, 3)) * 5
    n = rng.integers(1, 5, wp * hp).astype(np.uint32)
    a = orc.film_resolve_filter(w, h, ts, s, n, 0, (0.5, 0.5))
    b = orc.film_resolve(w, h, ts, s, n)
    assert np.allclose(a, b, rtol=2e-6)


def test_gaussian_single_pixel_reproduces_filter_weights():
    w, h, ts = 32, 32, 16
    p0 = (13, 17)
    s = np.zeros((w * h, 3))
    s[p0[0] + p0[1] * w] = 1.0
    n = np.ones(w * h, np.uint32)
    img = orc.film_resolve_filter(w, h, ts, s, n, 1, (2.0, 2.0), 2.0)
    den = sum(_g(o + 0.5) for o in range(-2, 2)) ** 2
    for qy in range(p0[1] - 3, p0[1] + 3):
        for qx in range(p0[0] - 3, p0[0] + 3):
            dx, dy = qx - p0[0] + 0.5, qy - p0[1] + 0.5
            want = (_g(dx) * _g(dy) if -2 <= qx - p0[0] <= 1 and -2 <= qy - p0[1] <= 1 else 0.0) / den
            assert abs(img[qy, qx, 0] - want) < 1e-6 * max(1.0, want / 1e-3), (qx, qy, img[qy, qx, 0], want)


def test_srgb8_known_answers_and_flip():
    v = np.array([0.0, 1.0, 0.5, 0.003, 0.0031308, -1.0, 7.0, np.nan, 0.18, 1e-9], np.float32)
    img = np.zeros((2, len(v), 3), np.float32)
    img[0, :, :] = v[:, None]
    img[1, :, :] = 0.5
    out = orc.image_srgb8(img, flip=False)
    s = np.where(v > 0.0031308, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055, v * 12.92)
    s = np.nan_to_num(np.clip(s, 0, 1), nan=0.0)
    want = np.clip(np.floor(s * 255.5 + 0.5), 0, 255)
    assert np.array_equal(out[0, :, 0], want.astype(np.uint8)), (out[0, :, 0], want)
    assert out[0, 1, 0] == 255 and out[0, 0, 0] == 0 and out[0, 3, 0] == 10
    flipped = orc.image_srgb8(img, flip=True)
    assert np.array_equal(flipped[0], out[1]) and np.array_equal(flipped[1], out[0])
