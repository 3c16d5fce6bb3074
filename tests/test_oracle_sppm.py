"""Pins the oracle's SPPM (SPPMRenderer, sppm.cpp:35-303; VisiblePointSearcher,
photon_mapping.h:70-151).

- Visible points: TracerVP's direct light on an all-Lambert scene (the visible
  point is the first hit) is, per pixel, exactly the path tracer's first-bounce
  sample with the same per-pixel stream (the oracle's path tracer is pinned
  by closed forms in test_oracle_pt.py), and it is accumulated once per
  iteration (iterations repeat the same visible points).
- The hash grid: a visible point registers in the cells [min, max) of its
  radius box, so it matches the photons inside its sphere that fall in those
  cells. The expected match count for uniformly scattered photons is the
  volume of sphere ∩ (registered cells) — computed here by numerical
  integration, independently of the grid code; a radius box inside one cell
  registers nowhere and matches nothing.
- UpdatePixel (sppm.cpp:145-167): the progressive radius reduction's closed
  forms (N_k = alpha M k, R_k^2 = R_0^2 prod alpha (j + 1) / (alpha j + 1))
  and the consistency of the density estimate tau_k / (k pi R_k^2) for a
  flux proportional to the gathering area.
- The photon pass end to end (TracerPhoton, AddPhoton, UpdatePixel,
  ComputePixelRadiance, sppm.cpp:108-178): on a floor lit only through a
  wall, the photon part of the radiance equals the one-bounce indirect
  light by quadrature, averaged over the region each visible point gathers
  from. Light::SampleEmit, the photons' source, is pinned in
  test_oracle_emit.py.
"""
import dataclasses
import math

import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes


def _spec(**kw):
    return dataclasses.replace(scenes.SPPMSpec(), **kw)


def test_visible_point_direct_light_equals_path_tracer_first_bounce():
    sc = scenes.cornell_area(("quad", "cube"), res=(64, 64), spp=1, env=0.1)
    sc.pt = dataclasses.replace(sc.pt, spp=1, max_bounces=1, min_bounces=3)
    o = orc.OracleScene(sc)
    try:
        pt = o.render(args=o.args(spp=1, seed_mode=0), per_sample=True)["per_sample"]
        _, st1, n1 = o.render_sppm(_spec(iteration_count=1, photons_per_iteration=0))
        _, st3, _ = o.render_sppm(_spec(iteration_count=3, photons_per_iteration=0))
    finally:
        o.close()
    d1 = st1["direct"].reshape(-1, 3)
    assert n1 == 0
    assert np.array_equal(d1, pt.astype(np.float32))
    assert st1["valid"].mean() > 0.95            # nearly every pixel sees the box
    # three iterations: the same terms added three times (term by term, so
    # equal to 3 x d1 up to float32 rounding)
    np.testing.assert_allclose(st3["direct"].reshape(-1, 3), 3.0 * d1, rtol=1e-6, atol=0)
    # every visible point is a first hit on a Lambert surface: throughput 1
    v = st1["valid"].reshape(-1) == 1
    assert np.all(st1["beta"].reshape(-1, 3)[v] == 1.0)


def _sphere_box_volume(c, r, lo, hi, n=160):
    """Volume of the ball (c, r) inside the box [lo, hi) by the midpoint rule
    over the ball's bounding cube."""
    g = (np.arange(n) + 0.5) / n * 2 * r - r
    x, y, z = np.meshgrid(c[0] + g, c[1] + g, c[2] + g, indexing="ij")
    inside = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
    inbox = ((x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) &
             (z < hi[2]))
    return (inside & inbox).sum() * (2 * r / n) ** 3


def test_grid_matches_photons_in_the_registered_cells_only():
    rng = np.random.default_rng(4)
    side, r = 1.05, 1.0
    bmin = np.zeros(3, np.float32)
    # spread out visible points (no two share a neighbourhood)
    centres = np.array([[3.3, 4.1, 5.7], [11.6, 2.4, 9.2], [6.5, 13.1, 3.4], [15.2, 15.8, 14.9],
                        [21.4, 7.7, 18.3]], np.float32)
    n_ph = 200000
    tot = 0
    for c in centres:
        ph = (c + rng.uniform(-r, r, (n_ph, 3))).astype(np.float32)
        cnt = orc.sppm_grid_probe(bmin, side, c[None], [r], ph)[0]
        lo_cell = np.floor((c - r) / side)
        hi_cell = np.floor((c + r) / side)     # exclusive upper cell
        frac = _sphere_box_volume(c.astype(np.float64), r, lo_cell * side, hi_cell * side) / (2 * r) ** 3
        full = math.pi / 6                     # the whole ball inside its cube
        expect = n_ph * frac
        sd = math.sqrt(n_ph * frac * (1 - frac)) + 1.0
        assert abs(cnt - expect) < 5 * sd + 0.01 * expect, (c, cnt, expect)
        # the upper cells are missing: fewer than all photons in the ball
        assert cnt < 0.9 * n_ph * full
        tot += cnt
    assert tot < 0.8 * len(centres) * n_ph * math.pi / 6


def test_radius_box_inside_one_cell_registers_nowhere():
    side = 1.05
    c = np.array([[2.5, 2.5, 2.5]], np.float32)   # cell (2, 2, 2) spans [2.1, 3.15)
    ph = (c + np.random.default_rng(1).normal(0, 0.05, (1000, 3))).astype(np.float32)
    assert orc.sppm_grid_probe(np.zeros(3), side, c, [0.3], ph)[0] == 0
    # a radius box over cells 1 and 2 per axis registers in cell (1, 1, 1)
    # only: photons near its centre (cell 2) are not matched, photons inside
    # the ball but in cell 1 are
    c2 = np.array([[2.2, 2.2, 2.2]], np.float32)
    rng = np.random.default_rng(2)
    near_c = (c2 + rng.uniform(-0.04, 0.04, (500, 3))).astype(np.float32)
    near_lo = (np.float32(2.0) + rng.uniform(-0.04, 0.04, (500, 3))).astype(np.float32)
    assert orc.sppm_grid_probe(np.zeros(3), side, c2, [0.45], near_c)[0] == 0
    assert orc.sppm_grid_probe(np.zeros(3), side, c2, [0.45], near_lo)[0] == 500


def test_sppm_render_deterministic_and_photons_add_light():
    sc = scenes.cornell_area(("quad",), res=(64, 64), spp=1)
    spec = _spec(iteration_count=3, photons_per_iteration=20000, photon_seed_base=9)
    o = orc.OracleScene(sc)
    try:
        r1, s1, n1 = o.render_sppm(spec)
        r2, _, n2 = o.render_sppm(spec)
        r0, _, _ = o.render_sppm(dataclasses.replace(spec, photon_seed_base=10))
    finally:
        o.close()
    assert n1 == n2 > 0 and np.array_equal(r1, r2)
    assert np.isfinite(r1).all() and (r1 >= 0).all()
    direct = s1["direct"] / 3.0
    assert (r1 >= direct - 1e-6).all() and r1.mean() > 1.05 * direct.mean()
    # radii only shrink, N grows where photons arrived
    assert (s1["radius"] <= 0.1 + 1e-7).all() and (s1["n"] >= 0).all()
    assert not np.array_equal(r0, r1)


def test_update_pixel_progressive_radius_closed_forms():
    """UpdatePixel (sppm.cpp:145-167) is the progressive radius reduction of
    Hachisuka & Jensen 2009: N' = N + alpha M, R'^2 = R^2 (N + alpha M) /
    (N + M), tau' = (tau + beta Phi) R'^2 / R^2. Consequences checked here:
    with M photons every iteration from N = 0, N_k = alpha M k and
    R_k^2 = R_0^2 prod_{j<k} alpha (j + 1) / (alpha j + 1), whatever M is;
    an iteration without photons changes nothing; and a flux proportional to
    the gathering disk's area, Phi_k = c pi R_k^2, gives the density
    estimate tau_k / (k pi R_k^2) = beta c for every k, whatever the photon
    counts (the estimate is consistent), where a tau rescaled by R'/R
    instead of R'^2/R^2, or Phi added after the rescale, drifts."""
    alpha, r0 = 2.0 / 3.0, 0.1
    for M in (1, 7, 1000):
        k = np.arange(1, 41)
        r, n, _ = orc.sppm_update_probe(alpha, r0, [1, 1, 1], np.full(40, M), np.zeros((40, 3)))
        assert np.allclose(n, alpha * M * k, rtol=1e-6)
        want = r0 * r0 * np.cumprod(alpha * k / (alpha * (k - 1) + 1))
        assert np.allclose(r * r, want, rtol=2e-5), (r * r / want)
    rng = np.random.default_rng(4)
    beta = np.array([0.3, 0.5, 0.9])
    c = 2.0
    ms = rng.integers(0, 50, 60)
    ms[[3, 10, 11]] = 0
    r_prev = r0
    radius, phis = [], []
    # feed iteration by iteration so that Phi_k follows the current radius
    for i, m in enumerate(ms):
        phi = np.full(3, c * math.pi * r_prev * r_prev if m > 0 else 0.0)
        phis.append(phi)
        r, n, tau = orc.sppm_update_probe(alpha, r0, beta, ms[:i + 1], np.array(phis))
        if m == 0:
            assert r[-1] == r_prev
        r_prev = float(r[-1])
        radius.append(r_prev)
        k = i + 1
        est = tau[-1] / (k * math.pi * r_prev * r_prev)
        assert np.allclose(est * k / np.count_nonzero(ms[:k]), beta * c, rtol=2e-4), (k, est)


def _wall_floor_scene(res=64):
    """Floor (y = 0, albedo 0.7) seen from above, a wall (x = -1, y in
    [0, 1.4], albedo 0.8) and a down-facing one-sided emitter at y = 1.5
    with a black material: no surface shadows another on a light -> wall ->
    floor path."""
    le = 6.0
    prims = [scenes.quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), 0),
             scenes.quad((-1, 0, -1), (-1, 1.4, -1), (-1, 1.4, 1), (-1, 0, 1), 1),
             dataclasses.replace(scenes.quad((-0.6, 1.5, -0.3), (0.2, 1.5, -0.3), (0.2, 1.5, 0.3),
                                             (-0.6, 1.5, 0.3), 2, two_sided=False),
                                 emission=(le, le, le))]
    mats = [scenes.Material(0, (0.7, 0.7, 0.7)), scenes.Material(0, (0.8, 0.8, 0.8)),
            scenes.Material(0, (0.0, 0.0, 0.0))]
    # sees the floor in x in [0.28, 0.92] past the emitter's edge at x = 0.2
    cam = scenes.CameraSpec(pos=(0.6, 3.0, 0.0), look_at=(0.6, 0.0, 0.0), up=(0, 0, -1),
                            fov=12.0, res=(res, res))
    return scenes.Scene("wallfloor", prims, mats, [], cam, scenes.PTSpec(spp=1, tile_size=16)), le


def _one_bounce_indirect(x, le):
    """(rho_f / pi) Int_W (rho_w / pi) E_w(y) cos_x cos_y / d^2 dA_y, with
    E_w(y) = Le Int_E cos_e cos_w / d'^2 dA_e, by Gauss-Legendre on the
    wall (32 x 48) and the emitter (16 x 16)."""
    wy, wwy = _gl(0.0, 1.4, 32)
    wz, wwz = _gl(-1.0, 1.0, 48)
    Y = np.stack(np.meshgrid(wy, wz, indexing="ij"), -1).reshape(-1, 2)
    WY = np.outer(wwy, wwz).ravel()
    wall = np.stack([np.full(len(Y), -1.0), Y[:, 0], Y[:, 1]], 1)
    ex, wex = _gl(-0.6, 0.2, 16)
    ez, wez = _gl(-0.3, 0.3, 16)
    E = np.stack(np.meshgrid(ex, ez, indexing="ij"), -1).reshape(-1, 2)
    WE = np.outer(wex, wez).ravel()
    em = np.stack([E[:, 0], np.full(len(E), 1.5), E[:, 1]], 1)
    v = em[None, :, :] - wall[:, None, :]
    d2 = np.sum(v * v, -1)
    g = (v[..., 0] * v[..., 1]) / (d2 * d2)   # cos at the wall (+x) and emitter (-y)
    e_wall = le * (g @ WE)
    lw = 0.8 / math.pi * e_wall * WY
    out = np.empty(len(x))
    for i in range(0, len(x), 4096):
        u = wall[None, :, :] - x[i:i + 4096, None, :]
        d2 = np.sum(u * u, -1)
        gg = np.maximum(u[..., 1], 0) * np.maximum(-u[..., 0], 0) / (d2 * d2)
        out[i:i + 4096] = 0.7 / math.pi * (gg @ lw)
    return out


def _gl(a, b, n):
    x, w = np.polynomial.legendre.leggauss(n)
    return a + (b - a) * 0.5 * (x + 1), 0.5 * (b - a) * w


def _disk_lattice(p, r, side, bmin):
    """Points of a lattice (spacing r / 16, offset half a step so none lies
    on the circle) on the floor inside the disk of radius r about p, and
    which of them fall in the cells [min, max) of p's radius box: the
    photons a visible point at p matches (photon_mapping.h:70-151, the rule
    the grid tests above pin)."""
    g = (np.arange(-16, 16) + 0.5) * (r / 16)
    X, Z = np.meshgrid(p[0] + g, p[2] + g, indexing="ij")
    pts = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    pts = pts[np.sum((pts - p) ** 2, 1) < 0.999 * r * r]
    keep = np.ones(len(pts), bool)
    for a in (0, 2):
        lo = math.floor(max(p[a] - r - bmin[a], 0.0) / side)
        hi = math.floor(max(p[a] + r - bmin[a], 0.0) / side)
        c = np.floor(np.maximum(pts[:, a] - bmin[a], 0.0) / side)
        keep &= (c >= lo) & (c < hi)
    return pts, keep


def test_sppm_photon_estimate_converges_to_one_bounce_indirect():
    """The whole photon pass (SampleEmit, TracerPhoton's deposit from the
    second hit on, AddPhoton's BSDF-weighted flux, UpdatePixel and
    ComputePixelRadiance's tau / (N_e pi R^2), sppm.cpp:108-178): with
    photon paths of at most two hits, the photon part of each pixel
    (radiance minus the accumulated direct light / iterations) estimates
    the floor's light -> wall -> floor radiance averaged over the photons
    the visible point gathers. With alpha = 1 the radius stays at its
    initial value (UpdatePixel's closed forms above), so every iteration
    gathers over the same region: the disk about the visible point cut to
    the grid cells it registers in (the lattice there is checked against
    the oracle's grid probe). That average of the closed-form quadrature
    agrees, over 160 pixels, with four photon seeds within their spread
    and a 1% bound for the lattice's resolution."""
    sc, le = _wall_floor_scene()
    r0 = 0.05
    spec = _spec(iteration_count=16, photons_per_iteration=150000, photon_min_depth=8,
                 photon_max_depth=2, init_radius=r0, forward_max_depth=4, update_alpha=1.0)
    o = orc.OracleScene(sc)
    parts = []
    try:
        for seed in range(4):
            rad, st, _ = o.render_sppm(dataclasses.replace(spec, photon_seed_base=seed))
            parts.append((rad[..., 0] - st["direct"][..., 0] / spec.iteration_count).ravel())
        st = st.ravel()
    finally:
        o.close()
    assert (st["valid"] == 1).all() and np.allclose(st["pos"][:, 1], 0.0, atol=1e-5)
    assert np.all(st["radius"] == np.float32(r0))
    side = 1.05 * r0
    bmin = np.array([-1.0, 0.0, -1.0]) - r0
    sub = np.random.default_rng(0).choice(len(st), 160, replace=False)
    want = np.empty(len(sub))
    for k, i in enumerate(sub):
        p = st["pos"][i].astype(np.float64)
        pts, keep = _disk_lattice(p, r0, side, bmin)
        # the lattice's matched share equals the oracle grid's
        probe = orc.sppm_grid_probe(bmin, side, p[None], [r0], pts)[0]
        assert probe == keep.sum(), (probe, keep.sum())
        want[k] = _one_bounce_indirect(pts[keep], le).sum() / len(pts)
    got = np.array([q[sub].mean() for q in parts])
    m, se = got.mean(), got.std(ddof=1) / 2.0
    print(f"photon part {got} mean {m:.5g} want {want.mean():.5g} se {se:.3g}")
    assert abs(m - want.mean()) < 4 * se + 0.01 * want.mean(), (m, want.mean(), se)


def test_threaded_sppm_matches_one_thread():
    """The threaded oracle (the CPU baseline) does the same work: the same
    deposits and, up to the order photon fluxes are summed in, the same
    radiance."""
    sc = scenes.cornell_area(("quad", "twosided", "mesh", "cube"), res=(64, 64))
    spec = _spec(iteration_count=2, photons_per_iteration=30000, photon_seed_base=1,
                 init_radius=-1.0)
    o = orc.OracleScene(sc)
    try:
        r1, s1, n1 = o.render_sppm(spec)
        r8, s8, n8 = o.render_sppm(spec, threads=8)
    finally:
        o.close()
    assert n1 == n8 > 0
    assert np.array_equal(s1["valid"], s8["valid"]) and np.array_equal(s1["n"], s8["n"])
    assert np.allclose(r1, r8, rtol=1e-4, atol=1e-6)
