"""Seeded ray batches for traversal parity (batches B1-B4).

Rays are stored as raw records after the Ray constructor (ray.h:33-44):
origin offset by 1e-4 along the unit direction, t_min = 1e-4, t_max = +inf
(shadow rays: finite t_max).
"""
import numpy as np

from ajisai_render_amd._capi import RAY_DTYPE

EPS = np.float32(1e-4)


def _normalize(d):
    d = d.astype(np.float32)
    n = np.sqrt((d * d).sum(axis=1, dtype=np.float32)).astype(np.float32)
    return (d * (np.float32(1) / n)[:, None]).astype(np.float32)


def make_rays(o, d, tmin=EPS, tmax=np.inf, offset=True):
    d = _normalize(np.asarray(d, np.float32))
    o = np.asarray(o, np.float32)
    if offset:
        o = (o + d * EPS).astype(np.float32)
    r = np.zeros(o.shape[0], RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def camera_rays(cam_spec, n_max=None, seed=0):
    """B1: jittered pinhole primaries through each pixel (numpy restatement of
    a look-at camera; only used to produce inputs)."""
    w, h = cam_spec.res
    rng = np.random.default_rng(seed)
    pos = np.array(cam_spec.pos, np.float64)
    fwd = np.array(cam_spec.look_at, np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.array(cam_spec.up, np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    t = np.tan(np.radians(cam_spec.fov) / 2)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    xs = xs.reshape(-1) + rng.uniform(0, 1, w * h)
    ys = ys.reshape(-1) + rng.uniform(0, 1, w * h)
    sx = (2 * xs / w - 1) * t * (w / h)
    sy = (1 - 2 * ys / h) * t
    d = fwd[None] + sx[:, None] * right[None] + sy[:, None] * up[None]
    o = np.repeat(pos[None], d.shape[0], axis=0)
    r = make_rays(o, d)
    return r if n_max is None else r[:n_max]


def random_rays(box, n, seed):
    """B2: origins uniform in 1.1x the world box, directions uniform on S^2."""
    rng = np.random.default_rng(seed)
    mn, mx = np.asarray(box[:3], np.float64), np.asarray(box[3:], np.float64)
    c, half = (mn + mx) / 2, (mx - mn) / 2 * 1.1
    o = c + rng.uniform(-1, 1, (n, 3)) * half
    d = rng.normal(size=(n, 3))
    return make_rays(o, d)


def secondary_rays(prim_rays, hits, n, seed, shadow_dir=None, shadow_tmax=None):
    """B3: continuation rays (cosine-ish about a random direction) and shadow
    rays from the hit points of a primary batch."""
    rng = np.random.default_rng(seed)
    ids = hits["tri_id"] if "tri_id" in hits.dtype.names else hits["id"]
    ok = ids != 0xFFFFFFFF
    o = prim_rays["o"][ok] + prim_rays["d"][ok] * hits["t"][ok][:, None]
    if o.shape[0] == 0:
        return make_rays(np.zeros((1, 3)), np.ones((1, 3)))
    idx = rng.integers(0, o.shape[0], n)
    o = o[idx].astype(np.float32)
    if shadow_dir is None:
        d = rng.normal(size=(n, 3))
        return make_rays(o - 1e-4 * prim_rays["d"][ok][idx], d)
    d = np.repeat(np.asarray(shadow_dir, np.float32)[None], n, axis=0)
    d = d + rng.normal(scale=1e-3, size=d.shape)
    return make_rays(o - 1e-4 * prim_rays["d"][ok][idx], d, tmax=np.float32(shadow_tmax))


def near_threshold_rays(tri, n, rng):
    """Rays through triangle centroids at grazing angles chosen so that
    Moller-Trumbore's determinant |e0 . (d x e1)| = |d . (e0 x e1)| lies
    around the reference's 1e-6 rejection threshold (0.5x .. 2x of it, both
    signs): the det window of ploc_aggregate.cpp:194."""
    t = tri[rng.integers(0, tri.shape[0], n)].astype(np.float64)
    e0, e1 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    nrm = np.cross(e0, e1)
    nl = np.linalg.norm(nrm, axis=1)
    keep = nl > 1e-12
    t, e0, nrm, nl = t[keep], e0[keep], nrm[keep], nl[keep]
    nh = nrm / nl[:, None]
    tang = e0 / np.linalg.norm(e0, axis=1, keepdims=True)
    target = rng.uniform(0.5e-6, 2e-6, t.shape[0]) * rng.choice([-1.0, 1.0], t.shape[0])
    d = tang + (target / nl)[:, None] * nh
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = t.mean(axis=1)
    size = np.linalg.norm(e0, axis=1)
    src = c - d * size[:, None]
    return make_rays(src.astype(np.float32), d.astype(np.float32), offset=False)


def adversarial_rays(tri, box, n, seed):
    """B4: axis-aligned directions with +-0 components, rays aimed at shared
    vertices and edge midpoints, origins on box planes, and grazing rays
    whose MT determinant lies around the 1e-6 threshold."""
    rng = np.random.default_rng(seed)
    tri = tri.reshape(-1, 3, 3)
    mn, mx = np.asarray(box[:3], np.float32), np.asarray(box[3:], np.float32)
    out = []
    q = n // 5
    # 1) axis-aligned with signed zeros
    o = rng.uniform(mn, mx, (q, 3)).astype(np.float32)
    axis = rng.integers(0, 3, q)
    sgn = rng.choice([-1.0, 1.0], q)
    d = np.zeros((q, 3), np.float32)
    d[np.arange(q), axis] = sgn
    zs = rng.choice([0.0, -0.0], (q, 3)).astype(np.float32)
    d = np.where(d == 0, zs, d).astype(np.float32)
    r = np.zeros(q, RAY_DTYPE)
    r["o"], r["d"], r["tmin"], r["tmax"] = o, d, EPS, np.inf
    out.append(r)
    # 2) aimed at vertices
    t = tri[rng.integers(0, tri.shape[0], q)]
    target = t[np.arange(q), rng.integers(0, 3, q)]
    src = target + rng.normal(size=(q, 3)).astype(np.float32)
    out.append(make_rays(src, target - src, offset=False))
    # 3) aimed at edge midpoints
    t = tri[rng.integers(0, tri.shape[0], q)]
    k = rng.integers(0, 3, q)
    target = (t[np.arange(q), k] + t[np.arange(q), (k + 1) % 3]) * np.float32(0.5)
    src = target + rng.normal(size=(q, 3)).astype(np.float32)
    out.append(make_rays(src, target - src, offset=False))
    # 4) MT determinant near its rejection threshold
    out.append(near_threshold_rays(tri, q, rng))
    # 5) origins on the planes of the world box, random directions
    m = n - sum(len(x) for x in out)
    o = rng.uniform(mn, mx, (m, 3)).astype(np.float32)
    ax = rng.integers(0, 3, m)
    side = rng.integers(0, 2, m)
    o[np.arange(m), ax] = np.where(side == 0, mn[ax], mx[ax])
    r = np.zeros(m, RAY_DTYPE)
    r["o"] = o
    r["d"] = _normalize(rng.normal(size=(m, 3)))
    r["tmin"], r["tmax"] = EPS, np.inf
    out.append(r)
    return np.concatenate(out)
