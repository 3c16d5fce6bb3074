"""Seeded procedural scenes for the Ajisai path-tracing hot path.

The scenes are built from the same geometry kinds the reference's factory
creates (Quad: geometry/quad.cpp, Cube: geometry/cube.cpp, MeshGeo:
geometry/mesh.cpp, optionally wrapped in TwoSided), with MeshView buffers laid
out byte for byte as those classes expose them:

* Quad   — 4 Vector3f positions, indices {0,1,2},{0,2,3} (quad.cpp:33-34,115-127)
* Cube   — 24 Vector3f positions, std::tuple<u32,u32,u32> triangles
           (idx, idx+2, idx+1), (idx, idx+3, idx+2); libstdc++ stores a tuple's
           elements in reverse, so the index buffer holds (idx+1, idx+2, idx)...
           (cube.cpp:66-75, 131-143)
* MeshGeo — Vector4f position_and_us (stride 16), Vector4f normal_and_vs, u32
           index triples (mesh.cpp:237-249)

Configurations follow the benchmark definitions: cornell (256x256 Cornell box
with two rotated boxes, directional + constant environment) and heavy_mesh
(displaced icosphere of 1,310,720 triangles on a 2-triangle ground, Disney
materials, 1024x512 HDR environment).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

GEOM_QUAD, GEOM_CUBE, GEOM_MESH = 0, 1, 2
MAT_DIFFUSE, MAT_DISNEY = 0, 1
LIGHT_DIRECTIONAL, LIGHT_ENV = 0, 1


@dataclasses.dataclass
class Primitive:
    kind: int
    positions: np.ndarray          # (n, 3) float32, or (n, 4) for MESH
    index_words: np.ndarray        # (m, 3) uint32 in memory order
    two_sided: bool
    material: int
    normal_and_vs: Optional[np.ndarray] = None  # (n, 4) float32, MESH only
    emission: tuple = (0.0, 0.0, 0.0)           # non-black: an area light

    @property
    def n_tri(self) -> int:
        return int(self.index_words.shape[0])


@dataclasses.dataclass
class Material:
    type: int
    base_color: tuple
    subsurface: float = 0.0
    metallic: float = 0.0
    specular: float = 0.0
    specular_tint: float = 0.0
    roughness: float = 0.0
    anisotropic: float = 0.0
    sheen: float = 0.0
    sheen_tint: float = 0.0
    clearcoat: float = 0.0
    clearcoat_gloss: float = 0.0


@dataclasses.dataclass
class Light:
    type: int
    dir: tuple = (0.0, 1.0, 0.0)
    angle: float = 0.0
    radiance: tuple = (0.0, 0.0, 0.0)
    env: Optional[np.ndarray] = None   # (h, w, 3) float32
    rot: float = 0.0
    scaler: float = 1.0


@dataclasses.dataclass
class CameraSpec:
    pos: tuple
    look_at: tuple
    up: tuple
    fov: float
    res: tuple
    lens_radius: float = 0.0
    focal_distance: float = 1.0


@dataclasses.dataclass
class PTSpec:
    spp: int = 16
    tile_size: int = 16
    min_bounces: int = 5
    max_bounces: int = 10
    cont_prob: float = 0.9
    use_mis: bool = True   # False: PathTracing::Trace (BSDF sampling only)


@dataclasses.dataclass
class SPPMSpec:
    """SPPMRendererArgs (renderer.h:82-97) with the creator's defaults
    (renderer_creators.cpp:63-71); photon_seed_base seeds the photon streams."""
    forward_max_depth: int = 8
    init_radius: float = 0.1
    iteration_count: int = 100
    photons_per_iteration: int = 100000
    photon_min_depth: int = 5
    photon_max_depth: int = 10
    photon_cont_prob: float = 0.9
    update_alpha: float = 2.0 / 3.0
    photon_seed_base: int = 0


@dataclasses.dataclass
class VCMSpec:
    """VCMRendererArgs (renderer.h:102-114) with the creator's defaults
    (renderer_creators.cpp:111-118); seed_base seeds the path streams."""
    iteration: int = 1
    base_radius: float = 0.003
    radius_alpha: float = 0.75
    use_vm: bool = True
    use_vc: bool = True
    min_path_length: int = 0
    max_path_length: int = 10
    cont_prob: float = 0.9
    seed_base: int = 0


KEEP_MATERIAL = -1


@dataclasses.dataclass
class Instance:
    """A mesh placed in the world: row-major 3x4 object->world affine map;
    material KEEP_MATERIAL keeps the mesh primitive's own."""
    mesh: int
    o2w: np.ndarray
    material: int = KEEP_MATERIAL


def normal_map(o2w) -> np.ndarray:
    """cbrt|det A| (A^-1)^T of the linear part A: maps normals to world space
    and is the rotation itself for a rotation with uniform scale."""
    a = np.asarray(o2w, np.float64)[:, :3]
    return np.linalg.inv(a).T * np.cbrt(abs(np.linalg.det(a)))


@dataclasses.dataclass
class Scene:
    name: str
    primitives: List[Primitive]
    materials: List[Material]
    lights: List[Light]
    camera: CameraSpec
    pt: PTSpec
    search_radius: int = 14
    instances: Optional[List[Instance]] = None  # set: primitives are the meshes

    @property
    def n_tri(self) -> int:
        if self.instances is not None:
            return sum(self.primitives[i.mesh].n_tri for i in self.instances)
        return sum(p.n_tri for p in self.primitives)

    def soup(self) -> np.ndarray:
        """Flattened (n, 9) float32 triangles in PLOC order
        (ploc_aggregate.cpp:257-284)."""
        out = []
        for p in self.primitives:
            pos = p.positions[:, :3]
            out.append(pos[p.index_words.astype(np.int64)].reshape(-1, 9))
        return np.ascontiguousarray(np.concatenate(out, axis=0).astype(np.float32))

    def scene_box(self) -> np.ndarray:
        if self.instances is not None:
            mn, mx = np.full(3, np.inf), np.full(3, -np.inf)
            for ins in self.instances:
                pos = self.primitives[ins.mesh].positions[:, :3].astype(np.float64)
                w = pos @ np.asarray(ins.o2w, np.float64)[:, :3].T + np.asarray(ins.o2w)[:, 3]
                mn, mx = np.minimum(mn, w.min(axis=0)), np.maximum(mx, w.max(axis=0))
            return np.concatenate([mn, mx]).astype(np.float32)
        mn = np.full(3, np.finfo(np.float32).max, np.float32)
        mx = np.full(3, -np.finfo(np.float32).max, np.float32)
        for p in self.primitives:
            pos = p.positions[:, :3]
            pmn, pmx = pos.min(axis=0), pos.max(axis=0)
            if np.all(pmn == pmx):
                continue
            mn, mx = np.minimum(mn, pmn), np.maximum(mx, pmx)
        return np.concatenate([mn, mx]).astype(np.float32)


# --------------------------------------------------------------- geometry
def quad(a, b, c, d, material, two_sided=True) -> Primitive:
    pos = np.array([a, b, c, d], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return Primitive(GEOM_QUAD, pos, idx, two_sided, material)


_CUBE_FACES = np.array([
    [[-.5, -.5, -.5], [-.5, -.5, .5], [.5, -.5, .5], [.5, -.5, -.5]],
    [[-.5, .5, .5], [-.5, .5, -.5], [.5, .5, -.5], [.5, .5, .5]],
    [[-.5, .5, -.5], [-.5, -.5, -.5], [.5, -.5, -.5], [.5, .5, -.5]],
    [[.5, .5, .5], [.5, -.5, .5], [-.5, -.5, .5], [-.5, .5, .5]],
    [[-.5, .5, .5], [-.5, -.5, .5], [-.5, -.5, -.5], [-.5, .5, -.5]],
    [[.5, .5, -.5], [.5, -.5, -.5], [.5, -.5, .5], [.5, .5, .5]],
], np.float64)


def cube(local2world: np.ndarray, material, two_sided=False) -> Primitive:
    """Cube of cube.cpp:31-75 transformed by a 4x4 (row-major, column vectors)."""
    v = _CUBE_FACES.reshape(-1, 3)
    h = np.concatenate([v, np.ones((v.shape[0], 1))], axis=1) @ local2world.T
    pos = (h[:, :3] / h[:, 3:4]).astype(np.float32)
    words = []
    for i in range(6):
        idx = 4 * i
        # tuple (idx, idx+2, idx+1) and (idx, idx+3, idx+2), stored reversed
        words.append([idx + 1, idx + 2, idx])
        words.append([idx + 2, idx + 3, idx])
    return Primitive(GEOM_CUBE, pos, np.array(words, np.uint32), two_sided, material)


def rot_y(deg: float) -> np.ndarray:
    t = math.radians(deg)
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float64)


def translate(x, y, z) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scale3(x, y, z) -> np.ndarray:
    return np.diag([x, y, z, 1.0])


def icosphere(subdiv: int):
    """Unit icosphere: 20*4^subdiv triangles, 10*4^subdiv + 2 vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(subdiv):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // (nv + 1), uniq % (nv + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], axis=0)
        m = inv.reshape(3, -1) + nv
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    return v, f


def _fbm(p: np.ndarray, seed: int, octaves=6, lacunarity=2.0, gain=0.5) -> np.ndarray:
    """Seeded 3-D gradient-noise fBm (Perlin-style, quintic fade)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm])
    g = rng.normal(size=(256, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)

    def noise(x):
        i0 = np.floor(x).astype(np.int64)
        fr = x - i0
        w = fr * fr * fr * (fr * (fr * 6 - 15) + 10)
        out = np.zeros(x.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ii = (i0 + np.array([dx, dy, dz])) & 255
                    h = perm[perm[perm[ii[:, 0]] + ii[:, 1]] + ii[:, 2]]
                    d = fr - np.array([dx, dy, dz])
                    val = np.einsum("ij,ij->i", g[h], d)
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    out += wx * wy * wz * val
        return out

    total = np.zeros(p.shape[0])
    amp, freq = 1.0, 1.5
    for _ in range(octaves):
        total += amp * noise(p * freq)
        amp *= gain
        freq *= lacunarity
    return total


def displaced_icosphere(subdiv: int, seed: int, amplitude=0.15, radius=1.0,
                        center=(0.0, 0.0, 0.0), scale=1.0, rotation=None):
    """MeshGeo buffers of a seeded fBm-displaced icosphere."""
    v, f = icosphere(subdiv)
    disp = _fbm(v * 1.0, seed)
    disp = disp / (np.abs(disp).max() + 1e-12)
    p = v * (radius + amplitude * disp)[:, None]
    if rotation is not None:
        p = p @ rotation.T
    p = p * scale + np.asarray(center)
    # area-weighted vertex normals of the displaced surface
    e0 = p[f[:, 1]] - p[f[:, 0]]
    e1 = p[f[:, 2]] - p[f[:, 0]]
    fn = np.cross(e0, e1)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.linalg.norm(n, axis=1, keepdims=True) + 1e-30
    # spherical uv of the undisplaced direction
    u = 0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi)
    vv = np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi
    pos4 = np.concatenate([p, u[:, None]], axis=1).astype(np.float32)
    nrm4 = np.concatenate([n, vv[:, None]], axis=1).astype(np.float32)
    return pos4, nrm4, f.astype(np.uint32)


def mesh(pos4, nrm4, faces, material, two_sided=True) -> Primitive:
    return Primitive(GEOM_MESH, np.ascontiguousarray(pos4, np.float32),
                     np.ascontiguousarray(faces, np.uint32), two_sided, material,
                     np.ascontiguousarray(nrm4, np.float32))


# ------------------------------------------------------------ environment
def hdr_env(width: int, height: int, seed: int, sigma=2.0, sun=8, sun_gain=1e4):
    """Procedural HDR map: smooth log-normal luminance field (std `sigma` in
    log space) with one hot sun x sun patch; FP32 RGB, row 0 at theta = 0."""
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid((np.arange(height) + 0.5) / height, (np.arange(width) + 0.5) / width,
                         indexing="ij")
    field = np.zeros((height, width))
    for _ in range(24):
        kx, ky = rng.integers(1, 9), rng.integers(1, 7)
        ph = rng.uniform(0, 2 * np.pi, 2)
        field += rng.normal() * np.sin(2 * np.pi * kx * xx + ph[0]) * np.sin(np.pi * ky * yy + ph[1])
    field = (field - field.mean()) / (field.std() + 1e-12) * sigma
    lum = np.exp(field)
    tint = np.stack([1.0 + 0.15 * np.sin(2 * np.pi * xx), np.ones_like(xx),
                     1.0 + 0.25 * np.cos(np.pi * yy)], axis=-1)
    rgb = lum[..., None] * tint
    med = float(np.median(lum))
    sx = int(rng.integers(0, width - sun))
    sy = int(rng.integers(height // 8, height // 2 - sun))
    rgb[sy:sy + sun, sx:sx + sun, :] = med * sun_gain
    return np.ascontiguousarray(rgb.astype(np.float32))


def constant_env(value: float):
    return np.full((1, 1, 3), value, np.float32)


# --------------------------------------------------------------- configs
def cornell(seed: int = 20261203, res=(256, 256), spp=16) -> Scene:
    """Cornell-box-style scene: open-front unit box (5 two-sided quads) and two
    boxes rotated about y by seeded angles in [0, 30] degrees; Lambert walls
    0.73, left (0.65, 0.05, 0.05), right (0.12, 0.45, 0.15); a directional
    light of radiance 5 from a seeded upper-hemisphere direction facing the
    open side, plus a constant environment of 0.1. Pinhole 256x256, 16 spp,
    max 8 bounces, Russian roulette after bounce 3."""
    rng = np.random.default_rng(seed)
    grey, red, green = 0, 1, 2
    mats = [Material(MAT_DIFFUSE, (0.73, 0.73, 0.73)),
            Material(MAT_DIFFUSE, (0.65, 0.05, 0.05)),
            Material(MAT_DIFFUSE, (0.12, 0.45, 0.15))]
    prims = [
        quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), grey),      # floor
        quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), grey),      # ceiling
        quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), grey),      # back
        quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), red),       # left
        quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), green),     # right
    ]
    a1, a2 = rng.uniform(0, 30, 2)
    prims.append(cube(translate(0.33, 0.3, 0.38) @ rot_y(a1) @ scale3(0.3, 0.6, 0.3), grey))
    prims.append(cube(translate(0.68, 0.15, 0.65) @ rot_y(-a2) @ scale3(0.3, 0.3, 0.3), grey))
    while True:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if d[1] > 0.2 and d[2] > 0.3:
            break
    lights = [Light(LIGHT_DIRECTIONAL, dir=tuple(float(x) for x in d), angle=0.0,
                    radiance=(5.0, 5.0, 5.0)),
              Light(LIGHT_ENV, env=constant_env(0.1), rot=0.0, scaler=1.0)]
    cam = CameraSpec(pos=(0.5, 0.5, 2.2), look_at=(0.5, 0.5, 0.0), up=(0.0, 1.0, 0.0),
                     fov=40.0, res=res)
    return Scene("cornell", prims, mats, lights, cam,
                 PTSpec(spp=spp, tile_size=16, min_bounces=3, max_bounces=8, cont_prob=0.9))


def cornell_area(emitters=("quad",), res=(64, 64), spp=4, env=0.0, seed: int = 20261203) -> Scene:
    """The Cornell box of `cornell` lit by emissive primitives (area lights)
    instead of the directional light; `env` > 0 adds a constant environment.
    Emitters (each one primitive, Lambert 0.5 grey underneath):
      "quad"      ceiling lamp [0.35, 0.65]^2 at y = 0.998, one-sided, facing
                  down, radiance (12, 10, 8)
      "twosided"  a vertical two-sided quad near the left wall, radiance 3
      "mesh"      three separate small triangles above the right box
                  (vertices 0..8 in order), radiance (2, 6, 3), two-sided.
                  MeshGeo::Sample reads the index array at the triangle
                  number, so the reference samples the vertex windows
                  (0 1 2), (1 2 3), (2 3 4); a window that repeats a
                  position (a closed mesh sharing vertices) has a NaN normal
                  and gives NaN samples, which this mesh avoids
      "cube"      a small cube (side 0.12) on the floor at the front,
                  radiance (6, 3, 1)"""
    base = cornell(seed=seed, res=res, spp=spp)
    prims = list(base.primitives)
    mats = list(base.materials) + [Material(MAT_DIFFUSE, (0.5, 0.5, 0.5))]
    lamp_mat = len(mats) - 1
    for kind in emitters:
        if kind == "quad":
            p = quad((0.35, 0.998, 0.35), (0.65, 0.998, 0.35), (0.65, 0.998, 0.65),
                     (0.35, 0.998, 0.65), lamp_mat, two_sided=False)
            p.emission = (12.0, 10.0, 8.0)
        elif kind == "twosided":
            p = quad((0.08, 0.45, 0.3), (0.08, 0.75, 0.3), (0.08, 0.75, 0.6), (0.08, 0.45, 0.6),
                     lamp_mat, two_sided=True)
            p.emission = (3.0, 3.0, 3.0)
        elif kind == "mesh":
            v = np.array([[0.60, 0.62, 0.55], [0.70, 0.62, 0.55], [0.65, 0.64, 0.66],
                          [0.72, 0.66, 0.60], [0.82, 0.64, 0.62], [0.76, 0.70, 0.72],
                          [0.66, 0.74, 0.70], [0.74, 0.80, 0.74], [0.62, 0.78, 0.80]], np.float32)
            faces = np.arange(9, dtype=np.uint32).reshape(3, 3)
            nrm = np.zeros((9, 3), np.float32)
            for f in faces:
                n = np.cross(v[f[1]] - v[f[0]], v[f[2]] - v[f[0]])
                nrm[f] = n / np.linalg.norm(n)
            pos4 = np.concatenate([v, np.zeros((9, 1), np.float32)], axis=1)
            nrm4 = np.concatenate([nrm, np.zeros((9, 1), np.float32)], axis=1)
            p = mesh(pos4, nrm4, faces, lamp_mat, two_sided=True)
            p.emission = (2.0, 6.0, 3.0)
        elif kind == "cube":
            p = cube(translate(0.3, 0.06, 0.85) @ scale3(0.12, 0.12, 0.12), lamp_mat)
            p.emission = (6.0, 3.0, 1.0)
        else:
            raise ValueError(kind)
        prims.append(p)
    lights = [Light(LIGHT_ENV, env=constant_env(env), rot=0.0, scaler=1.0)] if env > 0 else []
    return Scene("cornell_area", prims, mats, lights, base.camera, base.pt)


def _disney_from(rng) -> Material:
    return Material(MAT_DISNEY, tuple(float(x) for x in rng.uniform(0, 1, 3)),
                    subsurface=0.0, metallic=float(rng.uniform()), specular=float(rng.uniform()),
                    specular_tint=0.0, roughness=float(rng.uniform()), anisotropic=0.0,
                    sheen=0.0, sheen_tint=0.0, clearcoat=float(rng.uniform()),
                    clearcoat_gloss=0.5)


def heavy_mesh(seed: int = 20261204, subdiv: int = 8, res=(1024, 1024), spp=64,
               env_size=(1024, 512)) -> Scene:
    """Displaced icosphere (20*4^subdiv triangles; 1,310,720 at subdiv 8) on a
    2-triangle ground quad; Disney materials drawn per object; seeded HDR env
    map; 1024x1024, 64 spp, max 16 bounces (RR defaults: min 5, p 0.9)."""
    rng = np.random.default_rng(seed)
    mats = [_disney_from(rng), _disney_from(rng)]
    pos4, nrm4, f = displaced_icosphere(subdiv, seed + 1, amplitude=0.15)
    prims = [mesh(pos4, nrm4, f, 0, two_sided=True),
             quad((-6, -1.2, -6), (-6, -1.2, 6), (6, -1.2, 6), (6, -1.2, -6), 1)]
    env = hdr_env(env_size[0], env_size[1], seed + 2)
    lights = [Light(LIGHT_ENV, env=env, rot=0.0, scaler=1.0)]
    cam = CameraSpec(pos=(0.0, 0.6, 4.2), look_at=(0.0, -0.1, 0.0), up=(0.0, 1.0, 0.0),
                     fov=38.0, res=res)
    return Scene("heavy_mesh", prims, mats, lights, cam,
                 PTSpec(spp=spp, tile_size=16, min_bounces=5, max_bounces=16, cont_prob=0.9))


def flat_mesh(pos4, faces, material, two_sided=True) -> Primitive:
    """MeshGeo with unshared vertices whose normals are the face normals, so
    the interpolated shading normal equals the geometric one."""
    p = np.asarray(pos4, np.float64)[:, :3]
    f = np.asarray(faces, np.int64)
    tri = p[f]                                   # (m, 3, 3)
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True) + 1e-30
    m = f.shape[0]
    pos = np.concatenate([tri.reshape(-1, 3), np.zeros((3 * m, 1))], axis=1)
    nrm = np.concatenate([np.repeat(fn, 3, axis=0), np.zeros((3 * m, 1))], axis=1)
    idx = np.arange(3 * m, dtype=np.uint32).reshape(m, 3)
    return mesh(pos.astype(np.float32), nrm.astype(np.float32), idx, material, two_sided)


def occluded_room(seed: int = 20261206, res=(512, 512), spp=64, env_size=(1024, 512),
                  furniture=((6, 0.55), (6, 0.5), (5, 0.35), (5, 0.3)), smooth=False,
                  sun=None) -> Scene:
    """Closed 4 x 3 x 4 room (x, y, z in [0,4] x [0,3] x [0,4]) whose x = 4
    wall is pierced by a 0.2 x 1.0 slit (z in [1.9, 2.1], y in [1.2, 2.2]);
    lit only by the heavy-mesh scene's HDR env map through the slit. Furniture:
    flat-shaded displaced icospheres (20*4^s triangles each; 204,800 for the
    default set) on box pedestals, Lambert or Disney (clearcoat 0) from the
    seed. 512 x 512, max 16 bounces, Russian roulette after bounce 3 (p 0.9).
    smooth=True gives the icospheres their interpolated vertex normals
    (shading normal != geometric normal); sun=(radiance, toward-light
    direction) adds a directional light, e.g. one shining through the slit
    onto the floor."""
    rng = np.random.default_rng(seed)
    wall = Material(MAT_DIFFUSE, (0.75, 0.75, 0.75))
    floor = Material(MAT_DIFFUSE, (0.62, 0.56, 0.5))
    mats = [wall, floor]
    X, Y, Z = 4.0, 3.0, 4.0
    z0, z1, y0, y1 = 1.9, 2.1, 1.2, 2.2
    prims = [
        quad((0, 0, 0), (X, 0, 0), (X, 0, Z), (0, 0, Z), 1),          # floor
        quad((0, Y, 0), (0, Y, Z), (X, Y, Z), (X, Y, 0), 0),          # ceiling
        quad((0, 0, 0), (0, 0, Z), (0, Y, Z), (0, Y, 0), 0),          # x = 0
        quad((0, 0, 0), (0, Y, 0), (X, Y, 0), (X, 0, 0), 0),          # z = 0
        quad((0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z), 0),          # z = Z
        # x = X wall around the slit
        quad((X, 0, 0), (X, Y, 0), (X, Y, z0), (X, 0, z0), 0),
        quad((X, 0, z1), (X, Y, z1), (X, Y, Z), (X, 0, Z), 0),
        quad((X, 0, z0), (X, y0, z0), (X, y0, z1), (X, 0, z1), 0),
        quad((X, y1, z0), (X, Y, z0), (X, Y, z1), (X, y1, z1), 0),
    ]
    spots = [(1.2, 1.1), (2.7, 1.0), (1.0, 2.9), (2.6, 3.1)]
    for k, ((sub, rad), (cx, cz)) in enumerate(zip(furniture, spots)):
        if rng.random() < 0.5:
            m = Material(MAT_DIFFUSE, tuple(float(c) for c in rng.uniform(0.2, 0.9, 3)))
        else:
            m = _disney_from(rng)
            m.clearcoat = 0.0
        mats.append(m)
        ped = 0.25 + 0.35 * rng.random()
        prims.append(cube(translate(cx, ped / 2, cz) @ scale3(0.7, ped, 0.7), 0))
        pos4, nrm4, f = displaced_icosphere(sub, seed + 10 + k, amplitude=0.15, radius=rad,
                                            center=(cx, ped + rad, cz))
        if smooth:
            prims.append(mesh(pos4, nrm4, f, len(mats) - 1, two_sided=True))
        else:
            prims.append(flat_mesh(pos4, f, len(mats) - 1, two_sided=True))
    env = hdr_env(env_size[0], env_size[1], 20261204 + 2)
    lights = [Light(LIGHT_ENV, env=env, rot=0.0, scaler=1.0)]
    if sun is not None:
        rad_, d = sun
        d = np.asarray(d, np.float64)
        lights.append(Light(LIGHT_DIRECTIONAL, dir=tuple(float(x) for x in d / np.linalg.norm(d)),
                            radiance=tuple(float(x) for x in rad_)))
    cam = CameraSpec(pos=(0.5, 1.5, 0.5), look_at=(3.3, 0.8, 3.0), up=(0.0, 1.0, 0.0),
                     fov=75.0, res=res)
    return Scene("occluded_room", prims, mats, lights, cam,
                 PTSpec(spp=spp, tile_size=16, min_bounces=3, max_bounces=16, cont_prob=0.9))


_ICO_CACHE = {}


def _icosphere_cached(subdiv: int):
    if subdiv not in _ICO_CACHE:
        _ICO_CACHE[subdiv] = icosphere(subdiv)
    return _ICO_CACHE[subdiv]


def wavy_icosphere(subdiv: int, rng, radius=1.0, center=(0.0, 0.0, 0.0), amplitude=0.15):
    """MeshGeo buffers of an icosphere displaced along its normals by a
    seeded sum of 6 sinusoidal waves (frequencies doubling, amplitudes
    halving: an fBm-like spectrum that costs O(vertices) to evaluate)."""
    v, f = _icosphere_cached(subdiv)
    disp = np.zeros(v.shape[0])
    amp, freq = 1.0, 1.5
    for _ in range(6):
        k = rng.normal(size=3)
        k *= freq / np.linalg.norm(k)
        disp += amp * np.sin(np.pi * (v @ k) + rng.uniform(0, 2 * np.pi))
        amp *= 0.5
        freq *= 2.0
    disp /= np.abs(disp).max() + 1e-12
    p = v * (radius * (1.0 + amplitude * disp))[:, None] + np.asarray(center)
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    n = np.stack([sum(np.bincount(f[:, k], weights=fn[:, c], minlength=p.shape[0])
                      for k in range(3)) for c in range(3)], axis=1)
    n /= np.linalg.norm(n, axis=1, keepdims=True) + 1e-30
    pos4 = np.concatenate([p, np.zeros((p.shape[0], 1))], axis=1).astype(np.float32)
    nrm4 = np.concatenate([n, np.zeros((p.shape[0], 1))], axis=1).astype(np.float32)
    return pos4, nrm4, f.astype(np.uint32)


BATCH_RES = (128, 256, 512, 1024)
BATCH_SPP = (4, 16, 64)


def batch_scene(rng, env, tri_target: float, res: int, spp: int, name="batch") -> Scene:
    """One job of the batched uneven workload: a set of 1-8 displaced
    icospheres whose triangle total approaches tri_target (each sphere's
    subdivision picked for an even share of what is left), 1-8 Disney
    materials, a 2-triangle ground quad, the shared env map, res x res,
    spp samples, max 12 bounces (Russian roulette after bounce 5, p 0.9)."""
    n_sph = int(rng.integers(1, 9))
    n_mat = int(rng.integers(1, 9))
    mats = [_disney_from(rng) for _ in range(n_mat)]
    left = float(tri_target)
    prims = []
    for i in range(n_sph):
        share = max(left / (n_sph - i), 20.0)
        sub = int(np.clip(np.round(np.log(share / 20.0) / np.log(4.0)), 0, 8))
        ang = rng.uniform(0, 2 * np.pi)
        dist = 0.0 if i == 0 else rng.uniform(0.8, 2.2)
        rad = rng.uniform(0.3, 0.8) if i else 0.9
        c = (dist * np.cos(ang), rad - 1.0, dist * np.sin(ang))
        pos4, nrm4, f = wavy_icosphere(sub, rng, radius=rad, center=c)
        prims.append(mesh(pos4, nrm4, f, i % n_mat, two_sided=True))
        left -= f.shape[0]
    prims.append(quad((-8, -1.0, -8), (-8, -1.0, 8), (8, -1.0, 8), (8, -1.0, -8), 0))
    lights = [Light(LIGHT_ENV, env=env, rot=0.0, scaler=1.0)]
    az = rng.uniform(0, 2 * np.pi)
    cam = CameraSpec(pos=(5.0 * np.cos(az), 1.2, 5.0 * np.sin(az)), look_at=(0.0, -0.3, 0.0),
                     up=(0.0, 1.0, 0.0), fov=45.0, res=(res, res))
    return Scene(name, prims, mats, lights, cam,
                 PTSpec(spp=spp, tile_size=16, min_bounces=5, max_bounces=12, cont_prob=0.9))


def batched_jobs(n_jobs: int = 64, seed: int = 20261205, env_size=(1024, 512),
                 tri_range=(1e3, 2e6), res_choices=BATCH_RES, spp_choices=BATCH_SPP):
    """The batched uneven workload: n_jobs independent scenes with triangle
    totals log-uniform in tri_range, resolution uniform among res_choices
    (square), spp uniform among spp_choices, max 12 bounces, all lit by the
    heavy-mesh scene's env map."""
    rng = np.random.default_rng(seed)
    env = hdr_env(env_size[0], env_size[1], 20261204 + 2)
    out = []
    for j in range(n_jobs):
        t = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        res = int(rng.choice(res_choices))
        spp = int(rng.choice(spp_choices))
        out.append(batch_scene(rng, env, t, res, spp, name=f"batch{j}"))
    return out


def flatten_instances(scene: Scene) -> Scene:
    """World-space copy of an instanced scene (one primitive per instance):
    positions through the object->world map, mesh normals through the normal
    map (not renormalised: interpolating them equals mapping the interpolated
    object-space normal), instance materials applied."""
    prims = []
    for ins in scene.instances:
        p = scene.primitives[ins.mesh]
        m = np.asarray(ins.o2w, np.float64)
        pos = p.positions.astype(np.float64).copy()
        pos[:, :3] = pos[:, :3] @ m[:, :3].T + m[:, 3]
        nrm = None
        if p.normal_and_vs is not None:
            nrm = p.normal_and_vs.astype(np.float64).copy()
            nrm[:, :3] = nrm[:, :3] @ normal_map(m).T
            nrm = nrm.astype(np.float32)
        mat = p.material if ins.material == KEEP_MATERIAL else ins.material
        prims.append(Primitive(p.kind, pos.astype(np.float32), p.index_words, p.two_sided, mat,
                               nrm))
    return dataclasses.replace(scene, primitives=prims, instances=None)


def random_rotation(rng) -> np.ndarray:
    """Uniform random rotation (normalised Gaussian quaternion)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def instanced_field(seed: int = 20261207, n_instances: int = 2000, n_meshes: int = 8,
                    tri_range=(5e4, 5e5), ground=200.0, res=(1920, 1080), spp=256,
                    env_size=(2048, 1024), max_bounces=24) -> Scene:
    """Large instanced scene: n_meshes seeded displaced icospheres (triangle
    counts log-uniform in tri_range, rounded to the nearest icosphere level
    20*4^s), n_instances instances with uniform random rotations, scale
    log-uniform in [0.2, 5] and positions on a ground x ground quad (resting
    on it), one Disney material per mesh; HDR env map (sigma 2.5) plus a
    directional light; 1920x1080, 256 spp, max 24 bounces."""
    rng = np.random.default_rng(seed)
    mats = [_disney_from(rng) for _ in range(n_meshes)]
    mats.append(Material(MAT_DIFFUSE, (0.45, 0.42, 0.38)))
    prims = []
    for m in range(n_meshes):
        t = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        sub = int(np.clip(np.round(np.log(t / 20.0) / np.log(4.0)), 0, 8))
        pos4, nrm4, f = wavy_icosphere(sub, rng, radius=1.0)
        prims.append(mesh(pos4, nrm4, f, m, two_sided=True))
    g = ground / 2
    prims.append(quad((-g, 0.0, -g), (-g, 0.0, g), (g, 0.0, g), (g, 0.0, -g), n_meshes))
    insts = []
    for _ in range(n_instances):
        m = int(rng.integers(0, n_meshes))
        sc = float(np.exp(rng.uniform(np.log(0.2), np.log(5.0))))
        rot = random_rotation(rng) * sc
        x, z = rng.uniform(-g * 0.95, g * 0.95, 2)
        o2w = np.concatenate([rot, np.array([[x], [sc * 0.9], [z]])], axis=1)
        insts.append(Instance(m, o2w))
    insts.append(Instance(n_meshes, np.concatenate([np.eye(3), np.zeros((3, 1))], axis=1)))
    env = hdr_env(env_size[0], env_size[1], seed + 1, sigma=2.5)
    d = rng.normal(size=3)
    d[1] = abs(d[1]) + 0.5
    lights = [Light(LIGHT_ENV, env=env, rot=0.0, scaler=1.0),
              Light(LIGHT_DIRECTIONAL, dir=tuple(d / np.linalg.norm(d)), radiance=(3.0, 2.9, 2.7))]
    cam = CameraSpec(pos=(0.0, 14.0, -g * 0.9), look_at=(0.0, 2.0, 0.0), up=(0.0, 1.0, 0.0),
                     fov=50.0, res=res)
    return Scene("instanced_field", prims, mats, lights, cam,
                 PTSpec(spp=spp, tile_size=16, min_bounces=5, max_bounces=max_bounces,
                        cont_prob=0.9), instances=insts)


def random_soup(n: int, seed: int, extent=1.0, size=0.05) -> np.ndarray:
    """(n, 9) float32 random triangle soup in [-extent, extent]^3."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-extent, extent, (n, 1, 3))
    d = rng.normal(scale=size, size=(n, 3, 3))
    return np.ascontiguousarray((c + d).reshape(n, 9).astype(np.float32))


def soup_scene(tri: np.ndarray) -> Scene:
    """Wrap a triangle soup as one MeshGeo primitive (for accelerator tests)."""
    n = tri.shape[0]
    pos = tri.reshape(-1, 3)
    pos4 = np.concatenate([pos, np.zeros((pos.shape[0], 1), np.float32)], axis=1)
    nrm4 = np.zeros_like(pos4)
    nrm4[:, 1] = 1.0
    faces = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return Scene("soup", [mesh(pos4, nrm4, faces, 0)], [Material(MAT_DIFFUSE, (0.5, 0.5, 0.5))],
                 [], CameraSpec((0, 0, 3), (0, 0, 0), (0, 1, 0), 40.0, (64, 64)), PTSpec())
