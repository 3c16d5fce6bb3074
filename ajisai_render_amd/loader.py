"""Ajisai YAML scene files for the GPU path.

Mirrors the reference's factory (ajisai/factory/factory.h:96-115: every node
is created by its "type" key) and creators (factory/creator/*.cpp) for what
the GPU path renders: the root keys "scene", "camera", "renderer"
(app/main.cpp:74-80); primitives of type "geometry" with "quad", "cube",
"mesh" (OBJ) and "twosided" geometries, emissive ones ("emission", an area
light, primitive_creators.cpp:65-72) included; "diffuse" and "disney"
materials with "constant" textures; "directional" and "envmap" lights
("constant" or "hdr" textures, .pfm or Radiance .hdr files); the "thin_lens"
camera with "box" or "gaussian" filters; the "pt", "sppm" and "vcm" renderers
(renderer_creators.cpp:27-135); the "ploc" aggregate, which also stands in for
"native", "embree" and a missing aggregate key (the reference's default is
the brute-force native aggregate, scene_creators.cpp:56-63; every aggregate
returns the closest hit, so only exact ties and Embree's own triangle test can
differ). Keys and defaults are the creators'. Anything else (spheres, media,
other materials, textures) raises SceneError naming it.

OBJ files are re-indexed as MeshGeoCreatorImpl does (geometry_creators.cpp:
104-240): polygons fan-triangulated, one vertex per distinct (position,
normal, texcoord) index triple in order of first use, normals computed from
the face when the file has none, positions through the transform's
transformPoint and normals through transformVector then normalised.
"""
import dataclasses
import math
import os
from typing import Optional

import numpy as np
import yaml

from . import scenes
from .scenes import (CameraSpec, Light, LIGHT_DIRECTIONAL, LIGHT_ENV, MAT_DIFFUSE, MAT_DISNEY,
                     Material, PTSpec, Primitive, Scene)


class SceneError(ValueError):
    pass


@dataclasses.dataclass
class FilterSpec:
    type: str = "box"           # "box" | "gaussian"
    radius: tuple = (0.5, 0.5)
    alpha: float = 2.0


@dataclasses.dataclass
class LoadedScene:
    scene: Scene
    filter: FilterSpec
    use_mis: bool = True
    renderer: str = "pt"                  # "pt" | "sppm" | "vcm"
    sppm: Optional[scenes.SPPMSpec] = None
    vcm: Optional[scenes.VCMSpec] = None
    aggregate: str = "ploc"               # the file's aggregate type (or "native")


# ----------------------------------------------------------------- helpers
def _vec(node, n, default=None, what="value"):
    if node is None:
        if default is None:
            raise SceneError(f"missing {what}")
        return tuple(float(x) for x in default)
    if not isinstance(node, (list, tuple)) or len(node) != n:
        raise SceneError(f"{what}: expected a sequence of {n} numbers")
    return tuple(float(x) for x in node)


def _matrix(node):
    """convert<Matrix4f>::decode (factory/creator/helper.h:68-80): 16 numbers
    row by row; the decoder reads the first 15, so the last element stays
    the identity's 1. Returned as a float32 4x4 (rows)."""
    if not isinstance(node, (list, tuple)) or len(node) != 16:
        raise SceneError("transform matrix: expected 16 numbers")
    m = np.eye(4, dtype=np.float32)
    for i in range(15):
        m[i // 4, i % 4] = np.float32(node[i])
    return m


def _matmul(a, b):
    """Matrix product in float32, k accumulated in order from 0."""
    out = np.zeros((4, 4), np.float32)
    for r in range(4):
        for c in range(4):
            acc = np.float32(0)
            for k in range(4):
                acc = np.float32(acc + np.float32(a[r, k] * b[k, c]))
            out[r, c] = acc
    return out


def _transform(node):
    """GetTransform (helper.h:101-129): absent -> identity; a sequence ->
    the product of its entries left to right; {type: matrix, matrix: [...]}."""
    if node is None:
        return np.eye(4, dtype=np.float32)
    if isinstance(node, list):
        ret = np.eye(4, dtype=np.float32)
        for n in node:
            ret = _matmul(ret, _transform(n))
        return ret
    if not isinstance(node, dict) or node.get("type", "") != "matrix" or "matrix" not in node:
        raise SceneError("transform: only {type: matrix, matrix: [16 numbers]} is defined")
    return _matrix(node["matrix"])


def _xform(m, p, w):
    """Matrix4::transformPoint (w = 1, divided by the result's w) or
    transformVector (w = 0) with the column order of the product
    (rectangular_matrix.h:199-212), in float32."""
    p = np.asarray(p, np.float32)
    cols = [p[:, 0], p[:, 1], p[:, 2], np.full(p.shape[0], np.float32(w))]
    out = []
    for r in range(4):
        acc = np.zeros(p.shape[0], np.float32)
        for c in range(4):
            acc = (acc + (np.float32(m[r, c]) * cols[c]).astype(np.float32)).astype(np.float32)
        out.append(acc)
    res = np.stack(out[:3], axis=1)
    if w == 1:
        res = (res / out[3][:, None]).astype(np.float32)
    return res


def _normalize32(v):
    v = np.asarray(v, np.float32)
    n = np.sqrt((v * v).sum(axis=1, dtype=np.float32)).astype(np.float32)
    inv = (np.float32(1) / n).astype(np.float32)
    return (v * inv[:, None]).astype(np.float32)


# --------------------------------------------------------------- file data
def read_obj(path):
    """OBJ positions, normals, texcoords and triangulated faces (each corner
    a (v, n, t) index triple, -1 when absent; negative indices resolved)."""
    vs, ns, ts, faces = [], [], [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] == "v":
                vs.append([float(x) for x in parts[1:4]])
            elif parts[0] == "vn":
                ns.append([float(x) for x in parts[1:4]])
            elif parts[0] == "vt":
                ts.append([float(x) for x in parts[1:3]] + [0.0] * max(0, 2 - len(parts[1:3])))
            elif parts[0] == "f":
                corner = []
                for tok in parts[1:]:
                    idx = tok.split("/")
                    def res(s, n):
                        if s == "":
                            return -1
                        k = int(s)
                        return k - 1 if k > 0 else n + k
                    v = res(idx[0], len(vs))
                    t = res(idx[1], len(ts)) if len(idx) > 1 else -1
                    nn = res(idx[2], len(ns)) if len(idx) > 2 else -1
                    corner.append((v, nn, t))
                for i in range(1, len(corner) - 1):   # fan triangulation
                    faces.append([corner[0], corner[i], corner[i + 1]])
    return (np.asarray(vs, np.float32).reshape(-1, 3), np.asarray(ns, np.float32).reshape(-1, 3),
            np.asarray(ts, np.float32).reshape(-1, 2), faces)


def mesh_from_obj(path, transform, material, two_sided=False):
    vs, ns, ts, faces = read_obj(path)
    if len(faces) == 0:
        raise SceneError(f"{path}: no faces")
    has_n = len(ns) > 0
    key_index = {}
    pos, nrm, uv, idx = [], [], [], []
    for f in faces:
        for c in f:
            if c in key_index:
                idx.append(key_index[c])
                continue
            v, n, t = c
            if has_n:
                if n < 0:
                    raise SceneError(f"{path}: a face without normal indices in a file with normals")
                normal = ns[n]
            else:
                a, b, cc = (vs[f[k][0]] for k in range(3))
                normal = np.cross((b - a).astype(np.float32), (cc - a).astype(np.float32))
                normal = _normalize32(normal[None])[0]
            key_index[c] = len(pos)
            idx.append(len(pos))
            pos.append(vs[v])
            nrm.append(normal)
            uv.append(ts[t] if t >= 0 else np.zeros(2, np.float32))
    pos = _xform(transform, np.asarray(pos, np.float32), 1)
    nrm = _normalize32(_xform(transform, np.asarray(nrm, np.float32), 0))
    uv = np.asarray(uv, np.float32)
    pos4 = np.concatenate([pos, uv[:, :1]], axis=1)
    nrm4 = np.concatenate([nrm, uv[:, 1:2]], axis=1)
    return scenes.mesh(pos4, nrm4, np.asarray(idx, np.uint32).reshape(-1, 3), material, two_sided)


def read_pfm(path):
    """ReadPFM (texture2d_creators.cpp:178-243): header words, rows stored
    bottom to top, |scale| multiplies, sign of scale = endianness."""
    with open(path, "rb") as f:
        data = f.read()
    words, pos = [], 0
    while len(words) < 4:
        while data[pos:pos + 1] in (b" ", b"\n", b"\t", b"\r"):
            pos += 1
        start = pos
        while data[pos:pos + 1] not in (b" ", b"\n", b"\t", b"\r"):
            pos += 1
        words.append(data[start:pos].decode())
    pos += 1  # the single whitespace after the scale word
    ch = {"PF": 3, "Pf": 1}.get(words[0])
    if ch is None:
        raise SceneError(f"{path}: not a PFM file")
    w, h, scale = int(words[1]), int(words[2]), float(words[3])
    arr = np.frombuffer(data, dtype="<f4" if scale < 0 else ">f4", count=w * h * ch, offset=pos)
    img = arr.reshape(h, w, ch)[::-1].astype(np.float32) * np.float32(abs(scale))
    if ch == 1:
        img = np.repeat(img, 3, axis=2)
    return np.ascontiguousarray(img)


def read_hdr(path):
    """Radiance RGBE (.hdr) as stbi_loadf reads it: '-Y h +X w' rows top to
    bottom, flat or new-style run-length scanlines, value = m * 2^(e-136)."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    lines = []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode(errors="replace")
        pos = end + 1
        if line.strip() == "" and lines:
            break
        lines.append(line)
    if not lines[0].startswith("#?"):
        raise SceneError(f"{path}: not a Radiance file")
    end = data.index(b"\n", pos)
    dims = data[pos:end].decode().split()
    pos = end + 1
    if len(dims) != 4 or dims[0] != "-Y" or dims[2] != "+X":
        raise SceneError(f"{path}: only -Y h +X w orientation is read")
    h, w = int(dims[1]), int(dims[3])
    rgbe = np.zeros((h, w, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(h):
        if w >= 8 and w < 32768 and buf[pos] == 2 and buf[pos + 1] == 2 and not (buf[pos + 2] & 0x80):
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = int(buf[pos])
                    pos += 1
                    if n > 128:
                        n -= 128
                        rgbe[y, x:x + n, c] = buf[pos]
                        pos += 1
                    else:
                        rgbe[y, x:x + n, c] = buf[pos:pos + n]
                        pos += n
                    x += n
        else:
            rgbe[y] = buf[pos:pos + 4 * w].reshape(w, 4)
            pos += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1), e - 136), 0).astype(np.float32)
    return np.ascontiguousarray(rgbe[..., :3].astype(np.float32) * f[..., None])


# ------------------------------------------------------------------ nodes
def _texture_constant(node, what):
    if not isinstance(node, dict) or "type" not in node:
        raise SceneError(f"{what}: texture node with a type expected")
    if node["type"] != "constant":
        raise SceneError(f"{what}: texture type '{node['type']}' is not supported on the GPU path "
                         "(constant textures only)")
    t = node.get("texel")
    if isinstance(t, (list, tuple)):
        return _vec(t, 3, what=f"{what}.texel")
    return (float(t),) * 3


def _material(node):
    t = node.get("type")
    if t == "diffuse":
        return Material(MAT_DIFFUSE, _texture_constant(node.get("albedo"), "diffuse.albedo"))
    if t == "disney":
        keys = ["subsurface", "metallic", "specular", "specular_tint", "roughness", "anisotropic",
                "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"]
        vals = {k: _texture_constant(node.get(k), f"disney.{k}")[0] for k in keys}
        return Material(MAT_DISNEY, _texture_constant(node.get("base_color"), "disney.base_color"),
                        **vals)
    raise SceneError(f"material type '{t}' is not supported on the GPU path (diffuse, disney)")


def _geometry(node, material, base_dir, two_sided=False):
    t = node.get("type")
    if t == "twosided":
        return _geometry(node["internal"], material, base_dir, True)
    m = _transform(node.get("transform"))
    if t == "quad":
        pts = [_vec(node.get(k), 3, d, f"quad.{k}") for k, d in
               (("a", (-1, -1, 0)), ("b", (1, -1, 0)), ("c", (1, 1, 0)), ("d", (-1, 1, 0)))]
        p = _xform(m, np.asarray(pts, np.float32), 1)
        return scenes.quad(*[tuple(float(x) for x in q) for q in p], material, two_sided)
    if t == "cube":
        # Cube (cube.cpp:31-75): the unit cube's 24 face corners through
        # transformPoint; index tuples as scenes.cube stores them
        c = scenes.cube(np.eye(4), material, two_sided)
        pos = _xform(m, scenes._CUBE_FACES.reshape(-1, 3).astype(np.float32), 1)
        return dataclasses.replace(c, positions=np.ascontiguousarray(pos, np.float32))
    if t == "mesh":
        fn = node["filename"]
        path = fn if os.path.isabs(fn) or os.path.exists(fn) else os.path.join(base_dir, fn)
        return mesh_from_obj(path, m, material, two_sided)
    raise SceneError(f"geometry type '{t}' is not supported on the GPU path (quad, cube, mesh, twosided)")


def _env_texels(node, base_dir):
    if node.get("type") == "constant":
        return np.asarray(_texture_constant(node, "envmap.texture"), np.float32).reshape(1, 1, 3)
    if node.get("type") == "hdr":
        fn = node["filename"]
        path = fn if os.path.isabs(fn) or os.path.exists(fn) else os.path.join(base_dir, fn)
        return read_pfm(path) if path.endswith(".pfm") else read_hdr(path)
    raise SceneError(f"envmap texture type '{node.get('type')}' is not supported (constant, hdr)")


def _light(node, base_dir):
    t = node.get("type")
    if t == "directional":
        d = _vec(node.get("direction"), 3, what="directional.direction")
        return Light(LIGHT_DIRECTIONAL, dir=d, angle=math.radians(float(node.get("angle", 0.0))),
                     radiance=_vec(node.get("emission"), 3, what="directional.emission"))
    if t == "envmap":
        return Light(LIGHT_ENV, env=_env_texels(node["texture"], base_dir),
                     rot=float(node.get("rot", 0.0)), scaler=float(node.get("scaler", 1.0)))
    raise SceneError(f"light type '{t}' is not supported on the GPU path (directional, envmap)")


def load(source, base_dir: Optional[str] = None) -> LoadedScene:
    """Load an Ajisai scene file (path) or an already parsed dict."""
    if isinstance(source, (str, os.PathLike)):
        base_dir = base_dir or os.path.dirname(os.path.abspath(source))
        with open(source) as f:
            cfg = yaml.safe_load(f)
    else:
        cfg = source
    base_dir = base_dir or "."
    sc_node, cam_node, r_node = cfg.get("scene"), cfg.get("camera"), cfg.get("renderer")
    if sc_node is None or cam_node is None or r_node is None:
        raise SceneError("scene file needs 'scene', 'camera' and 'renderer'")
    prims, mats = [], []
    for p in sc_node.get("primitives", []) or []:
        if p.get("type") != "geometry":
            raise SceneError(f"primitive type '{p.get('type')}' is not supported (geometry)")
        emission = tuple(float(e) for e in _vec(p.get("emission"), 3, (0, 0, 0),
                                                  "primitive.emission"))
        if "med_in" in p or "med_out" in p:
            raise SceneError("participating media are not supported on the GPU path")
        if "material" not in p:
            raise SceneError("a primitive without material (null material) is not supported")
        mats.append(_material(p["material"]))
        prims.append(dataclasses.replace(_geometry(p["geometry"], len(mats) - 1, base_dir),
                                         emission=emission))
    lights = [_light(l, base_dir) for l in (sc_node.get("lights", []) or [])]
    agg = sc_node.get("aggregate") or {"type": "native"}
    if agg.get("type") not in ("ploc", "native", "embree"):
        raise SceneError(f"aggregate '{agg.get('type')}' is not supported (ploc, native, embree)")
    radius = int(agg.get("search_radius", 14)) if agg.get("type") == "ploc" else 14
    if cam_node.get("type") != "thin_lens":
        raise SceneError(f"camera type '{cam_node.get('type')}' is not supported (thin_lens)")
    res = _vec(cam_node.get("resolution"), 2, what="camera.resolution")
    cam = CameraSpec(pos=_vec(cam_node.get("position"), 3, what="camera.position"),
                     look_at=_vec(cam_node.get("look_at"), 3, what="camera.look_at"),
                     up=_vec(cam_node.get("up"), 3, what="camera.up"),
                     fov=float(cam_node["fov"]), res=(int(res[0]), int(res[1])),
                     lens_radius=float(cam_node.get("lens_radius", 0.0)),
                     focal_distance=float(cam_node.get("focal_distance", 1.0)))
    fnode = cam_node.get("filter") or {"type": "box"}
    if fnode.get("type") == "box":
        filt = FilterSpec("box", (float(fnode.get("xwidth", 0.5)), float(fnode.get("ywidth", 0.5))))
    elif fnode.get("type") == "gaussian":
        filt = FilterSpec("gaussian", (float(fnode.get("xwidth", 2.0)), float(fnode.get("ywidth", 2.0))),
                          float(fnode.get("alpha", 2.0)))
    else:
        raise SceneError(f"filter '{fnode.get('type')}' is not supported (box, gaussian)")
    rtype = r_node.get("type")
    sppm = vcm = None
    use_mis = True
    if rtype == "pt":
        use_mis = bool(r_node.get("use_mis", True))
        pt = PTSpec(spp=int(r_node["spp"]), tile_size=int(r_node.get("tile_size", 16)),
                    min_bounces=int(r_node.get("min_bounces", 5)),
                    max_bounces=int(r_node.get("max_bounces", 10)),
                    cont_prob=float(r_node.get("cont_prob", 0.9)), use_mis=use_mis)
    elif rtype == "sppm":  # SPPMCreatorImpl (renderer_creators.cpp:56-90)
        pt = PTSpec()
        sppm = scenes.SPPMSpec(
            forward_max_depth=int(r_node.get("forward_max_depth", 8)),
            init_radius=float(r_node.get("init_radius", 0.1)),
            iteration_count=int(r_node.get("iteration_count", 100)),
            photons_per_iteration=int(r_node.get("photons_per_iteration", 100000)),
            photon_min_depth=int(r_node.get("photon_min_depth", 5)),
            photon_max_depth=int(r_node.get("photon_max_depth", 10)),
            photon_cont_prob=float(r_node.get("photon_cont_prob", 0.9)),
            update_alpha=float(r_node.get("update_alpha", 2.0 / 3)))
    elif rtype == "vcm":  # VCMCreatorImpl (renderer_creators.cpp:105-135)
        pt = PTSpec()
        vcm = scenes.VCMSpec(
            iteration=int(r_node.get("iteration", 1)),
            base_radius=float(r_node.get("radius", 0.003)),
            radius_alpha=float(r_node.get("alpha", 0.75)),
            use_vm=bool(r_node.get("use_vm", True)), use_vc=bool(r_node.get("use_vc", True)),
            min_path_length=int(r_node.get("min_path_length", 0)),
            max_path_length=int(r_node.get("max_path_length", 10)),
            cont_prob=float(r_node.get("cont_prob", 0.9)))
    else:
        raise SceneError(f"renderer '{rtype}' is not supported (pt, sppm, vcm)")
    scene = Scene("yaml", prims, mats, lights, cam, pt, search_radius=radius)
    return LoadedScene(scene, filt, use_mis, rtype, sppm, vcm, agg.get("type"))
