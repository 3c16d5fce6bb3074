"""Ajisai YAML scene files (ajisai_render_amd.loader): keys and defaults of
the reference's creators, OBJ re-indexing, PFM/HDR readers, and rejection of
what the GPU path does not render."""
import os

import numpy as np
import pytest
import yaml

from ajisai_render_amd import loader, scenes


def _cornell_yaml(tmp, obj=True, env="constant", lamp=False):
    """A Cornell-style scene in the reference's file format; `lamp` adds an
    emissive one-sided ceiling quad (an area light)."""
    mat = lambda c: {"type": "diffuse", "albedo": {"type": "constant", "texel": list(c)}}  # noqa: E731
    quad = lambda a, b, c, d, col: {"type": "geometry", "material": mat(col),  # noqa: E731
                                    "geometry": {"type": "twosided", "internal":
                                                 {"type": "quad", "a": a, "b": b, "c": c, "d": d}}}
    prims = [quad([0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1], [0.73] * 3),
             quad([0, 1, 0], [0, 1, 1], [1, 1, 1], [1, 1, 0], [0.73] * 3),
             {"type": "geometry", "material": {"type": "disney",
                                               **{k: {"type": "constant", "texel": v} for k, v in
                                                  [("base_color", [0.8, 0.3, 0.2]), ("subsurface", 0),
                                                   ("metallic", 0.2), ("specular", 0.5),
                                                   ("specular_tint", 0), ("roughness", 0.4),
                                                   ("anisotropic", 0), ("sheen", 0), ("sheen_tint", 0),
                                                   ("clearcoat", 0.3), ("clearcoat_gloss", 0.5)]}},
              "geometry": {"type": "cube", "transform": {"type": "matrix", "matrix":
                           [0.3, 0, 0, 0.35, 0, 0.3, 0, 0, 0, 0, 0.3, 0.35, 0, 0, 0, 1]}}}]
    if obj:
        with open(os.path.join(tmp, "tet.obj"), "w") as f:
            f.write("v 0.2 0.0 0.2\nv 0.5 0.0 0.2\nv 0.35 0.0 0.5\nv 0.35 0.4 0.3\n"
                    "f 1 2 3\nf 1 2 4\nf 2 3 4\nf 3 1 4\n")
        prims.append({"type": "geometry", "material": mat([0.1, 0.6, 0.2]),
                      "geometry": {"type": "mesh", "filename": "tet.obj",
                                   "transform": [{"type": "matrix", "matrix":
                                                  [1, 0, 0, 0.1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]}]}})
    if lamp:
        prims.append({"type": "geometry", "material": mat([0.5, 0.5, 0.5]), "emission": [9, 8, 6],
                      "geometry": {"type": "quad", "a": [0.35, 0.99, 0.35], "b": [0.65, 0.99, 0.35],
                                   "c": [0.65, 0.99, 0.65], "d": [0.35, 0.99, 0.65]}})
    lights = [{"type": "directional", "direction": [0.3, 0.8, 0.5], "emission": [5, 5, 5]}]
    if env == "constant":
        lights.append({"type": "envmap", "texture": {"type": "constant", "texel": 0.1}})
    else:
        lights.append({"type": "envmap", "texture": {"type": "hdr", "filename": env}, "rot": 30,
                       "scaler": 2})
    cfg = {"scene": {"type": "default", "primitives": prims, "lights": lights,
                     "aggregate": {"type": "ploc", "search_radius": 10}},
           "camera": {"type": "thin_lens", "resolution": [48, 32], "position": [0.5, 0.5, 2.2],
                      "look_at": [0.5, 0.5, 0], "up": [0, 1, 0], "fov": 40,
                      "filter": {"type": "gaussian"}},
           "renderer": {"type": "pt", "spp": 4}}
    path = os.path.join(tmp, "scene.yaml")
    with open(path, "w") as f:
        yaml.safe_dump(cfg, f)
    return path


def test_keys_defaults_and_geometry(tmp_path):
    ls = loader.load(_cornell_yaml(str(tmp_path)))
    sc = ls.scene
    assert ls.filter.type == "gaussian" and ls.filter.radius == (2.0, 2.0) and ls.filter.alpha == 2.0
    assert (sc.pt.spp, sc.pt.tile_size, sc.pt.min_bounces, sc.pt.max_bounces, sc.pt.cont_prob) == \
        (4, 16, 5, 10, 0.9)
    assert sc.search_radius == 10 and sc.camera.res == (48, 32) and sc.camera.lens_radius == 0.0
    assert [p.kind for p in sc.primitives] == [scenes.GEOM_QUAD, scenes.GEOM_QUAD, scenes.GEOM_CUBE,
                                               scenes.GEOM_MESH]
    assert sc.primitives[0].two_sided and not sc.primitives[2].two_sided
    # the cube: the reference's cube (centred, side 1) scaled by 0.3, moved to (0.35, 0, 0.35)
    cpos = sc.primitives[2].positions
    assert np.allclose(cpos.min(0), [0.2, -0.15, 0.2]) and np.allclose(cpos.max(0), [0.5, 0.15, 0.5])
    # the mesh: 4 faces, no normals in the file -> one vertex per (v, n, t)
    # first use, normals from the first face that used the vertex
    m = sc.primitives[3]
    assert m.index_words.shape == (4, 3) and m.positions.shape[0] == 4
    assert np.allclose(m.positions[0, :3], [0.3, 0.0, 0.2])
    n = m.normal_and_vs[:, :3]
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    assert np.allclose(n[0], [0, -1, 0], atol=1e-6) or np.allclose(n[0], [0, 1, 0], atol=1e-6)
    assert sc.materials[2].type == scenes.MAT_DISNEY and sc.materials[2].clearcoat == 0.3
    assert sc.lights[0].type == scenes.LIGHT_DIRECTIONAL and sc.lights[1].env.shape == (1, 1, 3)


def test_obj_reindexing_with_normals_and_texcoords(tmp_path):
    p = os.path.join(str(tmp_path), "q.obj")
    with open(p, "w") as f:
        f.write("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                "vn 0 0 1\nvn 0 0 -1\nf 1/1/1 2/2/1 3/3/1 4/4/1\nf 1/1/2 3/3/2 2/2/2\n")
    mesh = loader.mesh_from_obj(p, np.eye(4, dtype=np.float32), 0)
    # quad fan-triangulated (2 faces) + a back face reusing positions with
    # another normal: 4 + 3 distinct (v, n, t) triples
    assert mesh.index_words.tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6]]
    assert mesh.positions.shape[0] == 7
    assert np.allclose(mesh.normal_and_vs[4, :3], [0, 0, -1])
    assert np.allclose(mesh.positions[2, 3], 1) and np.allclose(mesh.normal_and_vs[3, 3], 1)  # u, v


def test_pfm_and_hdr_readers(tmp_path):
    rng = np.random.default_rng(3)
    img = (rng.random((5, 7, 3)) * 4).astype(np.float32)
    p = os.path.join(str(tmp_path), "e.pfm")
    with open(p, "wb") as f:
        f.write(b"PF\n7 5\n-1.0\n")
        f.write(img[::-1].astype("<f4").tobytes())     # rows bottom to top
    assert np.array_equal(loader.read_pfm(p), img)
    # Radiance: a flat scanline file and a run-length file of the same image
    vals = (rng.random((3, 40, 3)) * 10).astype(np.float32)
    m, e = np.frexp(vals.max(axis=2))
    rgbe = np.zeros((3, 40, 4), np.uint8)
    rgbe[..., :3] = np.clip(np.floor(vals / np.ldexp(1.0, e)[..., None] * 256), 0, 255)
    rgbe[..., 3] = e + 128
    want = rgbe[..., :3].astype(np.float32) * np.ldexp(np.float32(1), rgbe[..., 3].astype(np.int32) - 136)[..., None]
    hdr = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 3 +X 40\n"
    flat = os.path.join(str(tmp_path), "f.hdr")
    with open(flat, "wb") as f:
        f.write(hdr + rgbe.tobytes())
    assert np.allclose(loader.read_hdr(flat), want)
    rle = os.path.join(str(tmp_path), "r.hdr")
    with open(rle, "wb") as f:
        f.write(hdr)
        for y in range(3):
            f.write(bytes([2, 2, 0, 40]))
            for c in range(4):
                row = rgbe[y, :, c]
                f.write(bytes([128 + 10, row[0]]))            # a run of 10 copies of row[0]
                f.write(bytes([30]) + row[10:].tobytes())      # 30 literals
    got = loader.read_hdr(rle)
    exp = want.copy()
    exp[:, :10] = want[:, :1]
    assert np.allclose(got, exp)


def test_unsupported_inputs_are_named(tmp_path):
    cfg = yaml.safe_load(open(_cornell_yaml(str(tmp_path), obj=False)))
    bad = [("scene", "primitives", 0, "material", {"type": "glass"}),
           ("scene", "primitives", 0, "geometry", {"type": "sphere", "radius": 1, "center": [0, 0, 0]})]
    for path in bad:
        c = yaml.safe_load(yaml.safe_dump(cfg))
        node = c
        for k in path[:-2]:
            node = node[k]
        node[path[-2]] = path[-1]
        with pytest.raises(loader.SceneError):
            loader.load(c)

    # use_mis: false selects PathTracing::Trace (renderer_creators.cpp:38, pt.cpp:85-87)
    c = yaml.safe_load(yaml.safe_dump(cfg))
    c["renderer"]["use_mis"] = False
    assert loader.load(c).scene.pt.use_mis is False
    assert loader.load(cfg).scene.pt.use_mis is True


def test_emissive_primitive_is_an_area_light(tmp_path):
    """"emission" on a geometry primitive (primitive_creators.cpp:65-72) makes
    it an area light; the primitive keeps its geometry and material."""
    sc = loader.load(_cornell_yaml(str(tmp_path), obj=False, lamp=True)).scene
    lamp = sc.primitives[-1]
    assert lamp.emission == (9.0, 8.0, 6.0) and lamp.kind == scenes.GEOM_QUAD
    assert not lamp.two_sided
    assert all(max(p.emission) == 0 for p in sc.primitives[:-1])


def test_sppm_vcm_renderers_and_aggregates(tmp_path):
    """"sppm" and "vcm" renderer nodes with the creators' defaults and keys;
    "native" / "embree" / no aggregate load onto the GPU PLOC tree."""
    cfg = yaml.safe_load(open(_cornell_yaml(str(tmp_path), obj=False)))
    c = yaml.safe_load(yaml.safe_dump(cfg))
    c["renderer"] = {"type": "sppm", "photons_per_iteration": 5000}
    ls = loader.load(c)
    assert ls.renderer == "sppm" and ls.sppm == scenes.SPPMSpec(photons_per_iteration=5000)
    c["renderer"] = {"type": "vcm", "radius": 0.01, "use_vm": False, "max_path_length": 6}
    ls = loader.load(c)
    assert ls.renderer == "vcm" and ls.vcm == scenes.VCMSpec(base_radius=0.01, use_vm=False,
                                                               max_path_length=6)
    for agg in ({"type": "native"}, {"type": "embree"}, None):
        c2 = yaml.safe_load(yaml.safe_dump(cfg))
        if agg is None:
            del c2["scene"]["aggregate"]
        else:
            c2["scene"]["aggregate"] = agg
        ls = loader.load(c2)
        assert ls.aggregate == (agg["type"] if agg else "native")
        assert ls.scene.search_radius == 14
    c["renderer"] = {"type": "bdpt"}
    with pytest.raises(loader.SceneError):
        loader.load(c)
