"""ctypes wrapper of the CPU oracle (liborc.so) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or as the timed CPU baseline. The oracle
restates the reference (see orc.h for the per-function citations and how each
part is pinned).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

NODE_DTYPE = np.dtype([("mn", "<f4", 3), ("mx", "<f4", 3), ("count", "<u4"), ("first", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("id", "<u4"), ("u", "<f4"), ("v", "<f4")])


class MeshView(C.Structure):
    _fields_ = [("position_buffer", C.c_void_p), ("position_offset", C.c_size_t),
                ("position_stride", C.c_size_t), ("position_size", C.c_size_t),
                ("index_buffer", C.c_void_p), ("index_offset", C.c_size_t),
                ("index_stride", C.c_size_t), ("index_size", C.c_size_t)]


class Prim(C.Structure):
    _fields_ = [("view", MeshView), ("kind", C.c_uint32), ("two_sided", C.c_uint32),
                ("material", C.c_uint32), ("pad", C.c_uint32), ("normal_and_vs", C.c_void_p),
                ("emission", C.c_float * 3), ("pad2", C.c_uint32)]


class Mat(C.Structure):
    _fields_ = [("type", C.c_uint32), ("base_color", C.c_float * 3),
                ("subsurface", C.c_float), ("metallic", C.c_float), ("specular", C.c_float),
                ("specular_tint", C.c_float), ("roughness", C.c_float),
                ("anisotropic", C.c_float), ("sheen", C.c_float), ("sheen_tint", C.c_float),
                ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float)]


class Lt(C.Structure):
    _fields_ = [("type", C.c_uint32), ("dir", C.c_float * 3), ("angle", C.c_float),
                ("radiance", C.c_float * 3), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("env_texels", C.c_void_p), ("rot", C.c_float), ("scaler", C.c_float)]


class Cam(C.Structure):
    _fields_ = [("raster2camera", C.c_float * 16), ("camera2world", C.c_float * 16),
                ("resolution", C.c_float * 2), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float)]


class Args(C.Structure):
    _fields_ = [("spp", C.c_int32), ("tile_size", C.c_int32), ("min_bounces", C.c_int32),
                ("max_bounces", C.c_int32), ("cont_prob", C.c_float), ("use_mis", C.c_int32),
                ("seed_mode", C.c_int32), ("seed_base", C.c_uint32),
                ("env_cdf_fp64", C.c_int32), ("pad", C.c_int32)]


class SppmArgs(C.Structure):
    _fields_ = [("forward_max_depth", C.c_int32), ("init_radius", C.c_float),
                ("iteration_count", C.c_int32), ("photons_per_iteration", C.c_int32),
                ("photon_min_depth", C.c_int32), ("photon_max_depth", C.c_int32),
                ("photon_cont_prob", C.c_float), ("update_alpha", C.c_float),
                ("photon_seed_base", C.c_uint32), ("pad", C.c_uint32)]


VCM_PIXEL_UNITS, VCM_REVERSE_PDF, VCM_LT_ONLY = 1, 2, 4


class VcmArgs(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("base_radius", C.c_float),
                ("radius_alpha", C.c_float), ("use_vm", C.c_int32), ("use_vc", C.c_int32),
                ("min_path_length", C.c_int32), ("max_path_length", C.c_int32),
                ("cont_prob", C.c_float), ("seed_base", C.c_uint32), ("variant", C.c_uint32)]


SPPM_PIXEL_DTYPE = np.dtype([("pos", "<f4", 3), ("beta", "<f4", 3), ("direct", "<f4", 3),
                             ("tau", "<f4", 3), ("radius", "<f4"), ("n", "<f4"),
                             ("valid", "<u4"), ("pad", "<u4")])

_P = C.c_void_p
_lib = None


def build():
    """Compile liborc.so with the committed Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    sig = {
        "orc_morton_split": ([C.c_uint32], C.c_uint32),
        "orc_morton_encode": ([C.c_uint32] * 3, C.c_uint32),
        "orc_flatten": ([_P, C.c_uint32, _P, _P], C.c_uint32),
        "orc_morton_codes": ([_P, C.c_uint32, _P, _P], None),
        "orc_radix_sort": ([_P, _P, C.c_uint32, C.c_uint32], None),
        "orc_ploc_build": ([_P, C.c_uint32, _P, C.c_uint32, _P, _P], C.c_int),
        "orc_ploc_nearest": ([_P, C.c_uint32, C.c_uint32, C.c_uint32, _P], None),
        "orc_trace_closest": ([_P, _P, _P, _P, C.c_uint32, _P], None),
        "orc_trace_any": ([_P, _P, _P, _P, C.c_uint32, _P], None),
        "orc_brute_closest": ([_P, C.c_uint32, _P, C.c_uint32, _P], None),
        "orc_trace_closest_2l": ([_P, _P, _P, _P, _P, _P, C.c_uint32, _P, _P], None),
        "orc_ggx_parts": ([C.c_uint32, _P, _P, C.c_float, _P, _P, _P], None),
        "orc_ggx_sample_vn_batch": ([C.c_uint32, _P, _P, C.c_float, _P, _P], None),
        "orc_fresnel_conductor": ([C.c_float, C.c_float, C.c_float], C.c_float),
        "orc_trace_any_2l": ([_P, _P, _P, _P, _P, _P, C.c_uint32, _P], None),
        "orc_trace_closest_counts": ([_P, _P, _P, _P, C.c_uint32, _P, _P], None),
        "orc_pcg32_floats": ([C.c_uint64, C.c_uint32, _P], None),
        "orc_scene_create": ([_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32,
                              C.c_int32], _P),
        "orc_scene_destroy": ([_P], None),
        "orc_scene_tri_count": ([_P], C.c_uint32),
        "orc_scene_nodes": ([_P], _P),
        "orc_scene_perm": ([_P], _P),
        "orc_scene_tris": ([_P], _P),
        "orc_scene_world_bound": ([_P, _P], None),
        "orc_scene_env_tables": ([_P, C.c_uint32, _P, _P, _P, _P], C.c_int),
        "orc_env_sample": ([_P, C.c_uint32, C.c_float, C.c_float, _P, _P, _P], None),
        "orc_env_pdf": ([_P, C.c_uint32, _P], C.c_float),
        "orc_env_radiance": ([_P, C.c_uint32, _P, _P], None),
        "orc_bsdf_sample": ([_P, _P, _P, _P, _P, _P, _P, _P, _P], None),
        "orc_bsdf_sample_batch": ([_P, _P, _P, _P, C.c_uint32, _P, _P, _P], None),
        "orc_bsdf_eval": ([_P, _P, _P, _P, _P, _P], None),
        "orc_bsdf_pdf": ([_P, _P, _P, _P, _P], C.c_float),
        "orc_square_to_disk": ([_P, _P], None),
        "orc_cosine_hemisphere": ([_P, _P, _P], None),
        "orc_frame": ([_P, _P, _P], None),
        "orc_camera_ray": ([_P, C.c_float, C.c_float, _P, _P], None),
        "orc_camera_setup": ([_P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_float], None),
        "orc_render": ([_P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, _P], C.c_uint64),
        "orc_render_mt": ([_P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_int],
                          C.c_uint64),
        "orc_film_resolve": ([C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P], None),
        "orc_film_resolve_filter": ([C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_float,
                                     C.c_float, C.c_float, _P, _P, _P], None),
        "orc_image_srgb8": ([C.c_uint32, C.c_uint32, _P, C.c_int, _P], None),
        "orc_render_sppm": ([_P, _P, _P, C.c_int32, _P, _P], C.c_int64),
        "orc_render_sppm_mt": ([_P, _P, _P, C.c_int32, _P, _P, C.c_int], C.c_int64),
        "orc_render_vcm_mt": ([_P, _P, _P, _P, _P, C.c_int], C.c_int64),
        "orc_sppm_grid_probe": ([_P, C.c_float, C.c_uint32, _P, _P, C.c_uint32, _P, _P], None),
        "orc_light_emit_batch": ([_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P,
                                  _P], None),
        "orc_sppm_update_probe": ([C.c_float, C.c_float, _P, C.c_uint32, _P, _P, _P], None),
        "orc_render_vcm": ([_P, _P, _P, _P, _P], C.c_int64),
        "orc_vcm_to_raster": ([_P, _P, _P, _P, _P], C.c_int),
        "orc_vcm_image_to_solid": ([_P, _P], C.c_float),
    }
    for name, (a, r) in sig.items():
        f = getattr(lib, name)
        f.argtypes = a
        f.restype = r
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data


# ----------------------------------------------------------------- PLOC
def ploc_build(tri: np.ndarray, scene_box: np.ndarray, radius: int = 14):
    lib = load()
    tri = np.ascontiguousarray(tri, np.float32)
    n = tri.shape[0]
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    perm = np.zeros(n, np.uint32)
    box = np.ascontiguousarray(scene_box, np.float32)
    it = lib.orc_ploc_build(_p(tri), n, _p(box), radius, _p(nodes), _p(perm))
    return nodes, perm, it


def morton_codes(tri, scene_box):
    lib = load()
    tri = np.ascontiguousarray(tri, np.float32)
    out = np.zeros(tri.shape[0], np.uint32)
    box = np.ascontiguousarray(scene_box, np.float32)
    lib.orc_morton_codes(_p(tri), tri.shape[0], _p(box), _p(out))
    return out


def radix_sort(keys, vals, bits=30):
    lib = load()
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib.orc_radix_sort(_p(k), _p(v), k.shape[0], bits)
    return k, v


def nearest(nodes, begin, end, radius):
    lib = load()
    nodes = np.ascontiguousarray(nodes)
    out = np.zeros(nodes.shape[0], np.uint32)
    lib.orc_ploc_nearest(_p(nodes), begin, end, radius, _p(out))
    return out


def trace_closest(nodes, perm, tri, rays):
    lib = load()
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    lib.orc_trace_closest(_p(nodes), _p(perm), _p(np.ascontiguousarray(tri, np.float32)),
                          _p(rays), rays.shape[0], _p(hits))
    return hits


def trace_any(nodes, perm, tri, rays):
    lib = load()
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    occ = np.zeros(rays.shape[0], np.uint8)
    lib.orc_trace_any(_p(nodes), _p(perm), _p(np.ascontiguousarray(tri, np.float32)),
                      _p(rays), rays.shape[0], _p(occ))
    return occ


class _Blas(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("perm", C.c_void_p), ("tri", C.c_void_p)]


def trace_2l(tlas_nodes, tlas_perm, w2o, inst_mesh, blases, rays, closest=True):
    """Two-level traversal. blases: list of (nodes, perm, tri) per mesh;
    w2o: (n_inst, 12) float32; inst_mesh: (n_inst,) uint32. Returns
    (hits, instance) for closest, occluded flags otherwise."""
    lib = load()
    keep = [(np.ascontiguousarray(nd), np.ascontiguousarray(pm, np.uint32),
             np.ascontiguousarray(t, np.float32)) for nd, pm, t in blases]
    arr = (_Blas * len(keep))()
    for i, (nd, pm, t) in enumerate(keep):
        arr[i].nodes, arr[i].perm, arr[i].tri = _p(nd), _p(pm), _p(t)
    tn = np.ascontiguousarray(tlas_nodes)
    tp = np.ascontiguousarray(tlas_perm, np.uint32)
    m = np.ascontiguousarray(w2o, np.float32)
    im = np.ascontiguousarray(inst_mesh, np.uint32)
    rays = np.ascontiguousarray(rays)
    n = rays.shape[0]
    if closest:
        hits = np.zeros(n, HIT_DTYPE)
        inst = np.zeros(n, np.uint32)
        lib.orc_trace_closest_2l(_p(tn), _p(tp), _p(m), _p(im), C.addressof(arr), _p(rays), n,
                                 _p(hits), _p(inst))
        return hits, inst
    occ = np.zeros(n, np.uint8)
    lib.orc_trace_any_2l(_p(tn), _p(tp), _p(m), _p(im), C.addressof(arr), _p(rays), n, _p(occ))
    return occ


def ggx_parts(v, h, alpha):
    """(D(h), Smith G1(v, h), visible-normal pdf of h given v), shading frame."""
    lib = load()
    v = np.ascontiguousarray(v, np.float32)
    h = np.ascontiguousarray(h, np.float32)
    n = h.shape[0]
    d, g1, pv = (np.zeros(n, np.float32) for _ in range(3))
    lib.orc_ggx_parts(n, _p(v), _p(h), float(alpha), _p(d), _p(g1), _p(pv))
    return d, g1, pv


def ggx_sample_vn(v, u, alpha):
    lib = load()
    v = np.ascontiguousarray(v, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    n = v.shape[0]
    h = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib.orc_ggx_sample_vn_batch(n, _p(v), _p(u), float(alpha), _p(h), _p(pdf))
    return h, pdf


def brute_closest(tri, rays):
    lib = load()
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    tri = np.ascontiguousarray(tri, np.float32)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    lib.orc_brute_closest(_p(tri), tri.shape[0], _p(rays), rays.shape[0], _p(hits))
    return hits


def trace_counts(nodes, perm, tri, rays):
    lib = load()
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    nv = np.zeros(rays.shape[0], np.uint32)
    nt = np.zeros(rays.shape[0], np.uint32)
    lib.orc_trace_closest_counts(_p(nodes), _p(perm), _p(np.ascontiguousarray(tri, np.float32)),
                                 _p(rays), rays.shape[0], _p(nv), _p(nt))
    return nv, nt


def pcg32_floats(seed, n):
    lib = load()
    out = np.zeros(n, np.float32)
    lib.orc_pcg32_floats(seed, n, _p(out))
    return out


# ----------------------------------------------------------------- scene
class OracleScene:
    """orc_scene built from an ajisai_render_amd.scenes.Scene."""

    def __init__(self, scene, env_cdf_fp64=1):
        lib = load()
        self.lib = lib
        self.keep = []
        prims = (Prim * max(1, len(scene.primitives)))()
        for i, p in enumerate(scene.primitives):
            pos = np.ascontiguousarray(p.positions, np.float32)
            idx = np.ascontiguousarray(p.index_words, np.uint32)
            nrm = None if p.normal_and_vs is None else np.ascontiguousarray(p.normal_and_vs,
                                                                             np.float32)
            self.keep += [pos, idx, nrm]
            mv = MeshView(pos.ctypes.data, 0, pos.strides[0], pos.shape[0], idx.ctypes.data, 0,
                          12, idx.shape[0])
            prims[i] = Prim(mv, p.kind, int(p.two_sided), p.material, 0,
                            nrm.ctypes.data if nrm is not None else None,
                            (C.c_float * 3)(*getattr(p, "emission", (0.0, 0.0, 0.0))), 0)
        mats = (Mat * max(1, len(scene.materials)))()
        for i, m in enumerate(scene.materials):
            mats[i] = Mat(m.type, (C.c_float * 3)(*m.base_color), m.subsurface, m.metallic,
                          m.specular, m.specular_tint, m.roughness, m.anisotropic, m.sheen,
                          m.sheen_tint, m.clearcoat, m.clearcoat_gloss)
        lts = (Lt * max(1, len(scene.lights)))()
        for i, l in enumerate(scene.lights):
            env = None if l.env is None else np.ascontiguousarray(l.env, np.float32)
            self.keep.append(env)
            h, w = (env.shape[0], env.shape[1]) if env is not None else (0, 0)
            lts[i] = Lt(l.type, (C.c_float * 3)(*l.dir), l.angle, (C.c_float * 3)(*l.radiance),
                        w, h, env.ctypes.data if env is not None else None, l.rot, l.scaler)
        self.keep += [prims, mats, lts]
        self.h = lib.orc_scene_create(prims, len(scene.primitives), mats, len(scene.materials),
                                      lts, len(scene.lights), scene.search_radius,
                                      env_cdf_fp64)
        if not self.h:
            raise RuntimeError("orc_scene_create failed")
        self.scene = scene
        self.n_tri = lib.orc_scene_tri_count(self.h)

    def nodes(self):
        p = self.lib.orc_scene_nodes(self.h)
        return np.ctypeslib.as_array((C.c_uint8 * (32 * (2 * self.n_tri - 1))).from_address(p)
                                     ).view(NODE_DTYPE).copy()

    def perm(self):
        p = self.lib.orc_scene_perm(self.h)
        return np.ctypeslib.as_array((C.c_uint32 * self.n_tri).from_address(p)).copy()

    def tris(self):
        p = self.lib.orc_scene_tris(self.h)
        return np.ctypeslib.as_array((C.c_float * (9 * self.n_tri)).from_address(p)
                                     ).reshape(-1, 9).copy()

    def env_tables(self, li):
        l = self.scene.lights[li]
        h, w = l.env.shape[:2]
        cc = np.zeros((h, w + 1), np.float32)
        cf = np.zeros(h, np.float32)
        mc = np.zeros(h + 1, np.float32)
        mf = np.zeros(1, np.float32)
        assert self.lib.orc_scene_env_tables(self.h, li, _p(cc), _p(cf), _p(mc), _p(mf)) == 0
        return cc, cf, mc, mf[0]

    def env_sample(self, li, u0, u1):
        d = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        rgb = np.zeros(3, np.float32)
        self.lib.orc_env_sample(self.h, li, u0, u1, _p(d), _p(pdf), _p(rgb))
        return d, float(pdf[0]), rgb

    def light_emit(self, li, n, seed=0):
        """n draws of Light::SampleEmit from light li (emissive primitives'
        lights first, then the scene's): dict of pos, dir, normal, radiance
        (n x 3), pdf_pos, pdf_dir (n)."""
        out = {k: np.zeros((n, 3), np.float32) for k in ("pos", "dir", "normal", "radiance")}
        out["pdf_pos"] = np.zeros(n, np.float32)
        out["pdf_dir"] = np.zeros(n, np.float32)
        self.lib.orc_light_emit_batch(self.h, li, seed, n, *[_p(out[k]) for k in (
            "pos", "dir", "normal", "radiance", "pdf_pos", "pdf_dir")])
        return out

    def env_pdf(self, li, d):
        d = np.ascontiguousarray(d, np.float32)
        return float(self.lib.orc_env_pdf(self.h, li, _p(d)))

    def env_radiance(self, li, d):
        d = np.ascontiguousarray(d, np.float32)
        rgb = np.zeros(3, np.float32)
        self.lib.orc_env_radiance(self.h, li, _p(d), _p(rgb))
        return rgb

    def camera(self):
        cs = self.scene.camera
        cam = Cam()
        f = lambda v: np.ascontiguousarray(v, np.float32)  # noqa: E731
        a, b, c = f(cs.pos), f(cs.look_at), f(cs.up)
        self.lib.orc_camera_setup(C.byref(cam), _p(a), _p(b), _p(c), cs.fov, cs.res[0],
                                  cs.res[1], cs.lens_radius, cs.focal_distance)
        return cam

    def args(self, spp=None, seed_mode=0, seed_base=0, env_cdf_fp64=1, use_mis=True):
        pt = self.scene.pt
        return Args(pt.spp if spp is None else spp, pt.tile_size, pt.min_bounces,
                    pt.max_bounces, pt.cont_prob, 1 if use_mis else 0, seed_mode, seed_base,
                    env_cdf_fp64, 0)

    def traced_dims(self, args=None):
        ts = self.scene.pt.tile_size
        w, h = self.scene.camera.res
        return (w + ts - 1) // ts * ts, (h + ts - 1) // ts * ts

    def render(self, cam=None, args=None, pix_begin=0, pix_end=None, per_sample=False,
               threads=1):
        cam = cam or self.camera()
        args = args or self.args()
        wp, hp = self.traced_dims()
        n = wp * hp
        pix_end = n if pix_end is None else pix_end
        s = np.zeros((n, 3), np.float64)
        s2 = np.zeros((n, 3), np.float64)
        cnt = np.zeros(n, np.uint32)
        ps = np.zeros((n * args.spp, 3), np.float32) if per_sample else None
        if threads > 1:
            q = self.lib.orc_render_mt(self.h, C.byref(cam), C.byref(args), pix_begin, pix_end,
                                       _p(s), _p(s2), _p(cnt), threads)
        else:
            q = self.lib.orc_render(self.h, C.byref(cam), C.byref(args), pix_begin, pix_end,
                                    _p(s), _p(s2), _p(cnt), _p(ps) if ps is not None else None)
        return dict(sum=s, sum_sq=s2, count=cnt, per_sample=ps, queries=int(q))

    def render_sppm(self, spec, iterations=0, cam=None, threads=1):
        """SPPM (the scene's SPPMSpec): per-pixel radiance (h, w, 3), the
        per-pixel state (SPPM_PIXEL_DTYPE) and the photon deposit count."""
        cam = cam or self.camera()
        w, h = self.scene.camera.res
        a = SppmArgs(spec.forward_max_depth, spec.init_radius, spec.iteration_count,
                     spec.photons_per_iteration, spec.photon_min_depth, spec.photon_max_depth,
                     spec.photon_cont_prob, spec.update_alpha, spec.photon_seed_base, 0)
        rad = np.zeros((h, w, 3), np.float32)
        st = np.zeros(w * h, SPPM_PIXEL_DTYPE)
        if threads > 1:
            n = self.lib.orc_render_sppm_mt(self.h, C.byref(cam), C.byref(a), iterations, _p(rad),
                                            st.ctypes.data, threads)
        else:
            n = self.lib.orc_render_sppm(self.h, C.byref(cam), C.byref(a), iterations, _p(rad),
                                         st.ctypes.data)
        if n < 0:
            raise ValueError("sppm: film width and height must be multiples of 64")
        return rad, st.reshape(h, w), int(n)

    def render_vcm(self, spec, cam=None, variant=0, threads=1):
        """VCM (the scene's VCMSpec): (color_sum (h, w, 3), splat (h, w, 3),
        light vertices of the last iteration). variant: VCM_* bits of the
        test-only estimator variants (0: the reference as written)."""
        cam = cam or self.camera()
        w, h = self.scene.camera.res
        a = VcmArgs(spec.iteration, spec.base_radius, spec.radius_alpha, int(spec.use_vm),
                    int(spec.use_vc), spec.min_path_length, spec.max_path_length, spec.cont_prob,
                    spec.seed_base, variant)
        col = np.zeros((h, w, 3), np.float32)
        spl = np.zeros((h, w, 3), np.float32)
        if threads > 1:
            n = self.lib.orc_render_vcm_mt(self.h, C.byref(cam), C.byref(a), _p(col), _p(spl),
                                           threads)
        else:
            n = self.lib.orc_render_vcm(self.h, C.byref(cam), C.byref(a), _p(col), _p(spl))
        return col, spl, int(n)

    def close(self):
        if self.h:
            self.lib.orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def sppm_grid_probe(bmin, side, vp_pos, vp_r, ph_pos):
    """Per visible point, the photons the SPPM grid lookup matches to it."""
    lib = load()
    bmin = np.ascontiguousarray(bmin, np.float32)
    vp_pos = np.ascontiguousarray(vp_pos, np.float32).reshape(-1, 3)
    vp_r = np.ascontiguousarray(vp_r, np.float32)
    ph_pos = np.ascontiguousarray(ph_pos, np.float32).reshape(-1, 3)
    out = np.zeros(vp_pos.shape[0], np.uint32)
    lib.orc_sppm_grid_probe(_p(bmin), side, vp_pos.shape[0], _p(vp_pos), _p(vp_r),
                            ph_pos.shape[0], _p(ph_pos), _p(out))
    return out


def vcm_to_raster(cam, u2, pos):
    """ThinLensCamera::ToRaster: (ok, dir_to_camera, raster xy)."""
    lib = load()
    u2 = np.ascontiguousarray(u2, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    d = np.zeros(3, np.float32)
    r = np.zeros(2, np.float32)
    ok = lib.orc_vcm_to_raster(C.byref(cam), _p(u2), _p(pos), _p(d), _p(r))
    return bool(ok), d, r


def vcm_image_to_solid(cam, d):
    lib = load()
    d = np.ascontiguousarray(d, np.float32)
    return float(lib.orc_vcm_image_to_solid(C.byref(cam), _p(d)))


def film_resolve(w, h, ts, sums, counts):
    lib = load()
    s = np.ascontiguousarray(sums, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros((h, w, 3), np.float32)
    lib.orc_film_resolve(w, h, ts, _p(s), _p(c), _p(out))
    return out


def film_resolve_filter(w, h, ts, sums, counts, ftype=0, radius=(0.5, 0.5), alpha=2.0):
    lib = load()
    s = np.ascontiguousarray(sums, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros((h, w, 3), np.float32)
    lib.orc_film_resolve_filter(w, h, ts, ftype, radius[0], radius[1], alpha, _p(s), _p(c), _p(out))
    return out


def image_srgb8(rgb, flip=True):
    lib = load()
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    lib.orc_image_srgb8(w, h, _p(rgb), 1 if flip else 0, _p(out))
    return out


def bsdf_eval(mat, ng, ns, wi, wo):
    lib = load()
    f = np.zeros(3, np.float32)
    m = _mat(mat)
    lib.orc_bsdf_eval(C.byref(m), *[_p(np.ascontiguousarray(x, np.float32))
                                     for x in (ng, ns, wi, wo)], _p(f))
    return f


def bsdf_pdf(mat, ng, ns, wi, wo):
    lib = load()
    m = _mat(mat)
    return float(lib.orc_bsdf_pdf(C.byref(m), *[_p(np.ascontiguousarray(x, np.float32))
                                                for x in (ng, ns, wi, wo)]))


def bsdf_sample(mat, ng, ns, wo, sam):
    lib = load()
    m = _mat(mat)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    valid = np.zeros(1, np.int32)
    args = [np.ascontiguousarray(x, np.float32) for x in (ng, ns, wo, sam)]
    lib.orc_bsdf_sample(C.byref(m), *[_p(a) for a in args], _p(wi), _p(f), _p(pdf), _p(valid))
    return wi, f, float(pdf[0]), bool(valid[0])


def bsdf_sample_batch(mat, ng, ns, wo, sam):
    """Directions (n x 3) and validity (n) of bsdf_sample at one wo over the
    rows of sam (n x 3)."""
    lib = load()
    m = _mat(mat)
    sam = np.ascontiguousarray(sam, np.float32)
    n = sam.shape[0]
    wi = np.zeros((n, 3), np.float32)
    valid = np.zeros(n, np.int32)
    args = [np.ascontiguousarray(x, np.float32) for x in (ng, ns, wo)]
    lib.orc_bsdf_sample_batch(C.byref(m), *[_p(a) for a in args], n, _p(sam), _p(wi), _p(valid))
    return wi, valid.astype(bool)


def sppm_update_probe(alpha, r0, beta, m, phi):
    """UpdatePixel over len(m) iterations: (radius, N, tau (3)) after each."""
    lib = load()
    m = np.ascontiguousarray(m, np.uint32)
    phi = np.ascontiguousarray(phi, np.float32).reshape(-1, 3)
    out = np.zeros((len(m), 5), np.float32)
    lib.orc_sppm_update_probe(float(alpha), float(r0), _p(np.ascontiguousarray(beta, np.float32)),
                              len(m), _p(m), _p(phi), _p(out))
    return out[:, 0], out[:, 1], out[:, 2:]


def _mat(m):
    return Mat(m.type, (C.c_float * 3)(*m.base_color), m.subsurface, m.metallic, m.specular,
               m.specular_tint, m.roughness, m.anisotropic, m.sheen, m.sheen_tint, m.clearcoat,
               m.clearcoat_gloss)


def cosine_hemisphere(s2):
    lib = load()
    out = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib.orc_cosine_hemisphere(_p(np.ascontiguousarray(s2, np.float32)), _p(out), _p(pdf))
    return out, float(pdf[0])


def square_to_disk(s2):
    lib = load()
    out = np.zeros(2, np.float32)
    lib.orc_square_to_disk(_p(np.ascontiguousarray(s2, np.float32)), _p(out))
    return out


def frame(n):
    lib = load()
    dx = np.zeros(3, np.float32)
    dy = np.zeros(3, np.float32)
    lib.orc_frame(_p(np.ascontiguousarray(n, np.float32)), _p(dx), _p(dy))
    return dx, dy


def camera_ray(cam, rx, ry, lens=(0.5, 0.5)):
    lib = load()
    r = np.zeros(1, RAY_DTYPE)
    lib.orc_camera_ray(C.byref(cam), rx, ry, _p(np.ascontiguousarray(lens, np.float32)), _p(r))
    return r[0]
