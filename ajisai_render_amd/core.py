"""Host-side mirror of Ajisai's Aggregate and Renderer interfaces over the C-ABI.

* ``PLOCAggregate`` mirrors ``Aggregate`` / ``CreatePLOCAggregate(search_radius)``
  (ajisai/core/aggregate/aggregate.h:32-46): ``Build(primitives)``, batched
  ``Intersect(rays)`` and ``Occlude(rays)``.
* ``GpuScene`` mirrors ``DefaultScene`` (scene/default_scene.cpp:29-103).
* ``PathTracer`` mirrors ``CreatePTRenderer(PTRendererArgs)`` /
  ``Renderer::Render`` (renderer/renderer.h:37-80, pt.cpp:40-73).
* ``PathGuide`` is the learned directional sampler (no reference
  counterpart): an MLP giving an 8-lobe vMF mixture per shading point.

Device buffers are torch tensors on ``cuda``; torch is only used for device
memory and streams. Every compute call goes through libajisai_gpu.so.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _capi as capi
from .scenes import Light, Material, Primitive, Scene


def _torch():
    import torch  # noqa: WPS433 (imported lazily: plumbing only)
    return torch


def _ptr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())


class Context:
    """ajg_ctx: one per device."""

    def __init__(self, device: int = 0):
        self.lib = capi.load()
        h = C.c_void_p()
        st = self.lib.ajg_ctx_create(device, C.byref(h))
        if st != capi.AJG_OK:
            raise capi.AjgError(st, f"ajg_ctx_create(device={device})")
        self.h = h
        self.device = device

    def check(self, st, what=""):
        capi.check(st, self.h, what)

    def inject_failure(self, countdown: int, status: int):
        self.check(self.lib.ajg_ctx_inject_failure(self.h, countdown, status))

    def set_cancel(self, cancel: bool = True):
        """Raise (or clear) the context's cancel flag: renders in progress on
        this context stop at their next host checkpoint with AJG_ERR_CANCELLED."""
        self.check(self.lib.ajg_ctx_set_cancel(self.h, 1 if cancel else 0))

    def close(self):
        if self.h:
            self.lib.ajg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def mesh_views(prims: Sequence[Primitive]):
    """MeshView structs over the primitives' numpy buffers (kept alive by
    the returned list)."""
    arr = (capi.MeshView * max(1, len(prims)))()
    keep = []
    for i, p in enumerate(prims):
        pos = np.ascontiguousarray(p.positions, np.float32)
        idx = np.ascontiguousarray(p.index_words, np.uint32)
        keep += [pos, idx]
        arr[i] = capi.MeshView(pos.ctypes.data, 0, pos.strides[0], pos.shape[0],
                               idx.ctypes.data, 0, 12, idx.shape[0])
    return arr, keep


def c_primitives(prims: Sequence[Primitive]):
    views, keep = mesh_views(prims)
    arr = (capi.Primitive * max(1, len(prims)))()
    for i, p in enumerate(prims):
        nrm = None
        if p.normal_and_vs is not None:
            nrm = np.ascontiguousarray(p.normal_and_vs, np.float32)
            keep.append(nrm)
        arr[i] = capi.Primitive(views[i], p.kind, int(p.two_sided), p.material, 0,
                                nrm.ctypes.data if nrm is not None else None,
                                (C.c_float * 3)(*p.emission), 0)
    return arr, keep


def c_materials(mats: Sequence[Material]):
    arr = (capi.Material * max(1, len(mats)))()
    for i, m in enumerate(mats):
        arr[i] = capi.Material(m.type, (C.c_float * 3)(*m.base_color), m.subsurface, m.metallic,
                               m.specular, m.specular_tint, m.roughness, m.anisotropic,
                               m.sheen, m.sheen_tint, m.clearcoat, m.clearcoat_gloss)
    return arr


def c_lights(lights: Sequence[Light]):
    arr = (capi.Light * max(1, len(lights)))()
    keep = []
    for i, l in enumerate(lights):
        if l.env is not None:
            env = np.ascontiguousarray(l.env, np.float32)
            keep.append(env)
            h, w = env.shape[:2]
            ptr = env.ctypes.data
        else:
            h = w = 0
            ptr = None
        arr[i] = capi.Light(l.type, (C.c_float * 3)(*l.dir), l.angle,
                            (C.c_float * 3)(*l.radiance), w, h, ptr, l.rot, l.scaler)
    return arr, keep


class PLOCAggregate:
    """Aggregate "ploc" on the GPU (CreatePLOCAggregate, aggregate.h:45)."""

    def __init__(self, ctx: Context, search_radius: int = 14, handle=None, owner=True):
        self.ctx = ctx
        self.search_radius = search_radius
        self.h = handle
        self.owner = owner

    # Aggregate::Build (aggregate.h:36)
    def Build(self, primitives: Sequence[Primitive]):
        self._release()
        views, keep = mesh_views(primitives)
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.ajg_accel_build(self.ctx.h, views, len(primitives),
                                                    self.search_radius, C.byref(h)),
                       "ajg_accel_build")
        self.h = h
        del keep
        return self

    def build_soup(self, tri_dev, scene_box: Optional[np.ndarray] = None, stream=None):
        """Build from a device-resident (n, 9) float32 soup."""
        self._release()
        box = None if scene_box is None else np.ascontiguousarray(scene_box, np.float32)
        h = C.c_void_p()
        n = int(tri_dev.shape[0])
        self.ctx.check(self.ctx.lib.ajg_accel_build_soup(
            self.ctx.h, _ptr(tri_dev), n, None if box is None else box.ctypes.data,
            self.search_radius, stream, C.byref(h)), "ajg_accel_build_soup")
        self.h = h
        return self

    def info(self) -> capi.AccelInfo:
        out = capi.AccelInfo()
        self.ctx.check(self.ctx.lib.ajg_accel_get_info(self.h, C.byref(out)))
        return out

    def download(self):
        inf = self.info()
        nodes = np.zeros(inf.node_count, capi.NODE_DTYPE)
        perm = np.zeros(inf.tri_count, np.uint32)
        self.ctx.check(self.ctx.lib.ajg_accel_download(self.h, nodes.ctypes.data,
                                                       perm.ctypes.data), "download")
        return nodes, perm

    # Aggregate::Intersect (aggregate.h:38-39), batched over device rays
    def Intersect(self, rays_dev, hits_dev=None, stream=None):
        torch = _torch()
        n = rays_dev.numel() // 32
        if hits_dev is None:
            hits_dev = torch.empty(n * 16, dtype=torch.uint8, device=rays_dev.device)
        self.ctx.check(self.ctx.lib.ajg_accel_trace_closest(self.h, _ptr(rays_dev), n,
                                                            _ptr(hits_dev), stream),
                       "trace_closest")
        return hits_dev

    # Aggregate::Occlude (aggregate.h:41), batched
    def Occlude(self, rays_dev, occ_dev=None, stream=None):
        torch = _torch()
        n = rays_dev.numel() // 32
        if occ_dev is None:
            occ_dev = torch.empty(n, dtype=torch.uint8, device=rays_dev.device)
        self.ctx.check(self.ctx.lib.ajg_accel_trace_any(self.h, _ptr(rays_dev), n,
                                                        _ptr(occ_dev), stream), "trace_any")
        return occ_dev

    def counts(self, rays_dev, stream=None):
        torch = _torch()
        n = rays_dev.numel() // 32
        nodes = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        tris = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        self.ctx.check(self.ctx.lib.ajg_accel_trace_counts(self.h, _ptr(rays_dev), n,
                                                           _ptr(nodes), _ptr(tris), stream),
                       "trace_counts")
        return nodes, tris

    def _release(self):
        if self.h and self.owner:
            self.ctx.lib.ajg_accel_destroy(self.h)
        self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self._release()
        except Exception:
            pass


class GpuScene:
    """DefaultScene on the device (default_scene.cpp:29-103) with the GPU
    PLOC aggregate."""

    def __init__(self, ctx: Context, scene: Scene, env_cdf_fp64: int = 1):
        self.ctx = ctx
        self.scene = scene
        prims, keep = c_primitives(scene.primitives)
        mats = c_materials(scene.materials)
        lights, keep2 = c_lights(scene.lights)
        h = C.c_void_p()
        if scene.instances is not None:
            insts = (capi.Instance * max(1, len(scene.instances)))()
            for i, ins in enumerate(scene.instances):
                m = np.asarray(ins.o2w, np.float32).reshape(12)
                insts[i] = capi.Instance(ins.mesh, capi.INSTANCE_KEEP_MATERIAL if ins.material < 0
                                         else ins.material, (C.c_float * 12)(*m))
            ctx.check(ctx.lib.ajg_scene_create_instanced(
                ctx.h, prims, len(scene.primitives), insts, len(scene.instances), mats,
                len(scene.materials), lights, len(scene.lights), scene.search_radius,
                env_cdf_fp64, C.byref(h)), "ajg_scene_create_instanced")
        else:
            ctx.check(ctx.lib.ajg_scene_create(ctx.h, prims, len(scene.primitives), mats,
                                               len(scene.materials), lights, len(scene.lights),
                                               scene.search_radius, env_cdf_fp64, C.byref(h)),
                      "ajg_scene_create")
        self.h = h
        del keep, keep2
        acc = C.c_void_p()
        ctx.check(ctx.lib.ajg_scene_get_accel(self.h, C.byref(acc)))
        self.accel = PLOCAggregate(ctx, scene.search_radius, handle=acc, owner=False)

    # Scene::Intersect / Occlude batched (default_scene.cpp:63-70)
    def Intersect(self, rays_dev, stream=None):
        """Returns (hits n x 16 B as uint8 device tensor viewable as
        ajg_hit, instance ids)."""
        torch = _torch()
        n = rays_dev.numel() // 32
        hits = torch.empty(n * 16, dtype=torch.uint8, device=rays_dev.device)
        inst = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        self.ctx.check(self.ctx.lib.ajg_scene_trace_closest(self.h, _ptr(rays_dev), n, _ptr(hits),
                                                            _ptr(inst), stream), "scene trace")
        return hits, inst

    def Occlude(self, rays_dev, stream=None):
        torch = _torch()
        n = rays_dev.numel() // 32
        occ = torch.empty(n, dtype=torch.uint8, device=rays_dev.device)
        self.ctx.check(self.ctx.lib.ajg_scene_trace_any(self.h, _ptr(rays_dev), n, _ptr(occ),
                                                        stream), "scene occlude")
        return occ

    def instances(self):
        """(world->object maps (n, 12), padded world boxes (n, 6)) as used
        on the device."""
        n = len(self.scene.instances)
        w2o = np.zeros((n, 12), np.float32)
        box = np.zeros((n, 6), np.float32)
        self.ctx.check(self.ctx.lib.ajg_scene_instances(self.h, w2o.ctypes.data, box.ctypes.data),
                       "instances")
        return w2o, box

    def env_tables(self, li):
        l = self.scene.lights[li]
        h, w = l.env.shape[:2]
        cc = np.zeros((h, w + 1), np.float32)
        cf = np.zeros(h, np.float32)
        mc = np.zeros(h + 1, np.float32)
        mf = np.zeros(1, np.float32)
        self.ctx.check(self.ctx.lib.ajg_scene_env_tables(self.h, li, cc.ctypes.data,
                                                         cf.ctypes.data, mc.ctypes.data,
                                                         mf.ctypes.data), "env tables")
        return cc, cf, mc, mf[0]

    def close(self):
        if self.h:
            self.ctx.lib.ajg_scene_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class SPPMRenderer:
    """CreateSPPMRenderer(SPPMRendererArgs) (renderer.h:82-97, 113) /
    SPPMRenderer::Render (sppm.cpp:184-294) on the device. The per-pixel
    radiance is the value the reference adds to its film, one sample per
    pixel; `resolve_filtered` applies the film's filter."""

    def __init__(self, gscene: GpuScene, spec=None, timing=False):
        from .scenes import SPPMSpec
        self.g = gscene
        sp = spec if spec is not None else SPPMSpec()
        self.spec = sp
        self.args = capi.SppmArgs(sp.forward_max_depth, sp.init_radius, sp.iteration_count,
                                  sp.photons_per_iteration, sp.photon_min_depth,
                                  sp.photon_max_depth, sp.photon_cont_prob, sp.update_alpha,
                                  sp.photon_seed_base, capi.RENDER_TIMING if timing else 0)

    def render_device(self, cam_override=None, state=False, stream=None):
        """-> (radiance (H, W, 3) device tensor, state (H*W*64 bytes) or None,
        SppmStats)."""
        torch = _torch()
        w, h = self.g.scene.camera.res
        if cam_override is None:
            cam = camera_setup(self.g.scene.camera)
        else:
            cam = capi.Camera()
            C.memmove(C.byref(cam), C.byref(cam_override), C.sizeof(capi.Camera))
        rad = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        st = (torch.empty(h * w * capi.SPPM_PIXEL_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
              if state else None)
        stats = capi.SppmStats()
        self.g.ctx.check(self.g.ctx.lib.ajg_render_sppm(
            self.g.h, C.byref(cam), C.byref(self.args), _ptr(rad), _ptr(st), C.byref(stats),
            stream), "ajg_render_sppm")
        return rad.view(h, w, 3), st, stats

    # Renderer::Render (renderer.h:64-65)
    def Render(self, cam_override=None, state=False):
        rad, st, stats = self.render_device(cam_override, state)
        out = {"radiance": rad.cpu().numpy(), "stats": stats}
        if st is not None:
            w, h = self.g.scene.camera.res
            out["state"] = st.cpu().numpy().view(capi.SPPM_PIXEL_DTYPE).reshape(h, w)
        return out

    def resolve_filtered(self, rad_dev, ftype="box", radius=None, alpha=2.0):
        """Film::AddSample of one sample per pixel then MergeTile with the
        film's filter -> (H, W, 3) device tensor."""
        torch = _torch()
        h, w = rad_dev.shape[:2]
        t = capi.FILTER_BOX if ftype == "box" else capi.FILTER_GAUSSIAN
        r = radius if radius is not None else ((0.5, 0.5) if t == capi.FILTER_BOX else (2.0, 2.0))
        f = capi.Filter(t, (C.c_float * 2)(*r), alpha)
        ones = torch.ones(h * w, dtype=torch.int32, device="cuda")
        rgb = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        self.g.ctx.check(self.g.ctx.lib.ajg_film_resolve_filter(
            self.g.ctx.h, w, h, 64, C.byref(f), _ptr(rad_dev.reshape(-1)), _ptr(ones), _ptr(rgb),
            None), "ajg_film_resolve_filter")
        return rgb.view(h, w, 3)


class VCMRenderer:
    """CreateVCMRenderer(VCMRendererArgs) (renderer.h:102-116) /
    VCMRenderer::Render (vcm.cpp:135-147) on the device: per-pixel sums of
    the camera sub-paths' colors over iterations plus the light-tracing
    splats; `image` combines them as Film::WriteImage does."""

    def __init__(self, gscene: GpuScene, spec=None, timing=False):
        from .scenes import VCMSpec
        self.g = gscene
        sp = spec if spec is not None else VCMSpec()
        self.spec = sp
        self.args = capi.VcmArgs(sp.iteration, sp.base_radius, sp.radius_alpha, int(sp.use_vm),
                                 int(sp.use_vc), sp.min_path_length, sp.max_path_length,
                                 sp.cont_prob, sp.seed_base,
                                 capi.RENDER_TIMING if timing else 0)

    def render_device(self, cam_override=None, stream=None):
        """-> (color_sum (H, W, 3), splat (H, W, 3)) device tensors, VcmStats."""
        torch = _torch()
        w, h = self.g.scene.camera.res
        if cam_override is None:
            cam = camera_setup(self.g.scene.camera)
        else:
            cam = capi.Camera()
            C.memmove(C.byref(cam), C.byref(cam_override), C.sizeof(capi.Camera))
        col = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        spl = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        stats = capi.VcmStats()
        self.g.ctx.check(self.g.ctx.lib.ajg_render_vcm(
            self.g.h, C.byref(cam), C.byref(self.args), _ptr(col), _ptr(spl), C.byref(stats),
            stream), "ajg_render_vcm")
        return col.view(h, w, 3), spl.view(h, w, 3), stats

    def image(self, col_dev, spl_dev, ftype="box", radius=None, alpha=2.0):
        """Film::WriteImage: the filtered camera-path colors (weight 1 per
        iteration) plus the splats -> (H, W, 3) device tensor."""
        torch = _torch()
        h, w = col_dev.shape[:2]
        t = capi.FILTER_BOX if ftype == "box" else capi.FILTER_GAUSSIAN
        r = radius if radius is not None else ((0.5, 0.5) if t == capi.FILTER_BOX else (2.0, 2.0))
        f = capi.Filter(t, (C.c_float * 2)(*r), alpha)
        cnt = torch.full((h * w,), int(self.args.iteration), dtype=torch.int32, device="cuda")
        rgb = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        self.g.ctx.check(self.g.ctx.lib.ajg_film_resolve_filter(
            self.g.ctx.h, w, h, 1, C.byref(f), _ptr(col_dev.reshape(-1)), _ptr(cnt), _ptr(rgb),
            None), "ajg_film_resolve_filter")
        return rgb.view(h, w, 3) + spl_dev

    # Renderer::Render (renderer.h:64-65)
    def Render(self, cam_override=None):
        col, spl, stats = self.render_device(cam_override)
        return {"color_sum": col.cpu().numpy(), "splat": spl.cpu().numpy(), "stats": stats}


class PathTracer:
    """CreatePTRenderer(PTRendererArgs) (renderer.h:37-45, 80) for TraceMIS on
    the device. Renders into per-traced-pixel film sums."""

    def __init__(self, gscene: GpuScene, spp=None, tile_size=None, min_bounces=None,
                 max_bounces=None, cont_prob=None, seed_mode=0, seed_base=0, env_cdf_fp64=1,
                 max_slots=0, flags=0, use_mis=None):
        pt = gscene.scene.pt
        self.g = gscene
        mis = getattr(pt, "use_mis", True) if use_mis is None else use_mis
        self.args = capi.PTArgs(pt.spp if spp is None else spp,
                                pt.tile_size if tile_size is None else tile_size,
                                pt.min_bounces if min_bounces is None else min_bounces,
                                pt.max_bounces if max_bounces is None else max_bounces,
                                pt.cont_prob if cont_prob is None else cont_prob,
                                1 if mis else 0, seed_mode, seed_base, env_cdf_fp64, max_slots,
                                flags)

    def traced_dims(self):
        w, h = self.g.scene.camera.res
        ts = self.args.tile_size
        return (w + ts - 1) // ts * ts, (h + ts - 1) // ts * ts

    def camera(self, cam_override=None) -> capi.Camera:
        if cam_override is None:
            return camera_setup(self.g.scene.camera)
        cam = capi.Camera()
        C.memmove(C.byref(cam), C.byref(cam_override), C.sizeof(capi.Camera))
        return cam

    def render_device(self, cam_override=None, per_sample=False, stream=None, film=None,
                      guide=None, guide_args=None):
        """Render into device tensors; returns (film dict, RenderStats), or
        (film dict, RenderStats, GuideStats) when a PathGuide is given."""
        torch = _torch()
        wp, hp = self.traced_dims()
        n = wp * hp
        if film is None:
            film = dict(sum=torch.empty(n * 3, dtype=torch.float32, device="cuda"),
                        sum_sq=torch.empty(n * 3, dtype=torch.float32, device="cuda"),
                        count=torch.empty(n, dtype=torch.int32, device="cuda"))
            if per_sample:
                film["per_sample"] = torch.empty(n * self.args.spp * 3, dtype=torch.float32,
                                                 device="cuda")
        cf = capi.Film(_ptr(film["sum"]), _ptr(film["sum_sq"]), _ptr(film["count"]),
                       _ptr(film.get("per_sample")))
        stats = capi.RenderStats()
        cam = self.camera(cam_override)
        if guide is not None:
            ga = guide_args if guide_args is not None else guide_defaults()
            gs = capi.GuideStats()
            self.g.ctx.check(self.g.ctx.lib.ajg_render_pt_guided(
                self.g.h, C.byref(cam), C.byref(self.args), guide.h, C.byref(ga), C.byref(cf),
                C.byref(stats), C.byref(gs), stream), "ajg_render_pt_guided")
            return film, stats, gs
        self.g.ctx.check(self.g.ctx.lib.ajg_render_pt(self.g.h, C.byref(cam), C.byref(self.args),
                                                      C.byref(cf), C.byref(stats), stream),
                         "ajg_render_pt")
        return film, stats

    # Renderer::Render (renderer.h:64-65)
    def Render(self, cam_override=None, per_sample=False, guide=None, guide_args=None):
        res = self.render_device(cam_override, per_sample, guide=guide, guide_args=guide_args)
        film, stats = res[0], res[1]
        out = {k: v.cpu().numpy() for k, v in film.items()}
        out["sum"] = out["sum"].reshape(-1, 3)
        out["sum_sq"] = out["sum_sq"].reshape(-1, 3)
        out["count"] = out["count"].view(np.uint32)
        if "per_sample" in out:
            out["per_sample"] = out["per_sample"].reshape(-1, 3)
        out["stats"] = stats
        if guide is not None:
            out["guide_stats"] = res[2]
        return out

    def resolve_filtered(self, film_dev, ftype="box", radius=None, alpha=2.0):
        """Film resolve with the reference's box (radius 0.5) or gaussian
        (radius 2, alpha 2) filter -> (H, W, 3) device tensor."""
        torch = _torch()
        w, h = self.g.scene.camera.res
        t = capi.FILTER_BOX if ftype == "box" else capi.FILTER_GAUSSIAN
        r = radius if radius is not None else ((0.5, 0.5) if t == capi.FILTER_BOX else (2.0, 2.0))
        f = capi.Filter(t, (C.c_float * 2)(*r), alpha)
        rgb = torch.empty(h * w * 3, dtype=torch.float32, device="cuda")
        self.g.ctx.check(self.g.ctx.lib.ajg_film_resolve_filter(
            self.g.ctx.h, w, h, self.args.tile_size, C.byref(f), _ptr(film_dev["sum"]),
            _ptr(film_dev["count"]), _ptr(rgb), None), "ajg_film_resolve_filter")
        return rgb.view(h, w, 3)

    def write_srgb8(self, rgb_dev, flip=True):
        """ImageWriter::Write pixel path -> (H, W, 3) uint8 numpy array (rows
        bottom to top when flip, as the reference stores its PNG)."""
        torch = _torch()
        h, w = rgb_dev.shape[:2]
        out = torch.empty(h * w * 3, dtype=torch.uint8, device="cuda")
        self.g.ctx.check(self.g.ctx.lib.ajg_image_srgb8(self.g.ctx.h, w, h, _ptr(rgb_dev),
                                                        1 if flip else 0, _ptr(out), None),
                         "ajg_image_srgb8")
        return out.cpu().numpy().reshape(h, w, 3)

    def resolve(self, film_dev):
        """Film resolve on the device -> (H, W, 3) numpy image."""
        torch = _torch()
        w, h = self.g.scene.camera.res
        rgb = torch.empty(w * h * 3, dtype=torch.float32, device="cuda")
        self.g.ctx.check(self.g.ctx.lib.ajg_film_resolve(self.g.ctx.h, w, h, self.args.tile_size,
                                                         _ptr(film_dev["sum"]),
                                                         _ptr(film_dev["count"]), _ptr(rgb), None),
                         "ajg_film_resolve")
        return rgb.cpu().numpy().reshape(h, w, 3)


def batch_defaults(max_slots=0, samples_per_item=0, trav_blocks_per_cu=0, lds_stack=0,
                   iters_per_check=0, use_graph=1, pixel_tile=0):
    """ajg_batch_args (0 = the library's default for that knob)."""
    return capi.BatchArgs(max_slots, samples_per_item, trav_blocks_per_cu, lds_stack,
                          iters_per_check, use_graph, 0, pixel_tile)


class BatchRenderer:
    """Many (scene, camera, PTRendererArgs) jobs in one wavefront
    (ajg_render_batch). Jobs are PathTracer objects; each keeps its own
    PTRendererArgs and receives its own film."""

    def __init__(self, ctx: Context, tracers):
        self.ctx = ctx
        self.tracers = list(tracers)
        if not self.tracers:
            raise ValueError("a batch needs at least one job")

    def _jobs(self, films):
        jobs = (capi.Job * len(self.tracers))()
        for i, (pt, f) in enumerate(zip(self.tracers, films)):
            jobs[i].scene = pt.g.h.value if isinstance(pt.g.h, C.c_void_p) else pt.g.h
            jobs[i].camera = pt.camera()
            C.memmove(C.byref(jobs[i].args), C.byref(pt.args), C.sizeof(capi.PTArgs))
            jobs[i].film = capi.Film(_ptr(f["sum"]), _ptr(f["sum_sq"]), _ptr(f["count"]),
                                     _ptr(f.get("per_sample")))
        return jobs

    def films(self, per_sample=False):
        torch = _torch()
        out = []
        for pt in self.tracers:
            wp, hp = pt.traced_dims()
            n = wp * hp
            f = dict(sum=torch.empty(n * 3, dtype=torch.float32, device="cuda"),
                     sum_sq=torch.empty(n * 3, dtype=torch.float32, device="cuda"),
                     count=torch.empty(n, dtype=torch.int32, device="cuda"))
            if per_sample:
                f["per_sample"] = torch.empty(n * pt.args.spp * 3, dtype=torch.float32,
                                              device="cuda")
            out.append(f)
        return out

    def render_device(self, bargs=None, films=None, per_sample=False, stream=None):
        """Returns (films, BatchStats, per-job completion ms as numpy)."""
        films = self.films(per_sample) if films is None else films
        jobs = self._jobs(films)
        ba = bargs if bargs is not None else batch_defaults()
        st = capi.BatchStats()
        done = np.zeros(len(self.tracers), np.float32)
        self.ctx.check(self.ctx.lib.ajg_render_batch(self.ctx.h, jobs, len(self.tracers),
                                                     C.byref(ba), C.byref(st), done.ctypes.data,
                                                     stream), "ajg_render_batch")
        return films, st, done

    def autotune(self, trial_spp=2, start=None, films=None):
        """Coordinate-descent autotuner; returns (BatchArgs, best paths/s)."""
        films = self.films() if films is None else films
        jobs = self._jobs(films)
        ba = capi.BatchArgs()
        C.memmove(C.byref(ba), C.byref(start if start is not None else batch_defaults()),
                  C.sizeof(capi.BatchArgs))
        best = C.c_float(0.0)
        self.ctx.check(self.ctx.lib.ajg_batch_autotune(self.ctx.h, jobs, len(self.tracers),
                                                       trial_spp, C.byref(ba), C.byref(best)),
                       "ajg_batch_autotune")
        return ba, float(best.value)


def guide_defaults(alpha=0.5, train=True, batch=1 << 18, lr=1e-3, train_slots=0, max_steps=0,
                   steps_per_batch=4, warmup_passes=0, warmup_spp=16, warmup_steps_per_batch=64):
    """ajg_guide_args: one-sample mixture with probability alpha, online Adam
    training on batches of 2^18 records at lr 1e-3, optional warm-up passes."""
    return capi.GuideArgs(alpha, 1 if train else 0, batch, lr, train_slots, max_steps,
                          steps_per_batch, warmup_passes, warmup_spp, warmup_steps_per_batch)


class PathGuide:
    """ajg_guide: MLP x(112) -> 64 -> 64 -> 64 -> 48 (BF16 MFMA, FP32
    accumulation) whose outputs parameterise an 8-lobe vMF mixture."""

    N_IN, N_HID, N_OUT, N_LOBES = 112, 64, 48, 8

    def __init__(self, ctx: Context, scene_box, seed: int = 1):
        self.ctx = ctx
        box = np.ascontiguousarray(np.asarray(scene_box, np.float32).reshape(6))
        h = C.c_void_p()
        ctx.check(ctx.lib.ajg_guide_create(ctx.h, seed, box.ctypes.data, C.byref(h)),
                  "ajg_guide_create")
        self.h = h
        self.box = box
        self.n_params = int(ctx.lib.ajg_guide_param_count())

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.lib.ajg_guide_get_params(self.h, out.ctypes.data), "get params")
        return out

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self.ctx.check(self.ctx.lib.ajg_guide_set_params(self.h, p.ctypes.data), "set params")

    def forward(self, pos_dev, nrm_dev, want_raw=True, stream=None):
        """Returns (mixture n x 40, raw n x 48) device tensors."""
        torch = _torch()
        n = int(pos_dev.shape[0])
        dev = pos_dev.device
        mix = torch.empty((n, 40), dtype=torch.float32, device=dev)
        raw = torch.empty((n, 48), dtype=torch.float32, device=dev) if want_raw else None
        self.ctx.check(self.ctx.lib.ajg_guide_forward(self.h, _ptr(pos_dev), _ptr(nrm_dev), n,
                                                      _ptr(mix), _ptr(raw), stream), "forward")
        return mix, raw

    def train_step(self, samples_dev, lr=1e-3, want_grad=False, stream=None):
        """One step on device samples (n x 10: pos, nrm, dir, weight).
        Returns (loss, grad or None)."""
        n = int(samples_dev.shape[0])
        loss = C.c_float(0.0)
        grad = np.zeros(self.n_params, np.float32) if want_grad else None
        self.ctx.check(self.ctx.lib.ajg_guide_train_step(
            self.h, _ptr(samples_dev), n, float(lr), C.addressof(loss),
            grad.ctypes.data if grad is not None else None, stream), "train step")
        return float(loss.value), grad

    NOALLOC_MAX_SAMPLES = 32768

    def train_step_noalloc(self, samples_dev, lr=1e-3, stream=None):
        """The same step through ajg_guide_train_step_noalloc: no device
        allocation, work buffers in the library's static scratch, at most
        NOALLOC_MAX_SAMPLES samples; enqueued only (no loss returned)."""
        n = int(samples_dev.shape[0])
        self.ctx.check(self.ctx.lib.ajg_guide_train_step_noalloc(
            self.h, _ptr(samples_dev), n, float(lr), stream), "train step (no allocation)")

    def close(self):
        if self.h:
            self.ctx.lib.ajg_guide_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def camera_setup(cam_spec) -> capi.Camera:
    lib = capi.load()
    cam = capi.Camera()
    f3 = lambda v: np.ascontiguousarray(v, np.float32)  # noqa: E731
    pos, la, up = f3(cam_spec.pos), f3(cam_spec.look_at), f3(cam_spec.up)
    st = lib.ajg_camera_setup(C.byref(cam), pos.ctypes.data, la.ctypes.data, up.ctypes.data,
                              cam_spec.fov, cam_spec.res[0], cam_spec.res[1],
                              cam_spec.lens_radius, cam_spec.focal_distance)
    capi.check(st, None, "ajg_camera_setup")
    return cam


def rays_to_bytes(rays: np.ndarray) -> np.ndarray:
    assert rays.dtype == capi.RAY_DTYPE
    return np.ascontiguousarray(rays).view(np.uint8).reshape(-1)


def hits_from_bytes(buf: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(buf).view(capi.HIT_DTYPE)
