"""Helpers of the instancing tests: small instanced scenes and the oracle's
two-level trees for them (test infrastructure)."""
import numpy as np

import orc
from ajisai_render_amd import scenes
from ajisai_render_amd.scenes import Instance, Material, MAT_DIFFUSE, PTSpec, CameraSpec, Scene


def small_instanced(seed=5, n_inst=40, subdivs=(2, 3, 1), extent=6.0, quad_mesh=True,
                    identity=False):
    """A few displaced icospheres (and a quad) placed by random rotations,
    scales in [0.2, 5] and translations; identity=True places every mesh at
    the origin unmoved (one instance each)."""
    rng = np.random.default_rng(seed)
    prims = []
    for k, s in enumerate(subdivs):
        pos4, nrm4, f = scenes.wavy_icosphere(s, rng, radius=1.0,
                                              center=(3.0 * k, 0.0, 0.0) if identity else (0, 0, 0))
        prims.append(scenes.mesh(pos4, nrm4, f, k % 2, two_sided=True))
    if quad_mesh:
        prims.append(scenes.quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), 0))
    insts = []
    if identity:
        for m in range(len(prims)):
            insts.append(Instance(m, np.concatenate([np.eye(3), np.zeros((3, 1))], axis=1)))
    else:
        for _ in range(n_inst):
            m = int(rng.integers(0, len(prims)))
            sc = float(np.exp(rng.uniform(np.log(0.2), np.log(5.0))))
            rot = scenes.random_rotation(rng) * sc
            t = rng.uniform(-extent, extent, 3)
            insts.append(Instance(m, np.concatenate([rot, t[:, None]], axis=1),
                                  int(rng.integers(0, 2)) if rng.random() < 0.3 else -1))
    mats = [Material(MAT_DIFFUSE, (0.6, 0.5, 0.4)), Material(MAT_DIFFUSE, (0.3, 0.6, 0.5))]
    cam = CameraSpec((0.0, 2.0, 3 * extent), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, (64, 48))
    return Scene("small_instanced", prims, mats, [], cam, PTSpec(), instances=insts)


def mesh_tree(prim):
    """Oracle PLOC tree of one mesh in its own space (as the device builds
    each bottom-level tree)."""
    one = Scene("mesh", [prim], [], [], None, PTSpec())
    tri = one.soup()
    nodes, perm, _ = orc.ploc_build(tri, one.scene_box())
    return nodes, perm, tri


def host_instances(scene):
    """World->object maps (float32, from a float64 inverse) and conservative
    world boxes of the instances, computed on the host (CPU tests)."""
    w2o, boxes = [], []
    for ins in scene.instances:
        m = np.asarray(ins.o2w, np.float64)
        inv = np.linalg.inv(m[:, :3])
        w2o.append(np.concatenate([inv, (-inv @ m[:, 3])[:, None]], axis=1).reshape(12))
        pos = scene.primitives[ins.mesh].positions[:, :3].astype(np.float64)
        w = pos @ m[:, :3].T + m[:, 3]
        mn, mx = w.min(axis=0), w.max(axis=0)
        pad = 1e-4 * (np.abs(w).max() + 1.0)
        boxes.append(np.concatenate([mn - pad, mx + pad]))
    return np.asarray(w2o, np.float32), np.asarray(boxes, np.float32)


def tlas_tree(boxes):
    """Oracle top-level tree: PLOC over the instance boxes, each given as the
    degenerate triangle (min, max, min)."""
    b = np.asarray(boxes, np.float32)
    tri = np.concatenate([b[:, :3], b[:, 3:], b[:, :3]], axis=1)
    wbox = np.concatenate([b[:, :3].min(axis=0), b[:, 3:].max(axis=0)]).astype(np.float32)
    nodes, perm, _ = orc.ploc_build(tri, wbox)
    return nodes, perm, wbox


def oracle_trace(scene, w2o, boxes, rays, closest=True):
    blases = [mesh_tree(p) for p in scene.primitives]
    tn, tp, _ = tlas_tree(boxes)
    inst_mesh = np.array([i.mesh for i in scene.instances], np.uint32)
    return orc.trace_2l(tn, tp, w2o, inst_mesh, blases, rays, closest)


def world_soup(scene):
    """(n, 9) world-space triangles of every instance (float32, from
    float64 transforms) and, per triangle, (instance, triangle in mesh)."""
    tris, owner = [], []
    for k, ins in enumerate(scene.instances):
        p = scene.primitives[ins.mesh]
        m = np.asarray(ins.o2w, np.float64)
        pos = p.positions[:, :3].astype(np.float64) @ m[:, :3].T + m[:, 3]
        t = pos[p.index_words.astype(np.int64)].reshape(-1, 9)
        tris.append(t)
        owner.append(np.stack([np.full(t.shape[0], k), np.arange(t.shape[0])], axis=1))
    return (np.ascontiguousarray(np.concatenate(tris).astype(np.float32)),
            np.concatenate(owner))


def _mesh_plus_far_triangle(prim, offset=(40.0, 25.0, -30.0)):
    """The mesh with one more triangle far away, appended last: the same
    triangles under the same numbers, a much larger box."""
    pos, nrm, f = prim.positions, prim.normal_and_vs, prim.index_words
    n = len(pos)
    extra = pos[:3].copy()
    extra[:, :3] += np.asarray(offset, np.float32)
    return scenes.mesh(np.concatenate([pos, extra]), np.concatenate([nrm, nrm[:3]]),
                       np.concatenate([f, np.array([[n, n + 1, n + 2]], f.dtype)]),
                       prim.material, two_sided=prim.two_sided)


def coincident_scene(seed=7, n_inst=30, dup_of=(3, 7, 7, 12, 20), front=False):
    """small_instanced plus copies of instances `dup_of` under the same maps:
    a copy of instance k uses either k's mesh or that mesh with one far
    triangle appended (alternately), so copies tie at every hit on k while
    their top-level boxes, and with them the visit order, differ. front:
    the copies take the lowest numbers. Returns the scene and, per instance,
    the number of the original it copies."""
    import dataclasses
    sc = small_instanced(seed=seed, n_inst=n_inst)
    prims = list(sc.primitives)
    ext = {}
    extra = []
    for j, k in enumerate(dup_of):
        ins = sc.instances[k]
        m = ins.mesh
        if j % 2 == 1 and prims[m].kind == scenes.GEOM_MESH:
            if m not in ext:
                ext[m] = len(prims)
                prims.append(_mesh_plus_far_triangle(prims[m]))
            m = ext[m]
        extra.append(dataclasses.replace(ins, mesh=m))
    if front:
        insts = extra + list(sc.instances)
        orig = list(dup_of) + list(range(len(sc.instances)))
    else:
        insts = list(sc.instances) + extra
        orig = list(range(len(sc.instances))) + list(dup_of)
    return dataclasses.replace(sc, primitives=prims, instances=insts), np.array(orig)
