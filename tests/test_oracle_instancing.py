"""Pins the oracle's two-level (instanced) traversal: identity instances
equal the one-level traversal bit for bit, and random rotations/scales/
translations agree with a brute-force search over world-space copies of
the instances."""
import numpy as np

import orc
from instancing_util import (coincident_scene, host_instances, mesh_tree, oracle_trace,
                             small_instanced, tlas_tree, world_soup)
from raygen import random_rays


def test_identity_instances_equal_one_level():
    sc = small_instanced(identity=True)
    w2o, boxes = host_instances(sc)
    assert np.array_equal(w2o.reshape(-1, 3, 4)[:, :, :3], np.broadcast_to(np.eye(3), (len(w2o), 3, 3)))
    wbox = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])
    rays = random_rays(wbox, 4000, seed=1)
    hits, inst = oracle_trace(sc, w2o, boxes, rays)
    # the same meshes as one flattened scene (identity maps: same coordinates)
    soup, owner = world_soup(sc)
    ref = orc.brute_closest(soup, rays)
    hit = ref["id"] != 0xFFFFFFFF
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    got = owner[ref["id"][hit]]
    # ids: equal wherever the closest t is unique among the meshes
    same = (inst[hit] == got[:, 0]) & (hits["id"][hit] == got[:, 1])
    assert same.mean() > 0.995, same.mean()
    assert np.all(inst[~hit] == 0xFFFFFFFF)
    # one instance: exactly the one-level traversal of that mesh's tree
    nodes, perm, tri = mesh_tree(sc.primitives[0])
    r1 = random_rays(np.concatenate([tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)]),
                     3000, seed=2)
    one = orc.trace_closest(nodes, perm, tri, r1)
    tn, tp, _ = tlas_tree(boxes[:1])
    h2, i2 = orc.trace_2l(tn, tp, w2o[:1], np.zeros(1, np.uint32), [(nodes, perm, tri)], r1)
    assert h2.tobytes() == one.tobytes()
    assert np.all(i2[one["id"] != 0xFFFFFFFF] == 0)


def test_random_instances_match_world_space_brute_force():
    sc = small_instanced(seed=7, n_inst=30)
    w2o, boxes = host_instances(sc)
    wbox = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])
    rays = random_rays(wbox, 3000, seed=3)
    hits, inst = oracle_trace(sc, w2o, boxes, rays)
    soup, owner = world_soup(sc)
    ref = orc.brute_closest(soup, rays)
    hr = ref["id"] != 0xFFFFFFFF
    hg = hits["id"] != 0xFFFFFFFF
    assert hr.sum() > 300
    assert np.mean(hr == hg) > 0.998
    both = hr & hg
    o = owner[ref["id"][both]]
    same = (inst[both] == o[:, 0]) & (hits["id"][both] == o[:, 1])
    assert same.mean() > 0.995, same.mean()
    rel = np.abs(hits["t"][both] - ref["t"][both]) / np.maximum(ref["t"][both], 1e-3)
    assert np.quantile(rel, 0.99) < 1e-4 and rel.max() < 1e-2
    # any hit agrees with closest hit (same geometry), finite t_max
    occ = oracle_trace(sc, w2o, boxes, rays, closest=False)
    assert np.array_equal(occ.astype(bool), hg)


def test_mistakes_are_caught():
    """A transposed map or a dropped translation fails the brute-force check."""
    sc = small_instanced(seed=9, n_inst=20)
    w2o, boxes = host_instances(sc)
    wbox = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])
    rays = random_rays(wbox, 2000, seed=4)
    soup, owner = world_soup(sc)
    ref = orc.brute_closest(soup, rays)
    m = w2o.reshape(-1, 3, 4)
    bad_t = np.concatenate([np.transpose(m[:, :, :3], (0, 2, 1)), m[:, :, 3:]], axis=2)
    bad_tr = m.copy()
    bad_tr[:, :, 3] = 0
    for bad in (bad_t, bad_tr):
        hits, _ = oracle_trace(sc, bad.reshape(-1, 12), boxes, rays)
        agree = np.mean((hits["id"] != 0xFFFFFFFF) == (ref["id"] != 0xFFFFFFFF))
        assert agree < 0.99, agree


def test_coincident_instances_resolve_to_the_lowest_number():
    """Copies of an instance (same map; the same mesh or the mesh with a far
    triangle appended) tie at every hit on it: the result is the
    lowest-numbered copy, with t, u, v and triangle of the scene without
    copies, whether the copies come first or last. The copies' larger boxes
    put some of them ahead of their original in the top-level visit order,
    so the first-found rule would pick them."""
    sc = small_instanced(seed=7, n_inst=30)
    w2o, boxes = host_instances(sc)
    wbox = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])
    rays = random_rays(wbox, 4000, seed=5)
    base, binst = oracle_trace(sc, w2o, boxes, rays)
    dup = (3, 7, 7, 12, 20)
    hit = base["id"] != 0xFFFFFFFF
    assert (hit & np.isin(binst, dup)).sum() > 50
    for front in (False, True):
        sd, orig = coincident_scene(dup_of=dup, front=front)
        w2, bx = host_instances(sd)
        h, inst = oracle_trace(sd, w2, bx, rays)
        # (a copy with the far triangle has its own tree: the same hit except
        # where a ray meets a shared edge exactly)
        same = h.tobytes() == base.tobytes() or np.mean(h == base) > 0.999
        assert same
        assert np.array_equal(orig[inst[hit]], binst[hit])
        low = np.array([np.flatnonzero(orig == o).min() for o in binst[hit]])
        assert np.array_equal(inst[hit], low)
