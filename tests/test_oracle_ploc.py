"""Pins the oracle's PLOC build (ploc_aggregate.cpp:33-556) without running
the reference: closed forms, invariants, an independent numpy replay of the
clustering and brute-force nearest neighbours."""
import numpy as np
import pytest

import orc
from ajisai_render_amd import scenes

F32MAX = np.float32(np.finfo(np.float32).max)


def naive_spread(x):
    out = 0
    for k in range(10):
        out |= ((x >> k) & 1) << (3 * k)
    return out


def test_morton_split_equals_naive_interleave_all_values():
    lib = orc.load()
    for x in range(1024):
        assert lib.orc_morton_split(x) == naive_spread(x)


def test_morton_encode_interleaves_xyz():
    lib = orc.load()
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, (500, 3)):
        x, y, z = int(x), int(y), int(z)
        assert lib.orc_morton_encode(x, y, z) == naive_spread(x) | (naive_spread(y) << 1) | (
            naive_spread(z) << 2)


def deinterleave(code):
    out = [0, 0, 0]
    for k in range(10):
        for a in range(3):
            out[a] |= ((int(code) >> (3 * k + a)) & 1) << k
    return out


def test_morton_grid_cells_known_answers():
    # Scene box [0,1]^3: world2grid = 1024, offset = 0; a triangle whose box
    # centre is the centre of cell (i, j, k) must encode (i, j, k); centres
    # on the max face clamp to 1023, below the min clamp to 0.
    cells = [(0, 0, 0), (1, 2, 3), (511, 512, 1), (1023, 1023, 1023), (7, 1000, 64)]
    tris = []
    for (i, j, k) in cells:
        c = (np.array([i, j, k], np.float64) + 0.5) / 1024.0
        h = 0.25 / 1024.0
        tris.append([c[0] - h, c[1] - h, c[2] - h, c[0] + h, c[1] - h, c[2] + h,
                     c[0] - h, c[1] + h, c[2] + h])
    tris.append([1, 1, 1, 1, 1, 1, 1, 1, 1])          # centre on max corner
    tris.append([0, 0, 0, 0, 0, 0, 0, 0, 0])          # centre on min corner
    tri = np.array(tris, np.float32)
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    codes = orc.morton_codes(tri, box)
    for c, cell in zip(codes[:len(cells)], cells):
        assert deinterleave(c) == list(cell)
    assert deinterleave(codes[-2]) == [1023, 1023, 1023]
    assert deinterleave(codes[-1]) == [0, 0, 0]


def test_radix_sort_equals_stable_argsort():
    rng = np.random.default_rng(2)
    for n in (1, 2, 17, 1000, 5000):
        keys = rng.integers(0, 1 << 30, n).astype(np.uint32)
        keys[rng.integers(0, n, n // 3)] = keys[0]  # many duplicates
        vals = np.arange(n, dtype=np.uint32)
        k, v = orc.radix_sort(keys, vals)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(v, order)
        assert np.array_equal(k, keys[order])


def np_join(a_mn, a_mx, b_mn, b_mx):
    """Join (bounds.h:108-113) on single boxes, compare-select min/max."""
    if np.all(a_mn == a_mx):
        return b_mn.copy(), b_mx.copy()
    if np.all(b_mn == b_mx):
        return a_mn.copy(), a_mx.copy()
    mn = np.where(b_mn < a_mn, b_mn, a_mn)
    mx = np.where(a_mx < b_mx, b_mx, a_mx)
    return mn.astype(np.float32), mx.astype(np.float32)


def np_area(mn, mx):
    d = (mx - mn).astype(np.float32)
    return np.float32(2) * np.float32(np.float32(d[0] * d[1]) + np.float32(d[0] * d[2])
                                      + np.float32(d[1] * d[2]))


def tri_boxes(tri):
    t = tri.reshape(-1, 3, 3)
    return t.min(axis=1).astype(np.float32), t.max(axis=1).astype(np.float32)


def check_tree(nodes, perm, tri):
    n = tri.shape[0]
    assert nodes.shape[0] == 2 * n - 1
    assert sorted(perm.tolist()) == list(range(n))
    seen_leaf = np.zeros(n, np.int64)
    parent_count = np.zeros(2 * n - 1, np.int64)
    tmn, tmx = tri_boxes(tri)
    for k, nd in enumerate(nodes):
        if nd["count"]:
            assert nd["count"] == 1
            seen_leaf[nd["first"]] += 1
            p = perm[nd["first"]]
            assert np.array_equal(nd["mn"], tmn[p]) and np.array_equal(nd["mx"], tmx[p])
        else:
            c0, c1 = int(nd["first"]), int(nd["first"]) + 1
            assert c0 > k and c1 < 2 * n - 1
            parent_count[c0] += 1
            parent_count[c1] += 1
            # merged bound = Join(input[j], input[i]) with children (i, j)
            mn, mx = np_join(nodes[c1]["mn"], nodes[c1]["mx"], nodes[c0]["mn"], nodes[c0]["mx"])
            assert np.array_equal(mn, nd["mn"]) and np.array_equal(mx, nd["mx"])
    assert np.all(seen_leaf == 1)
    assert parent_count[0] == 0 and np.all(parent_count[1:] == 1)


def np_ploc(tri, box, r):
    """Independent numpy replay of PLOCAggregate::Build's clustering."""
    n = tri.shape[0]
    codes = orc.morton_codes(tri, box)
    perm = np.argsort(codes, kind="stable").astype(np.uint32)
    tmn, tmx = tri_boxes(tri)
    nn_ = 2 * n - 1
    nodes = np.zeros(nn_, orc.NODE_DTYPE)
    begin, end, pre_end = nn_ - n, nn_, nn_
    for i in range(n):
        nodes[begin + i] = (tmn[perm[i]], tmx[perm[i]], 1, i)
    cur = nodes
    while end - begin > 1:
        out = cur.copy()
        nb = {}
        for i in range(begin, end):
            best, bj = F32MAX, None
            for j in range(max(begin, i - r), min(i + r + 1, end)):
                if j == i:
                    continue
                lo, hi = (j, i) if j < i else (i, j)
                d = np_area(*np_join(cur[lo]["mn"], cur[lo]["mx"], cur[hi]["mn"], cur[hi]["mx"]))
                if d < best:
                    best, bj = d, j
            nb[i] = bj
        flags = [1 if (nb[i] is not None and nb[nb[i]] == i and i < nb[i]) else 0
                 for i in range(begin, end)]
        m = np.cumsum(flags)
        mc = int(m[-1])
        cc = 2 * mc
        ub = end - (cc + (end - begin - mc))
        cb = end - cc
        for i in range(begin, end):
            j = nb[i]
            mi = m[i - begin]
            if j is not None and nb[j] == i:
                if i < j:
                    mj = m[j - begin]
                    first = cb + (mi - 1) * 2
                    mn, mx = np_join(cur[j]["mn"], cur[j]["mx"], cur[i]["mn"], cur[i]["mx"])
                    out[ub + j - begin - mj] = (mn, mx, 0, first)
                    out[first] = cur[i]
                    out[first + 1] = cur[j]
            else:
                out[ub + i - begin - mi] = cur[i]
        cur = out
        pre_end, begin, end = end, ub, cb
    return cur, perm


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (7, 3), (64, 4), (300, 5)])
def test_ploc_oracle_matches_numpy_replay(n, seed):
    tri = scenes.random_soup(n, seed, extent=1.0, size=0.08)
    box = np.concatenate([tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)]).astype(np.float32)
    nodes, perm, it = orc.ploc_build(tri, box, 14)
    assert it >= 0
    ref_nodes, ref_perm = np_ploc(tri, box, 14)
    assert np.array_equal(perm, ref_perm)
    assert nodes.tobytes() == ref_nodes.tobytes()
    check_tree(nodes, perm, tri)


@pytest.mark.parametrize("radius", [1, 3, 14])
def test_ploc_oracle_invariants_random(radius):
    tri = scenes.random_soup(3000, 11, extent=2.0, size=0.05)
    box = np.concatenate([tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)]).astype(np.float32)
    nodes, perm, it = orc.ploc_build(tri, box, radius)
    assert it > 0
    check_tree(nodes, perm, tri)
    # leaves in Morton order
    codes = orc.morton_codes(tri, box)
    assert np.all(np.diff(codes[perm].astype(np.int64)) >= 0)


def test_ploc_oracle_duplicates_and_ties():
    # many identical triangles (all pair distances tie) still build a tree
    base = scenes.random_soup(20, 5, size=0.1)
    tri = np.repeat(base, 10, axis=0)
    box = np.concatenate([tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)]).astype(np.float32)
    nodes, perm, it = orc.ploc_build(tri, box, 14)
    assert it > 0
    check_tree(nodes, perm, tri)


def test_nearest_neighbour_brute_force_with_ties():
    rng = np.random.default_rng(7)
    n = 200
    nodes = np.zeros(n, orc.NODE_DTYPE)
    mn = rng.integers(0, 4, (n, 3)).astype(np.float32)   # integer boxes: many ties
    nodes["mn"] = mn
    nodes["mx"] = mn + rng.integers(1, 3, (n, 3)).astype(np.float32)
    nodes["count"] = 1
    for r in (1, 2, 5, 14):
        got = orc.nearest(nodes, 0, n, r)
        for i in range(n):
            ds = []
            for j in range(max(0, i - r), min(n, i + r + 1)):
                if j == i:
                    continue
                lo, hi = (j, i) if j < i else (i, j)
                ds.append((np_area(*np_join(nodes[lo]["mn"], nodes[lo]["mx"],
                                            nodes[hi]["mn"], nodes[hi]["mx"])), j))
            best = min(d for d, _ in ds)
            assert got[i] == min(j for d, j in ds if d == best)


def test_cornell_scene_flatten_and_tree():
    sc = scenes.cornell()
    o = orc.OracleScene(sc)
    assert o.n_tri == 34
    tri = o.tris()
    assert np.array_equal(tri, sc.soup())
    check_tree(o.nodes(), o.perm(), tri)
