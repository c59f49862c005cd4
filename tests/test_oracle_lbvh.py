"""Pins of the oracle's LBVH (Morton codes, sort, Karras hierarchy, refit) and traversal.

The reference has no BVH code (the GPU driver builds it), so these checks pin the
build's own LBVH against hand-computed values, a bit-by-bit Morton interleave,
numpy's stable sort, a brute-force characterisation of the binary radix tree
(every internal node splits its key range at the first differing key bit), exact
box containment, and FP64 brute-force closest hits.
"""
import numpy as np
import pytest

import scenes


def interleave_bits(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_morton_known_answers(oracle_lib):
    assert oracle_lib.morton3(0, 0, 0) == 0
    assert oracle_lib.morton3(1, 0, 0) == 0b100
    assert oracle_lib.morton3(0, 1, 0) == 0b010
    assert oracle_lib.morton3(0, 0, 1) == 0b001
    assert oracle_lib.morton3(3, 0, 0) == 0b100100
    assert oracle_lib.morton3(1023, 1023, 1023) == (1 << 30) - 1
    assert oracle_lib.morton3(1023, 0, 0) == 0b100100100100100100100100100100
    rng = np.random.default_rng(5)
    for x, y, z in rng.integers(0, 1024, size=(500, 3)):
        assert oracle_lib.morton3(int(x), int(y), int(z)) == interleave_bits(int(x), int(y), int(z))


def deinterleave(codes):
    q = np.zeros((len(codes), 3), np.int64)
    c = codes.astype(np.int64)
    for b in range(10):
        q[:, 0] |= ((c >> (3 * b + 2)) & 1) << b
        q[:, 1] |= ((c >> (3 * b + 1)) & 1) << b
        q[:, 2] |= ((c >> (3 * b)) & 1) << b
    return q


def tiny_scene(tri_v, two_sided=True):
    """Scene with one mesh holding the given world triangles (identity instance)."""
    v = np.asarray(tri_v, np.float32).reshape(-1, 3)
    n = np.tile(np.array([0, 1, 0], np.float32), (len(v), 1))
    idx = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky()
    return scenes.Scene("tiny", [(v, n, idx)], inst, scenes.material((0.5, 0.5, 0.5), 0.5, 0.0, two_sided)[None],
                        scenes.dir_light((1, 1, 1), (0, 1, 0))[None], sky, col, np.eye(3, dtype=np.float32).reshape(-1))


def check_tree(bvh, tri_v):
    """Brute-force characterisation of a Karras radix tree over (code, index) keys."""
    codes = bvh["morton"][bvh["order"]].astype(np.uint64)
    n = len(codes)
    keys = (codes << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    children = bvh["children"]
    if n == 1:
        assert list(children[0]) == [-1, -1]
        return
    # walk from the root: every leaf reached exactly once, ranges contiguous
    seen = np.zeros(n, np.int32)
    ranges = {}
    stack = [0]
    order = []
    while stack:
        nd = stack.pop()
        order.append(nd)
        for c in children[nd]:
            if c < 0:
                seen[~c] += 1
            else:
                stack.append(c)
    assert np.all(seen == 1)
    assert len(order) == n - 1 and len(set(order)) == n - 1
    for nd in reversed(order):
        lo = hi = None
        parts = []
        for c in children[nd]:
            r = (~c, ~c) if c < 0 else ranges[c]
            parts.append(r)
        (a0, a1), (b0, b1) = parts
        assert a1 + 1 == b0, "left range must end right before the right range"
        ranges[nd] = (a0, b1)
        # the split lies at the first key bit where the range's keys differ
        span = keys[a0:b1 + 1]
        xor = span[:-1] ^ span[1:]
        highest = np.array([int(x).bit_length() for x in xor])
        split = int(np.argmax(highest))
        assert highest[split] == highest.max()
        assert a0 + split == a1, "split must be at the highest differing bit"
        assert np.sum(highest == highest.max()) == 1
    assert ranges[0] == (0, n - 1)
    # parents are consistent with children
    par = bvh["parents"]
    assert par[0] == -1
    for nd in range(n - 1):
        for c in children[nd]:
            assert par[(n - 1) + ~c if c < 0 else c] == nd
    # refit: every child box is exactly the bounds of the triangles below it
    v = tri_v.reshape(-1, 3, 3)
    leaf_lo = v.min(axis=1)[bvh["order"]]
    leaf_hi = v.max(axis=1)[bvh["order"]]
    pre_lo = np.minimum.reduceat
    for nd in range(n - 1):
        for k, c in enumerate(children[nd]):
            r = (~c, ~c) if c < 0 else ranges[c]
            lo = leaf_lo[r[0]:r[1] + 1].min(axis=0)
            hi = leaf_hi[r[0]:r[1] + 1].max(axis=0)
            box = bvh["bounds"][nd, 6 * k:6 * k + 6]
            assert np.array_equal(box[:3], lo) and np.array_equal(box[3:], hi)


def test_c0_morton_sort_tree(c0_oracle):
    bvh = c0_oracle.bvh()
    tri_v, _, _ = c0_oracle.world()
    n = c0_oracle.n
    assert n == 20504
    # stable sort by code
    assert np.array_equal(bvh["order"], np.argsort(bvh["morton"], kind="stable").astype(np.uint32))
    # quantisation: decoded grid cells agree with a float64 computation to within one cell
    v = tri_v.reshape(-1, 3, 3).astype(np.float64)
    ce = (v.min(axis=1) + v.max(axis=1)) * 0.5
    lo, hi = ce.min(axis=0), ce.max(axis=0)
    np.testing.assert_array_equal(bvh["centroid_bounds"], np.concatenate([lo, hi]).astype(np.float32))
    q64 = np.clip(np.floor((ce - lo) / (hi - lo) * 1024), 0, 1023)
    q = deinterleave(bvh["morton"])
    assert np.max(np.abs(q - q64)) <= 1
    assert np.mean(q == q64) > 0.999
    check_tree(bvh, tri_v)


def test_duplicate_codes_and_tiny_trees(oracle_lib):
    rng = np.random.default_rng(3)
    base = rng.uniform(-1, 1, size=(9,))
    # 37 identical triangles plus a few distinct ones: equal Morton codes everywhere
    tris = [base] * 37 + [rng.uniform(-1, 1, size=(9,)) for _ in range(5)]
    sc = oracle_lib.OracleScene(tiny_scene(np.concatenate(tris)))
    bvh = sc.bvh()
    check_tree(bvh, sc.world()[0])
    for count in (1, 2, 3):
        sc = oracle_lib.OracleScene(tiny_scene(rng.uniform(-1, 1, size=(9 * count,))))
        bvh = sc.bvh()
        check_tree(bvh, sc.world()[0])


def random_rays(rng, n, centre, radius, tmax=1000.0):
    o = centre + rng.uniform(-radius, radius, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def test_traversal_matches_fp64_brute_force(c0_oracle):
    rng = np.random.default_rng(11)
    rays = random_rays(rng, 3000, np.array([0.0, 2.0, 0.0]), 5.0)
    t, leaf, u, v = c0_oracle.trace(rays)
    t64, tri64 = c0_oracle.trace_brute64(rays)
    order = c0_oracle.bvh()["order"]
    tri32 = np.where(leaf == 0xFFFFFFFF, 0xFFFFFFFF, order[np.minimum(leaf, len(order) - 1)])
    agree = tri32 == tri64
    assert agree.mean() > 0.999
    hit = (tri64 != 0xFFFFFFFF) & agree
    np.testing.assert_allclose(t[hit], t64[hit], rtol=2e-5, atol=2e-6)
    assert np.all((u[hit] >= 0) & (v[hit] >= 0) & (u[hit] + v[hit] <= 1 + 1e-6))
    # any-hit with tmax = closest distance * 0.999 must be free; * 1.001 occluded
    occ = c0_oracle.trace(rays, any_hit=True)[0]
    np.testing.assert_array_equal(occ.astype(bool), tri32 != 0xFFFFFFFF)
    short = rays.copy()
    sel = hit & (t64 > 0.1)
    short[:, 7] = np.where(sel, t64 * 0.999, 1000.0)
    occ_short = c0_oracle.trace(short[sel], any_hit=True)[0]
    assert occ_short.mean() < 0.01


def test_traversal_empty_and_single(oracle_lib):
    sc = oracle_lib.OracleScene(tiny_scene(np.array([-1, 0, -1, 1, 0, -1, 0, 0, 1], np.float32)))
    rays = np.array([[0, 1, 0, 0.001, 0, -1, 0, 100], [5, 1, 0, 0.001, 0, -1, 0, 100]], np.float32)
    t, leaf, u, v = sc.trace(rays)
    assert leaf[0] == 0 and abs(t[0] - 1.0) < 1e-6
    assert leaf[1] == 0xFFFFFFFF and t[1] == np.float32(100)
