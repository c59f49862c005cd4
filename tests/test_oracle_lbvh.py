"""Pins of the oracle's LBVH (Morton codes, sort, Karras hierarchy, refit) and traversal.

The reference has no BVH code (the GPU driver builds it), so these checks pin the
build's own LBVH against hand-computed values, a bit-by-bit Morton interleave,
numpy's stable sort, a brute-force characterisation of the binary radix tree
(every internal node splits its key range at the first differing key bit), exact
box containment, the invariants of the greedy BVH4 collapse (same leaves and boxes
as the BVH2 subtree, full nodes unless only leaves remain, no unexpanded child
larger than an expanded one), and FP64 brute-force closest hits.
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


def check_tree_any(bvh, tri_v):
    """Invariants of any binary BVH over the sorted leaves: every leaf once, parents
    consistent with children, every child box exactly the bounds of the triangles below
    it, at most 63 levels.  Returns (internal-node half-area sum, depth)."""
    children, par = bvh["children"], bvh["parents"]
    n = len(bvh["order"])
    if n == 1:
        return 0.0, 0
    v = tri_v.reshape(-1, 3, 3)
    leaf_lo = v.min(axis=1)[bvh["order"]]
    leaf_hi = v.max(axis=1)[bvh["order"]]
    seen = np.zeros(n, np.int32)
    post, stack, depth = [], [(0, 0)], 0
    while stack:
        nd, d = stack.pop()
        post.append(nd)
        depth = max(depth, d + 1)
        for c in children[nd]:
            if c < 0:
                seen[~c] += 1
            else:
                stack.append((c, d + 1))
    assert np.all(seen == 1) and len(set(post)) == n - 1 == len(post)
    assert par[0] == -1
    for nd in range(n - 1):
        for c in children[nd]:
            assert par[(n - 1) + ~c if c < 0 else c] == nd
    box = {}
    area = 0.0
    for nd in reversed(post):
        los, his = [], []
        for k, c in enumerate(children[nd]):
            lo, hi = (leaf_lo[~c], leaf_hi[~c]) if c < 0 else box[c]
            b = bvh["bounds"][nd, 6 * k:6 * k + 6]
            assert np.array_equal(b[:3], lo) and np.array_equal(b[3:], hi)
            los.append(lo)
            his.append(hi)
        box[nd] = (np.minimum(*los), np.maximum(*his))
        e = (box[nd][1] - box[nd][0]).astype(np.float64)
        area += e[0] * e[1] + e[1] * e[2] + e[2] * e[0]
    assert depth <= 63
    return area, depth


def test_treelet_restructuring_c0(oracle_lib, c0_scene, c0_oracle):
    """The default build restructures the LBVH's treelets (two rounds): a valid tree over
    the same leaves with a lower surface-area cost, more rounds lower it further, and
    the closest hits stay the same bit for bit."""
    raw = oracle_lib.OracleScene(c0_scene)
    raw.set_treelet_rounds(0)
    tri_v = raw.world()[0]
    a0, d0 = check_tree_any(raw.bvh(), tri_v)
    a2, d2 = check_tree_any(c0_oracle.bvh(), tri_v)   # the default: two rounds
    one = oracle_lib.OracleScene(c0_scene)
    one.set_treelet_rounds(1)
    a1, d1 = check_tree_any(one.bvh(), tri_v)
    three = oracle_lib.OracleScene(c0_scene)
    three.set_treelet_rounds(3)
    a3, d3 = check_tree_any(three.bvh(), tri_v)
    print(f"internal half-area sum: LBVH {a0:.1f}, 1 round {a1:.1f}, 2 rounds {a2:.1f}, 3 rounds {a3:.1f}; "
          f"depth {d0} -> {d2}")
    assert a3 <= a2 <= a1 < 0.95 * a0
    assert np.array_equal(raw.bvh()["order"], c0_oracle.bvh()["order"])
    check_bvh4(three.bvh(), *three.bvh4())
    rays = random_rays(np.random.default_rng(5), 2000, np.array([0.0, 2.0, 0.0]), 5.0)
    t0, l0, u0, v0 = raw.trace(rays)
    t3, l3, u3, v3 = three.trace(rays)
    assert np.array_equal(l0, l3) and np.array_equal(t0.view(np.uint32), t3.view(np.uint32))
    assert np.array_equal(u0.view(np.uint32), u3.view(np.uint32))
    assert np.array_equal(raw.trace(rays, any_hit=True)[0], three.trace(rays, any_hit=True)[0])
    assert np.array_equal(c0_oracle.trace(rays)[1], l0)


def all_tree_costs(boxes):
    """Half-area sums of the internal nodes of every binary tree over the given leaf
    boxes (enumerated explicitly: (2k - 3)!! trees for k leaves)."""
    def union(idx):
        b = boxes[list(idx)]
        return np.concatenate([b[:, :3].min(axis=0), b[:, 3:].max(axis=0)])

    def area(b):
        e = (b[3:] - b[:3]).astype(np.float64)
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]

    def trees(items):
        if len(items) == 1:
            yield 0.0
            return
        first, rest = items[0], items[1:]
        own = area(union(items))
        for mask in range(2 ** len(rest)):
            left = (first,) + tuple(r for i, r in enumerate(rest) if (mask >> i) & 1)
            right = tuple(r for i, r in enumerate(rest) if not (mask >> i) & 1)
            if not right:
                continue
            for cl in trees(left):
                for cr in trees(right):
                    yield own + cl + cr

    return list(trees(tuple(range(len(boxes)))))


def test_treelet_root_is_optimal_for_seven_leaves(oracle_lib):
    """With 7 triangles the root's treelet holds every leaf, so the restructured tree
    must reach the minimum over all 10,395 binary trees (explicit enumeration).  The
    triangles sit at halving distances along x, so the Morton order builds a chain
    tall enough (height 6) for the root to be restructured."""
    rng = np.random.default_rng(11)
    for trial in range(4):
        # boxes centred on y = z = 0, so only x reaches the Morton codes
        cx = 3.0 * 0.5 ** np.arange(7) + rng.uniform(0.0, 0.001, 7)
        sz = 0.3 * 0.5 ** np.arange(7) * rng.uniform(0.5, 1.0, 7)
        sy = rng.uniform(0.05, 0.5, 7)
        tri = np.stack([np.stack([cx - sz, -sy, -sz], 1), np.stack([cx + sz, sy, -sz], 1),
                        np.stack([cx, 0.2 * sy, sz], 1)], 1).astype(np.float32).reshape(-1)
        sc = oracle_lib.OracleScene(tiny_scene(tri))
        sc.set_treelet_rounds(0)
        _, lbvh_height = check_tree_any(sc.bvh(), sc.world()[0])
        assert lbvh_height >= 5
        sc.set_treelet_rounds(1)
        v = sc.world()[0].reshape(-1, 3, 3)[sc.bvh()["order"]]
        boxes = np.concatenate([v.min(axis=1), v.max(axis=1)], axis=1)
        costs = all_tree_costs(boxes)
        assert len(costs) == 10395
        got, _ = check_tree_any(sc.bvh(), sc.world()[0])
        sc.set_treelet_rounds(0)
        lbvh, _ = check_tree_any(sc.bvh(), sc.world()[0])
        print(f"trial {trial}: LBVH {lbvh:.4f} (height {lbvh_height}), restructured {got:.4f}, "
              f"optimum {min(costs):.4f}")
        assert abs(got - min(costs)) <= 1e-5 * min(costs)
        assert got <= lbvh + 1e-9


def half_area32(b):
    """Surface measure of boxes (lo.xyz, hi.xyz) in float32, as the collapse orders them."""
    b = np.asarray(b, np.float32)
    dx, dy, dz = b[..., 3] - b[..., 0], b[..., 4] - b[..., 1], b[..., 5] - b[..., 2]
    return (dx * dy + dy * dz) + dz * dx


def check_bvh4(bvh, child4, box4):
    """BVH4 node k collapses BVH2 node k: checked against the BVH2 tree it came from."""
    children, bounds, parents = bvh["children"], bvh["bounds"], bvh["parents"]
    ni = len(children)
    n = len(bvh["order"])

    def box_of(c):
        """Box of BVH2 child c (internal index or ~leaf) as stored in its parent."""
        p = parents[c] if c >= 0 else parents[(n - 1 if n > 1 else 1) + ~c]
        slot = 0 if children[p, 0] == c else 1
        return bounds[p, 6 * slot:6 * slot + 6]

    def leaves2(k):
        out, st = [], [k]
        while st:
            x = st.pop()
            for c in children[x]:
                (out.append(~c) if c < 0 else st.append(c))
        return sorted(out)

    for k in range(ni):
        ch = child4[k]
        lo = box4[k, 0:12].reshape(3, 4).T
        hi = box4[k, 12:24].reshape(3, 4).T
        used = ~np.isinf(lo[:, 0])
        cnt = int(used.sum())
        assert used[:cnt].all() and not used[cnt:].any()
        assert np.all(np.isinf(lo[cnt:])) and np.all(np.isinf(hi[cnt:])) and np.all(ch[cnt:] == 0)
        assert 2 <= cnt <= 4
        if cnt < 4:
            assert np.all(ch[:cnt] < 0)      # stopped early only when every slot is a leaf
        if n == 1:
            continue
        # same leaves as the BVH2 subtree; every slot's box is its BVH2 child box
        got = []
        for j in range(cnt):
            c = int(ch[j])
            got += [~c] if c < 0 else leaves2(c)
            np.testing.assert_array_equal(np.concatenate([lo[j], hi[j]]).view(np.uint32), box_of(c).view(np.uint32))
        assert sorted(got) == leaves2(k)
        # greedy order: no internal slot larger than a BVH2 node that was expanded
        expanded = set()
        for j in range(cnt):
            x = int(ch[j])
            x = parents[x] if x >= 0 else parents[(n - 1) + ~x]
            while x != k:
                expanded.add(x)
                x = parents[x]
        if expanded:
            exp_min = min(half_area32(box_of(e)) for e in expanded)
            for j in range(cnt):
                if ch[j] >= 0:
                    assert half_area32(box_of(int(ch[j]))) <= exp_min


def test_c0_morton_sort_tree(oracle_lib, c0_scene):
    c0_oracle = oracle_lib.OracleScene(c0_scene)
    c0_oracle.set_treelet_rounds(0)   # the Morton/Karras tree as built
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
    check_bvh4(bvh, *c0_oracle.bvh4())


def test_duplicate_codes_and_tiny_trees(oracle_lib):
    rng = np.random.default_rng(3)
    base = rng.uniform(-1, 1, size=(9,))
    # 37 identical triangles plus a few distinct ones: equal Morton codes everywhere
    tris = [base] * 37 + [rng.uniform(-1, 1, size=(9,)) for _ in range(5)]
    sc = oracle_lib.OracleScene(tiny_scene(np.concatenate(tris)))
    sc.set_treelet_rounds(0)
    bvh = sc.bvh()
    check_tree(bvh, sc.world()[0])
    check_bvh4(bvh, *sc.bvh4())
    sc.set_treelet_rounds(2)
    check_tree_any(sc.bvh(), sc.world()[0])
    check_bvh4(sc.bvh(), *sc.bvh4())
    for count in (1, 2, 3, 5, 9):
        sc = oracle_lib.OracleScene(tiny_scene(rng.uniform(-1, 1, size=(9 * count,))))
        check_tree_any(sc.bvh(), sc.world()[0])
        check_bvh4(sc.bvh(), *sc.bvh4())
        sc.set_treelet_rounds(0)
        bvh = sc.bvh()
        if count <= 3:
            check_tree(bvh, sc.world()[0])
        check_bvh4(bvh, *sc.bvh4())


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


def test_anyhit_longest_rays_all_octants(c0_oracle):
    """Shadow rays reach t = 1.7e38 (half of FLT_MAX): tmax * kRobust stays finite,
    so empty BVH4 slots (+inf boxes) are rejected for every direction octant."""
    rng = np.random.default_rng(12)
    rays = random_rays(rng, 4000, np.array([0.0, 2.0, 0.0]), 5.0, tmax=1.70141173e38)
    signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    rays[:, 4:7] = np.abs(rays[:, 4:7]) * signs[np.arange(len(rays)) % 8]
    occ = c0_oracle.trace(rays, any_hit=True)[0]
    t64, tri64 = c0_oracle.trace_brute64(rays)
    agree = occ.astype(bool) == (tri64 != 0xFFFFFFFF)
    assert agree.mean() > 0.999
    closest = c0_oracle.trace(rays)[1]
    assert np.mean((closest != 0xFFFFFFFF) == occ.astype(bool)) > 0.999


def brute_closest32(tri_sorted, rays, kRobust=np.float32(1.00000036)):
    """Closest hit over every triangle with single-precision arithmetic written out
    operation by operation: Moller-Trumbore with division-free acceptance, a candidate
    only where the triangle's bounding box passes the robust slab test up to the hit
    distance, minimum of (t, leaf).  No BVH: the traversal must find the same answer."""
    f = np.float32
    v = tri_sorted.reshape(-1, 3, 3).astype(f)
    v0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    lo = np.minimum(np.minimum(v[:, 0], v[:, 1]), v[:, 2]) + f(0)
    hi = np.maximum(np.maximum(v[:, 0], v[:, 1]), v[:, 2]) + f(0)
    out_t = np.zeros(len(rays), f)
    out_leaf = np.full(len(rays), 0xFFFFFFFF, np.uint32)
    for i, r in enumerate(rays.astype(f)):
        o, tmin, d, tmax = r[0:3], r[3], r[4:7], r[7]
        ad = np.abs(d)
        dd = np.where(ad < f(1e-20), np.where(d < 0, f(-1e-20), f(1e-20)), d)
        inv = (f(1) / dd).astype(f)
        pv = np.stack([d[1] * e2[:, 2] - d[2] * e2[:, 1], d[2] * e2[:, 0] - d[0] * e2[:, 2],
                       d[0] * e2[:, 1] - d[1] * e2[:, 0]], 1)
        det = (e1[:, 0] * pv[:, 0] + e1[:, 1] * pv[:, 1]) + e1[:, 2] * pv[:, 2]
        tv = o - v0
        un = (tv[:, 0] * pv[:, 0] + tv[:, 1] * pv[:, 1]) + tv[:, 2] * pv[:, 2]
        qv = np.stack([tv[:, 1] * e1[:, 2] - tv[:, 2] * e1[:, 1], tv[:, 2] * e1[:, 0] - tv[:, 0] * e1[:, 2],
                       tv[:, 0] * e1[:, 1] - tv[:, 1] * e1[:, 0]], 1)
        vn = (d[0] * qv[:, 0] + d[1] * qv[:, 1]) + d[2] * qv[:, 2]
        pos = (det > 0) & ~((un < 0) | (vn < 0) | (un + vn > det))
        neg = (det < 0) & ~((un > 0) | (vn > 0) | (un + vn < det))
        tn = (e2[:, 0] * qv[:, 0] + e2[:, 1] * qv[:, 1]) + e2[:, 2] * qv[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = tn / det
        x0, x1 = (lo[:, 0] - o[0]) * inv[0], (hi[:, 0] - o[0]) * inv[0]
        y0, y1 = (lo[:, 1] - o[1]) * inv[1], (hi[:, 1] - o[1]) * inv[1]
        z0, z1 = (lo[:, 2] - o[2]) * inv[2], (hi[:, 2] - o[2]) * inv[2]
        entry = np.maximum(np.maximum(np.minimum(x0, x1), np.minimum(y0, y1)), np.maximum(np.minimum(z0, z1), tmin))
        ok = (pos | neg) & (t >= tmin) & (t <= tmax) & (entry <= t * kRobust)
        if ok.any():
            idx = np.nonzero(ok)[0]
            best = idx[np.lexsort((idx, t[idx]))[0]]
            out_t[i], out_leaf[i] = t[best], best
        else:
            out_t[i] = tmax
    return out_t, out_leaf


def test_closest_hit_is_order_independent(oracle_lib, c0_oracle):
    """The BVH traversal returns exactly the brute-force minimum over all candidates --
    on a tie-rich grid (rays aimed at shared vertices and edges) and on C0."""
    grid = scenes.grid_quad((-2, 0, -2), (4, 0, 0), (0, 0, 4), 8, 8, (0, 1, 0))
    tri = grid[0][grid[2]].reshape(-1)
    sc = oracle_lib.OracleScene(tiny_scene(tri))
    ts = sc.world()[0][sc.bvh()["order"]]
    rng = np.random.default_rng(21)
    pts = np.array([(-2 + 0.5 * i, 0.0, -2 + 0.5 * j) for i in range(9) for j in range(9)] +
                   [(-2 + 0.25 + 0.5 * i, 0.0, -2 + 0.5 * j) for i in range(8) for j in range(9)], np.float32)
    o = pts + rng.uniform(-1.5, 1.5, size=pts.shape).astype(np.float32) * np.array([1, 0, 1], np.float32) \
        + np.array([0, 2, 0], np.float32)
    d = pts - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(pts), 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 0.001, d, 100.0
    t, leaf, _, _ = sc.trace(rays)
    bt, bleaf = brute_closest32(ts, rays)
    assert np.array_equal(leaf, bleaf) and np.array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert np.mean(leaf != 0xFFFFFFFF) > 0.9
    rays = random_rays(rng, 300, np.array([0.0, 2.0, 0.0]), 5.0)
    t, leaf, _, _ = c0_oracle.trace(rays)
    bt, bleaf = brute_closest32(c0_oracle.world()[0][c0_oracle.bvh()["order"]], rays)
    assert np.array_equal(leaf, bleaf) and np.array_equal(t.view(np.uint32), bt.view(np.uint32))


def test_traversal_empty_and_single(oracle_lib):
    sc = oracle_lib.OracleScene(tiny_scene(np.array([-1, 0, -1, 1, 0, -1, 0, 0, 1], np.float32)))
    rays = np.array([[0, 1, 0, 0.001, 0, -1, 0, 100], [5, 1, 0, 0.001, 0, -1, 0, 100]], np.float32)
    t, leaf, u, v = sc.trace(rays)
    assert leaf[0] == 0 and abs(t[0] - 1.0) < 1e-6
    assert leaf[1] == 0xFFFFFFFF and t[1] == np.float32(100)
