"""Pins for the oracle itself (CPU only).

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned by mathematics that a plausible mistake would violate:
closed-form intersections, brute force over every triangle, structural
invariants of the builder, and an FP64 recompute.
"""
import numpy as np
import pytest

import oracle_rt
import rt_cases
from atlas_rt import workloads as W


# ---------------------------------------------------------------------------
# Triangle test (intersections.hsh:36-58, Ray.cpp:107-130): closed forms
# ---------------------------------------------------------------------------
TRI = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


def test_triangle_interior_exact():
    # hit point (0.25, 0.5, 0) from z = 1 straight down: t = 1, u = 0.25, v = 0.5
    hit, sol = oracle_rt.triangle_test([0.25, 0.5, 1.0], [0, 0, -1], *TRI)
    assert hit and tuple(sol) == (1.0, 0.25, 0.5)
    # same hit at t = 4 with a direction of length 1/4 ... t scales with |d|
    hit, sol = oracle_rt.triangle_test([0.25, 0.5, 1.0], [0, 0, -0.25], *TRI)
    assert hit and tuple(sol) == (4.0, 0.25, 0.5)


def test_triangle_u_v_are_weights_of_v1_v2():
    # u multiplies e0 = v1 - v0, v multiplies e1 = v2 - v0 (a transposed
    # operand would swap them)
    hit, sol = oracle_rt.triangle_test([0.125, 0.5, 2.0], [0, 0, -1], [0, 0, 0], [2, 0, 0], [0, 4, 0])
    assert hit and tuple(sol) == (2.0, 0.0625, 0.125)


@pytest.mark.parametrize("p,expect", [
    ([0.0, 0.0], (0.0, 0.0)),    # vertex v0
    ([1.0, 0.0], (1.0, 0.0)),    # vertex v1
    ([0.0, 1.0], (0.0, 1.0)),    # vertex v2
    ([0.5, 0.5], (0.5, 0.5)),    # hypotenuse: u + v = 1 accepted
    ([0.5, 0.0], (0.5, 0.0)),    # edge v = 0
])
def test_triangle_boundary_accepted(p, expect):
    hit, sol = oracle_rt.triangle_test([p[0], p[1], 3.0], [0, 0, -1], *TRI)
    assert hit and sol[0] == 3.0 and (sol[1], sol[2]) == expect


def test_triangle_outside_rejected():
    for p in ([0.5, 0.50001], [-1e-7, 0.5], [0.5, -1e-7], [1.2, 0.1]):
        hit, _ = oracle_rt.triangle_test([p[0], p[1], 1.0], [0, 0, -1], *TRI)
        assert not hit, p


def test_triangle_backface_and_behind():
    # reversed winding is still hit (no culling; BACKFACE_CULLING is never defined)
    hit, sol = oracle_rt.triangle_test([0.25, 0.25, -1.0], [0, 0, 1], *TRI)
    assert hit and sol[0] == 1.0
    # a triangle behind the origin has t < 0 and is rejected
    hit, sol = oracle_rt.triangle_test([0.25, 0.25, -1.0], [0, 0, -1], *TRI)
    assert not hit and sol[0] == -1.0


def test_triangle_parallel_and_degenerate_rejected():
    # ray in the triangle's plane: q.e0 = 0 -> inf/nan, never accepted
    hit, _ = oracle_rt.triangle_test([-1.0, 0.25, 0.0], [1, 0, 0], *TRI)
    assert not hit
    # zero-area triangle
    hit, _ = oracle_rt.triangle_test([0.5, 0.0, 1.0], [0, 0, -1], [0, 0, 0], [1, 0, 0], [2, 0, 0])
    assert not hit
    # zero direction
    hit, _ = oracle_rt.triangle_test([0.25, 0.25, 1.0], [0, 0, 0], *TRI)
    assert not hit


def test_triangle_matches_fp64_on_random():
    """Against the same formula evaluated in float64 on the same float32
    inputs, within the forward error bound of float32 dot products:
    |x - x64| <= 64 eps |a||b| / |q.e0| for x = (a.b) / (q.e0).  A dropped
    term, a sign error or a swapped operand is off by O(1)."""
    import fp64
    g = np.random.default_rng(0)
    verts = W.triangle_soup(2000, W.rng(0, 99))
    rays = np.zeros(2000, dtype=oracle_rt.RAY_DTYPE)
    sols = np.zeros((2000, 3), np.float32)
    hits = np.zeros(2000, bool)
    for k in range(2000):
        tri = verts[k].reshape(3, 3).astype(np.float64)
        target = g.dirichlet([1, 1, 1]) @ tri
        o = g.uniform(-12, 12, 3)
        d = (target - o) / np.linalg.norm(target - o)
        rays[k]["origin"], rays[k]["direction"] = o, d
        hits[k], sols[k] = oracle_rt.triangle_test(rays[k]["origin"], rays[k]["direction"], *verts[k].reshape(3, 3))
    assert hits.sum() > 1900
    t64, u64, v64, cond = fp64.recompute(verts, rays, np.arange(2000))
    tri = verts.reshape(-1, 3, 3).astype(np.float64)
    o, d = rays["origin"].astype(np.float64), rays["direction"].astype(np.float64)
    e0, e1, s = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], o - tri[:, 0]
    p, q = np.cross(s, e0), np.cross(d, e1)
    den = np.abs(np.einsum("ij,ij->i", q, e0))
    nrm = lambda x: np.linalg.norm(x, axis=1)
    eps = 2.0 ** -24
    bound_t = 64 * eps * nrm(p) * nrm(e1) / den
    bound_u = 64 * eps * nrm(q) * nrm(s) / den
    bound_v = 64 * eps * nrm(p) * nrm(d) / den
    h = hits
    assert np.all(np.abs(sols[h, 0] - t64[h]) <= bound_t[h])
    assert np.all(np.abs(sols[h, 1] - u64[h]) <= bound_u[h])
    assert np.all(np.abs(sols[h, 2] - v64[h]) <= bound_v[h])


# ---------------------------------------------------------------------------
# Slab test (intersections.hsh:18-34): closed forms
# ---------------------------------------------------------------------------
BOX = ([0, 0, 0], [1, 1, 1])


def test_slab_entry_exit():
    hit, dist = oracle_rt.slab_test([-1, 0.5, 0.5], [1, 0, 0], *BOX, 0.0, 1e12)
    assert hit and dist == 1.0
    hit, dist = oracle_rt.slab_test([-1, 0.5, 0.5], [0.5, 0, 0], *BOX, 0.0, 1e12)
    assert hit and dist == 2.0  # distance in units of |d|
    # diagonal: enters through the corner region at t = 1 on all axes
    hit, dist = oracle_rt.slab_test([-1, -1, -1], [1, 1, 1], *BOX, 0.0, 1e12)
    assert hit and dist == 1.0


def test_slab_inside_and_limits():
    hit, dist = oracle_rt.slab_test([0.5, 0.5, 0.5], [0, 1, 0], *BOX, 0.25, 1e12)
    assert hit and dist == 0.25  # origin inside: entry clamps to tmin
    hit, dist = oracle_rt.slab_test([-1, 0.5, 0.5], [1, 0, 0], *BOX, 0.0, 0.5)
    assert not hit and dist == 0.5  # box beyond tmax: miss, dist = tmax
    hit, dist = oracle_rt.slab_test([-1, 0.5, 0.5], [-1, 0, 0], *BOX, 0.0, 1e12)
    assert not hit  # pointing away


def test_slab_axis_parallel_infinite_reciprocal():
    # zero direction components give +-inf slab distances
    hit, dist = oracle_rt.slab_test([0.5, -1, 0.5], [0, 1, 0], *BOX, 0.0, 1e12)
    assert hit and dist == 1.0
    hit, _ = oracle_rt.slab_test([2.0, -1, 0.5], [0, 1, 0], *BOX, 0.0, 1e12)
    assert not hit  # outside the x slab: (0-2)/0 and (1-2)/0 are both -inf
    # origin exactly on a face with a zero component: 0/0 = NaN, which IEEE
    # minNum/maxNum drop, leaving +inf -> the grazing ray misses
    hit, _ = oracle_rt.slab_test([0.0, -1, 0.5], [0, 1, 0], *BOX, 0.0, 1e12)
    assert not hit


# ---------------------------------------------------------------------------
# Builder (BVH.cpp:14-869): structural invariants
# ---------------------------------------------------------------------------
def _area(lo, hi):
    d = hi.astype(np.float64) - lo.astype(np.float64)
    return 2 * (d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2] + d[..., 2] * d[..., 0])


def check_bvh_invariants(verts, ob):
    nodes, ref, eon = ob.nodes, ob.ref_idx, ob.end_of_node
    n = len(verts)
    # every input triangle is referenced at least once
    assert np.array_equal(np.unique(ref), np.arange(n))
    # leaves below the root hold exactly one reference; #nodes = #refs - 1
    leaf_ptrs = np.concatenate([nodes["left_ptr"][nodes["left_ptr"] < 0], nodes["right_ptr"][nodes["right_ptr"] < 0]])
    first = ~leaf_ptrs
    assert np.all(eon[first] == 1)
    assert len(np.unique(first)) == len(ref) and eon.sum() == len(ref)
    assert len(nodes) == len(ref) - 1
    # pre-order: children come after their parent; the larger-area child is left
    for side in ("left_ptr", "right_ptr"):
        inner = nodes[side] >= 0
        assert np.all(nodes[side][inner] > np.nonzero(inner)[0])
    al = _area(nodes["left_min"], nodes["left_max"])
    ar = _area(nodes["right_min"], nodes["right_max"])
    assert np.all(al.astype(np.float32) >= ar.astype(np.float32))
    # every inner child box contains the boxes of its own children
    for side in ("left", "right"):
        ptr = nodes[f"{side}_ptr"]
        for i in np.nonzero(ptr >= 0)[0][:2000]:
            c = nodes[ptr[i]]
            lo = np.minimum(c["left_min"], c["right_min"])
            hi = np.maximum(c["left_max"], c["right_max"])
            assert np.all(nodes[i][f"{side}_min"] <= lo) and np.all(nodes[i][f"{side}_max"] >= hi)
    # the union of a triangle's leaf boxes covers the triangle
    tri = verts.reshape(-1, 3, 3)
    box_lo = np.full((n, 3), np.inf, np.float32)
    box_hi = np.full((n, 3), -np.inf, np.float32)
    for side in ("left", "right"):
        ptr = nodes[f"{side}_ptr"]
        m = ptr < 0
        slot = ~ptr[m]
        t = ref[slot]
        np.minimum.at(box_lo, t, nodes[f"{side}_min"][m])
        np.maximum.at(box_hi, t, nodes[f"{side}_max"][m])
    assert np.all(box_lo <= tri.min(axis=1)) and np.all(box_hi >= tri.max(axis=1))


@pytest.mark.parametrize("n,salt", [(4, 1), (100, 2), (4096, 3), (32768, 4)])
def test_builder_invariants(n, salt):
    verts, _ = rt_cases.soup(n, 1, salt)
    check_bvh_invariants(verts, oracle_rt.build_triangles(verts))


def test_builder_invariants_heightfield():
    verts = W.heightfield(grid=96)
    check_bvh_invariants(verts, oracle_rt.build_triangles(verts))


def test_builder_root_spatial_split_duplicates():
    # long thin triangles spanning the scene force SBVH reference duplication
    g = np.random.default_rng(5)
    n = 512
    verts = np.zeros((n, 9), np.float32)
    y = g.uniform(-5, 5, n)
    z = g.uniform(-5, 5, n)
    verts[:, 0] = -50
    verts[:, 1] = y
    verts[:, 2] = z
    verts[:, 3] = 50
    verts[:, 4] = y + 0.01
    verts[:, 5] = z
    verts[:, 6] = 0
    verts[:, 7] = y
    verts[:, 8] = z + 0.01
    ob = oracle_rt.build_triangles(verts)
    check_bvh_invariants(verts, ob)


def test_tlas_builder_single_instance_special_node():
    # BVH.cpp:80-88: one instance -> one node with both pointers ~0, the right
    # box collapsed to the origin; the reference list holds the instance twice
    ob = oracle_rt.build_aabbs(np.array([[1, 2, 3, 4, 5, 6]], np.float32))
    assert len(ob.nodes) == 1
    nd = ob.nodes[0]
    assert nd["left_ptr"] == -1 and nd["right_ptr"] == -1
    assert tuple(nd["left_min"]) == (1, 2, 3) and tuple(nd["right_max"]) == (0, 0, 0)
    assert list(ob.ref_idx) == [0, 0] and list(ob.end_of_node) == [0, 1]


def test_tlas_builder_invariants():
    g = np.random.default_rng(7)
    c = g.uniform(-100, 100, (256, 3))
    s = g.uniform(1, 5, (256, 3))
    boxes = np.concatenate([c - s, c + s], axis=1).astype(np.float32)
    ob = oracle_rt.build_aabbs(boxes)
    assert len(ob.nodes) == 255 and np.array_equal(np.sort(ob.ref_idx), np.arange(256))


# ---------------------------------------------------------------------------
# Traversal: brute force pins the closest hit independently of the tree
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n_tris,n_rays,salt", [(1, 2000, 1), (3, 2000, 2), (64, 20000, 3), (1024, 50000, 4),
                                                (4096, 65536, 5)])
def test_rgpu_closest_equals_brute_force(n_tris, n_rays, salt):
    verts, rays = rt_cases.soup(n_tris, n_rays, salt)
    if n_tris >= 64:
        rays = np.concatenate([rays, rt_cases.edge_rays(verts, n_random=2048, salt=salt)])
    ob = oracle_rt.build_triangles(verts)
    h, s, rt = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, raw_t=True)
    hb, sb, rtb = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, mode=oracle_rt.MODE_BRUTE, raw_t=True)
    # the engine's float32 distances that drove both searches
    h["t"], hb["t"] = rt, rtb
    tdiff = h["t"].view(np.int32) != hb["t"].view(np.int32)
    # The engine's traversal is not watertight: a hit exactly on a triangle's
    # edge or vertex lies on the face of the leaf box, and the slab test's
    # rounding can cull that box.  Random rays never do this; rays aimed at
    # vertices/edges may.  Every t difference must be such a boundary hit
    # that the traversal missed (never a different closer hit).
    assert not tdiff[:n_rays].any()
    m = tdiff
    assert np.all(hb["t"][m] < h["t"][m])
    w = np.minimum(np.minimum(hb["u"][m], hb["v"][m]), 1.0 - hb["u"][m] - hb["v"][m])
    assert np.all(w < 2e-4), w  # float32 barycentrics from ~20 units away carry ~1e-5..1e-4 error
    assert m.sum() <= max(2, len(rays) // 500)
    diff = (h["prim_id"] != hb["prim_id"]) & ~tdiff
    # remaining id differences are equal-t ties (t equal), resolved by
    # traversal order vs slot order
    assert diff.sum() <= max(2, len(rays) // 5000)
    same = ~diff & ~tdiff
    assert np.array_equal(h["u"][same].view(np.int32), hb["u"][same].view(np.int32))


def test_rgpu_any_equals_brute_existence():
    verts, rays = rt_cases.soup(2048, 30000, 9)
    rays["tmax"] = np.random.default_rng(1).uniform(0.5, 30.0, len(rays)).astype(np.float32)
    ob = oracle_rt.build_triangles(verts)
    anyh = oracle_rt.trace_any(ob.nodes, ob.tris, rays)
    hb, _ = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, mode=oracle_rt.MODE_BRUTE)
    assert np.array_equal(anyh, (hb["prim_id"] >= 0).astype(np.uint32))


def test_rgpu_nan_direction_is_a_miss():
    verts, _ = rt_cases.soup(256, 1, 3)
    ob = oracle_rt.build_triangles(verts)
    rays = rt_cases.make_rays([[0, 0, 0]], [[np.nan, 0, 1]], tmax=123.0)
    h, s = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays)
    assert h["prim_id"][0] == -1 and h["t"][0] == 123.0 and s[0] == -1


def test_rcpu_vs_rgpu_disagreements_are_tmin_or_ties():
    """R-CPU (BVH.cpp:103-174) ignores tmin for triangles and multiplies by a
    reciprocal in the slab test; with tmin = 0 the two semantics agree except
    at equal-t ties."""
    verts, rays = rt_cases.soup(8192, 65536, 13)
    ob = oracle_rt.build_triangles(verts)
    rays0 = rays.copy()
    rays0["tmin"] = 0.0
    _, _, tg = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays0, raw_t=True)
    _, _, tc = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays0, mode=oracle_rt.MODE_RCPU, raw_t=True)
    assert np.array_equal(tg.view(np.int32), tc.view(np.int32))
    # with tmin = 1e-4 R-CPU can accept hits in [0, 1e-4]
    _, _, tg1 = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, raw_t=True)
    _, _, tc1 = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, mode=oracle_rt.MODE_RCPU, raw_t=True)
    d = tg1 != tc1
    assert np.all(tc1[d] <= 1e-4)


def _edge_aimed_rays(verts, n, salt):
    """Rays from above aimed at the midpoints of triangle edges, where two
    triangles of a grid mesh meet and may report the same float32 t."""
    g = np.random.default_rng(salt)
    tri = verts.reshape(-1, 3, 3).astype(np.float64)
    pick = g.integers(0, len(tri), n)
    k = g.integers(0, 3, n)
    a, b = tri[pick, k], tri[pick, (k + 1) % 3]
    tgt = 0.5 * (a + b)
    o = tgt + np.column_stack([g.uniform(-6, 6, n), g.uniform(15, 40, n), g.uniform(-6, 6, n)])
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rt_cases.make_rays(o.astype(np.float32), d.astype(np.float32))


def test_oracle_tie_flags_match_brute_force():
    """The tie flag marks a hit whose float32 t another tested slot also
    reaches (the traversal order then picks the slot).  Brute force over every
    slot except the winner's finds such a rival independently of the tree:
    every flagged ray has one, and a rival the flag misses is one that the
    traversal dropped at a box boundary."""
    verts = W.heightfield(grid=17, amplitude=4.0)
    rays = _edge_aimed_rays(verts, 6000, 17)
    ob = oracle_rt.build_triangles(verts)
    h, s, rt, tie = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, raw_t=True, ties=True)
    rival = np.zeros(len(rays), dtype=bool)
    for w in np.unique(s[s >= 0]):
        idx = np.nonzero(s == w)[0]
        others = np.delete(ob.tris, w)
        _, _, rtb = oracle_rt.trace_closest(ob.nodes, others, None, rays[idx], mode=oracle_rt.MODE_BRUTE, raw_t=True)
        rival[idx] = rtb.view(np.int32) == rt[idx].view(np.int32)
    assert tie.sum() >= 50, tie.sum()
    assert np.all(rival[tie])
    # Each unflagged rival lies on an edge the winner shares, i.e. on the face
    # of its own leaf box, and the traversal never tests it: once the ray's
    # distance is the hit t, some box on the rival's root-to-leaf path either
    # fails the slab test or is entered at exactly t.  The latter drops the
    # box too: a missed sibling's distance is the current t as well, so the
    # engine's near-first rule (hitL <= hitR, bvh.hsh:235) walks into the
    # missed sibling and pushes nothing.
    parent = {}
    for i, n in enumerate(ob.nodes):
        parent[int(n["left_ptr"])] = (i, "left")
        parent[int(n["right_ptr"])] = (i, "right")
    eon = ob.tris["v0"][:, 3] > 0  # end-of-leaf flag
    leaf_start = np.zeros(len(ob.tris), dtype=np.int64)
    for k in range(1, len(ob.tris)):
        leaf_start[k] = leaf_start[k - 1] if not eon[k - 1] else k

    def culled(i, slot):
        o, d, t = rays[i]["origin"], rays[i]["direction"], rt[i]
        key = ~int(leaf_start[slot])
        while key in parent:
            node, side = parent[key]
            n = ob.nodes[node]
            ok, dist = oracle_rt.slab_test(o, d, n[side + "_min"], n[side + "_max"], rays[i]["tmin"], t)
            if not ok or np.float32(dist) == t:
                return True
            key = node
        return False

    missed = np.nonzero(rival & ~tie)[0]
    for i in missed:
        o, d = rays[i]["origin"], rays[i]["direction"]
        for slot, tr in enumerate(ob.tris):
            if slot == s[i]:
                continue
            hit, sol = oracle_rt.triangle_test(o, d, tr["v0"][:3], tr["v1"][:3], tr["v2"][:3])
            if hit and np.float32(sol[0]) == rt[i]:
                assert culled(i, slot), (i, slot)
    print(f"{tie.sum()} flagged ties, {len(missed)} rivals dropped at a box boundary")
    # random rays through a triangle soup almost never tie
    verts2, rays2 = rt_cases.soup(4096, 20000, 18)
    ob2 = oracle_rt.build_triangles(verts2)
    *_, tie2 = oracle_rt.trace_closest(ob2.nodes, ob2.tris, ob2.ref_idx, rays2, ties=True)
    assert tie2.sum() == 0


def test_oracle_tie_flags_on_duplicated_triangles():
    """Every triangle present twice: each hit has a rival at the same t."""
    verts, rays = rt_cases.soup(8192, 20000, 19)
    dup = np.concatenate([verts, verts])
    ob = oracle_rt.build_triangles(dup)
    h, s, tie = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, ties=True)
    hit = s >= 0
    assert hit.sum() > 1000
    assert np.array_equal(tie, hit)


def test_chain_tree_oracle():
    nodes, tris, rays = rt_cases.chain_tree(80)
    h, _ = oracle_rt.trace_closest(nodes, tris, np.arange(len(tris), dtype=np.uint32), rays)
    assert np.all(h["prim_id"] == 0) and np.all(h["t"] == 105.0)


def test_visit_counts_reasonable_c0_sample():
    verts, rays = W.c0(n_rays=16384)
    ob = oracle_rt.build_triangles(verts)
    h, s, nv, tt = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays, stats=True)
    assert 20 < nv.mean() < 200 and 1 <= tt.mean() < 20
    assert 0.3 < (h["prim_id"] >= 0).mean() < 0.9
