"""GPU linear-BVH builder (atlas_rt_blas_build_device).

CPU: the oracle's restatement pinned by a textbook binary radix tree
(recursive split at the highest differing key bit), closed-form Morton order
and tree shape for eight corner triangles, tree invariants (one triangle per
leaf, pre-order, larger-area child left, boxes equal to the union of their
triangles) and brute-force nearest hits through the oracle's traversal.
GPU: the device build equals the oracle's arrays byte for byte, batched, and
the trace kernels on it match the oracle's traversal bit for bit.

Reference: BVHBuilder::Flatten (src/engine/volume/BVH.cpp:408-441), the node
and triangle packing of MeshData::BuildBVH (MeshData.cpp:172-269).
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import workloads as W


def _radix_tree_reference(keys):
    """Binary radix tree by definition: the range [lo, hi] splits after the
    last key whose highest differing bit (between keys[lo] and keys[hi]) is
    0; the node of a range is numbered by its split's side, as in Karras'
    layout (root = 0, children gamma and gamma + 1)."""
    n = len(keys)
    out = np.zeros((max(n - 1, 0), 2), np.int64)

    def rec(node, lo, hi):
        b = int(keys[lo] ^ keys[hi]).bit_length() - 1
        gamma = max(i for i in range(lo, hi) if not (int(keys[i]) >> b) & 1)
        c0 = ~gamma if lo == gamma else gamma
        c1 = ~(gamma + 1) if hi == gamma + 1 else gamma + 1
        out[node] = (c0, c1)
        if c0 >= 0:
            rec(gamma, lo, gamma)
        if c1 >= 0:
            rec(gamma + 1, gamma + 1, hi)

    if n > 1:
        rec(0, 0, n - 1)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_oracle_radix_tree_is_the_binary_radix_tree(seed):
    g = np.random.default_rng(seed)
    for n in (2, 3, 5, 64, 300):
        if seed % 2:
            keys = np.unique(g.integers(0, 1 << 62, n * 2, dtype=np.int64).astype(np.uint64))[:n]
        else:
            # long shared prefixes and runs of close keys
            base = np.uint64(0xABCD) << np.uint64(40)
            keys = np.unique(base + g.integers(0, 4 * n, n * 2).astype(np.uint64))[:n]
        keys = np.sort(keys)
        assert np.array_equal(oracle_rt.radix_tree(keys), _radix_tree_reference(keys)), (seed, n)


def _corner_soup():
    """Eight equal tiny triangles centred on the corners of [0, 1]^3, listed
    in reverse Morton order (x is the top bit, then y, then z)."""
    tri = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.0]], np.float32)
    tri -= tri.mean(axis=0)
    corners = [(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)][::-1]
    return np.stack([(tri + np.array(c, np.float32)).reshape(9) for c in corners]).astype(np.float32)


def test_oracle_lbvh_corner_closed_form():
    """Morton order is z fastest, then y, then x; every box of a level has the
    same area, so the radix tree's order stands: a complete tree of depth 3
    with the triangles in Morton order in slots 0..7."""
    verts = _corner_soup()
    nodes, tris, orig, depth = oracle_rt.build_lbvh(verts)
    assert len(nodes) == 7 and depth == 3
    assert list(orig) == [7, 6, 5, 4, 3, 2, 1, 0]
    # pre-order: 0 -> (1 -> (2, 3), 4 -> (5, 6)), inner-node children in order
    assert [int(p) for p in nodes["left_ptr"]] == [1, 2, ~0, ~2, 5, ~4, ~6]
    assert [int(p) for p in nodes["right_ptr"]] == [4, 3, ~1, ~3, 6, ~5, ~7]
    # the root's children split x at 0.5
    assert nodes[0]["left_max"][0] < 0.5 < nodes[0]["right_min"][0]


def _check_tree(verts, nodes, tris, orig, depth):
    n = len(verts)
    assert len(nodes) == n - 1 and len(tris) == n
    assert np.array_equal(np.sort(orig), np.arange(n))
    v = verts.reshape(-1, 3, 3)[orig]
    assert np.array_equal(np.stack([tris["v0"][:, :3], tris["v1"][:, :3], tris["v2"][:, :3]], axis=1), v)
    assert np.all(tris["v0"][:, 3] == 1.0) and np.all(tris["v2"][:, 3] == 1.0)
    lo, hi = v.min(axis=1), v.max(axis=1)
    if n == 1:
        assert depth == 1
        return
    next_slot = [0]
    max_depth = [0]

    def area(bmin, bmax):
        d = (np.asarray(bmax, np.float32) - np.asarray(bmin, np.float32)).astype(np.float32)
        return np.float32(2.0) * ((d[0] * d[1] + d[1] * d[2]) + d[2] * d[0])

    def walk(i, d):
        """Returns the slot range of node i's subtree; checks its boxes."""
        nd = nodes[i]
        assert area(nd["left_min"], nd["left_max"]) >= area(nd["right_min"], nd["right_max"])
        ranges = []
        for side in ("left", "right"):
            p = int(nd[f"{side}_ptr"])
            if p < 0:
                assert ~p == next_slot[0]
                next_slot[0] += 1
                max_depth[0] = max(max_depth[0], d + 1)
                r = (~p, ~p + 1)
            else:
                assert p > i
                r = walk(p, d + 1)
            assert np.array_equal(nd[f"{side}_min"], lo[r[0]:r[1]].min(axis=0))
            assert np.array_equal(nd[f"{side}_max"], hi[r[0]:r[1]].max(axis=0))
            ranges.append(r)
        assert ranges[0][1] == ranges[1][0]
        return ranges[0][0], ranges[1][1]

    assert walk(0, 0) == (0, n)
    assert depth == max_depth[0]


@pytest.mark.parametrize("case", ["one", "two", "three", "soup17", "soup2000", "identical", "plane", "corners"])
def test_oracle_lbvh_invariants_and_brute_force(case):
    g = np.random.default_rng(5)
    if case.startswith("soup") or case in ("one", "two", "three"):
        n = {"one": 1, "two": 2, "three": 3}.get(case) or int(case[4:])
        verts, _ = rt_cases.soup(n, 1, 80 + n)
    elif case == "identical":
        verts = np.tile(g.uniform(-1, 1, 9).astype(np.float32), (300, 1))
    elif case == "plane":
        verts, _ = rt_cases.soup(500, 1, 81)
        verts = verts.reshape(-1, 3, 3)
        verts[:, :, 1] = 0.25
        verts = verts.reshape(-1, 9)
    else:
        verts = _corner_soup()
    nodes, tris, orig, depth = oracle_rt.build_lbvh(verts)
    _check_tree(verts, nodes, tris, orig, depth)
    assert depth <= 64 or len(verts) > 1 << 20
    # the oracle's engine-order traversal on this tree finds the brute-force nearest hit
    rays = rt_cases.soup(4, 4096, 90)[1]
    hits, slots = oracle_rt.trace_closest(nodes, tris, orig, rays)
    brute, _ = oracle_rt.trace_closest(nodes, tris, orig, rays, mode=oracle_rt.MODE_BRUTE)
    assert np.array_equal(hits["prim_id"] >= 0, brute["prim_id"] >= 0)
    both = hits["prim_id"] >= 0
    assert np.allclose(hits["t"][both], brute["t"][both], rtol=1e-6, atol=0)


def _device_cases():
    g = np.random.default_rng(11)
    cases = [rt_cases.soup(n, 1, 100 + n)[0] for n in (1, 2, 3, 100, 4096, 65536)]
    cases.append(np.tile(g.uniform(-1, 1, 9).astype(np.float32), (500, 1)))
    cases.append(_corner_soup())
    return cases


@pytest.mark.gpu
def test_gpu_lbvh_equals_oracle(torch_cuda):
    """One batched device build of eight meshes: every node, triangle record,
    slot order and depth equals the oracle's."""
    print("case lbvh device build", flush=True)
    cases = _device_cases()
    blases, ms = atlas_rt.build_many_device(cases)
    print(f"lbvh: {sum(len(c) for c in cases)} triangles in {len(cases)} meshes, {ms:.3f} ms of GPU build")
    for verts, b in zip(cases, blases):
        nodes, tris, orig, depth = oracle_rt.build_lbvh(verts)
        gn, gt, go = b.export()
        assert gn.tobytes() == nodes.tobytes(), len(verts)
        assert gt.tobytes() == tris.tobytes(), len(verts)
        assert np.array_equal(go, orig), len(verts)
        assert b.info()[2] == depth, (len(verts), b.info(), depth)


def _fp32_t(rays, tris, i, slot):
    tr = tris[slot]
    hit, sol = oracle_rt.triangle_test(rays[i]["origin"], rays[i]["direction"], tr["v0"][:3], tr["v1"][:3],
                                       tr["v2"][:3])
    return sol[0] if hit else np.inf


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 2, 2 | 0x10, 3, 4 | 0x10])
def test_gpu_lbvh_traces_match_oracle(torch_cuda, kernel):
    """The engine-order kernels on a device-built tree give the oracle's
    traversal result on the same tree, bit for bit; the wide kernels (over the
    wide nodes collapsed from the device-built tree) find every engine hit at
    a float32 t no larger, with identical records where the slot agrees."""
    print(f"case lbvh traces kernel={kernel}", flush=True)
    torch = torch_cuda
    verts, rays = rt_cases.soup(65536, 262144, 120)
    (b,), _ = atlas_rt.build_many_device([verts])
    nodes, tris, orig = b.export()
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    b.trace_closest(d_rays, d_hits, d_slots, kernel=kernel)
    torch.cuda.synchronize()
    assert (sr >= 0).mean() > 0.3
    h = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
    s = d_slots.cpu().numpy()
    if kernel & 0xF in (1, 2):
        assert h.tobytes() == hr.tobytes()
        assert np.array_equal(s, sr)
        return
    assert np.all(s[sr >= 0] >= 0), "wide traversal lost a hit"
    diff = np.nonzero((s != sr) & (sr >= 0))[0]
    for i in diff:
        assert _fp32_t(rays, tris, i, s[i]) <= _fp32_t(rays, tris, i, sr[i]), i
    same = (s == sr) & (sr >= 0)
    assert np.array_equal(rt_cases.hit_bits(h[same]), rt_cases.hit_bits(hr[same]))
    assert len(diff) < 0.001 * len(rays)


@pytest.mark.gpu
def test_gpu_lbvh_many_meshes(torch_cuda):
    """1,100 meshes (two batches of launches) of 1-3,000 triangles: spot
    checks against the oracle, and nearest hits equal to the SAH-built trees'."""
    print("case lbvh many meshes", flush=True)
    torch = torch_cuda
    g = np.random.default_rng(12)
    sizes = g.integers(1, 3000, 1100)
    meshes = [rt_cases.soup(int(n), 1, 2000 + i)[0] for i, n in enumerate(sizes)]
    blases, ms = atlas_rt.build_many_device(meshes)
    print(f"lbvh: {int(sizes.sum())} triangles in {len(meshes)} meshes, {ms:.3f} ms of GPU build")
    for i in (0, 1, 517, 1023, 1024, 1099):
        nodes, tris, orig, depth = oracle_rt.build_lbvh(meshes[i])
        gn, gt, go = blases[i].export()
        assert gn.tobytes() == nodes.tobytes() and gt.tobytes() == tris.tobytes() and np.array_equal(go, orig), i
    rays = rt_cases.soup(4, 8192, 130)[1]
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    for i in (3, 600):
        sah = atlas_rt.Blas.build(meshes[i])
        out = []
        for b in (blases[i], sah):
            b.trace_closest(d_rays, d_hits, kernel=1)
            torch.cuda.synchronize()
            out.append(d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1).copy())
        assert np.array_equal(out[0]["prim_id"] >= 0, out[1]["prim_id"] >= 0)
        assert np.array_equal(out[0]["t"], out[1]["t"])
