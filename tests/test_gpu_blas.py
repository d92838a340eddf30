"""GPU parity of the single-BLAS traversal kernels against the oracle.

The oracle restates HitClosest / HitAny (data/shader/raytracer/bvh.hsh:191-273,
359-441) on the same packed arrays the kernels read, so the bar is bit
equality: hit slot, caller index, t, u and v must match exactly.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import workloads as W

pytestmark = pytest.mark.gpu


def _to_dev(torch, rays):
    return torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()


def _gpu_closest(torch, blas, rays, stats=False, kernel=0):
    d_rays = _to_dev(torch, rays)
    n = len(rays)
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(n, dtype=torch.int32, device="cuda")
    if stats:
        d_vis = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        blas.trace_closest_stats(d_rays, d_hits, d_vis, d_slots)
    else:
        blas.trace_closest(d_rays, d_hits, d_slots, kernel=kernel)
    torch.cuda.synchronize()
    hits = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
    slots = d_slots.cpu().numpy()
    if stats:
        return hits, slots, d_vis.cpu().numpy().astype(np.uint32)
    return hits, slots


def _gpu_any(torch, blas, rays):
    d_rays = _to_dev(torch, rays)
    d_hit = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    blas.trace_any(d_rays, d_hit)
    torch.cuda.synchronize()
    return d_hit.cpu().numpy().astype(np.uint32)


def _assert_same(h_gpu, s_gpu, h_ref, s_ref, what):
    bad = np.nonzero(np.any(rt_cases.hit_bits(h_gpu) != rt_cases.hit_bits(h_ref), axis=1) | (s_gpu != s_ref))[0]
    assert len(bad) == 0, (f"{what}: {len(bad)} rays differ, first {bad[:5]}: gpu {h_gpu[bad[:3]]} "
                           f"slots {s_gpu[bad[:3]]} vs oracle {h_ref[bad[:3]]} slots {s_ref[bad[:3]]}")


@pytest.mark.parametrize("n_tris,n_rays,salt", [(1, 4096, 1), (2, 4096, 2), (3, 4096, 3), (7, 8192, 4),
                                                (256, 65536, 5), (4096, 65536, 6), (65536, 131072, 7)])
def test_closest_bit_exact_random(torch_cuda, n_tris, n_rays, salt):
    print(f"case closest n_tris={n_tris} n_rays={n_rays} salt={salt}", flush=True)
    verts, rays = rt_cases.soup(n_tris, n_rays, salt)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    h, s = _gpu_closest(torch_cuda, blas, rays)
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    _assert_same(h, s, hr, sr, f"closest {n_tris}x{n_rays}")


@pytest.mark.parametrize("kernel", [2, 2 | 0x40, 2 | 0x50])
@pytest.mark.parametrize("n_rays", [1, 63, 1000, 65536, 1048576])
def test_persistent_kernel_bit_exact(torch_cuda, n_rays, kernel):
    """The persistent wave-refill kernels (plain, leaf-batched, leaf-batched
    with the register-topped stack) return the engine's hits, including
    batches smaller than one wave, and the self-resetting ray counter works
    across consecutive launches (each call is made twice)."""
    print(f"case persistent n_rays={n_rays} kernel={kernel}", flush=True)
    verts, rays = rt_cases.soup(65536, n_rays, 61)
    if n_rays >= 1000:
        rays[::97]["direction"] = np.nan  # NaN rays are misses
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    for _ in range(2):
        h, s = _gpu_closest(torch_cuda, blas, rays, kernel=kernel)
        _assert_same(h, s, hr, sr, f"persistent {n_rays} kernel {kernel}")


@pytest.mark.parametrize("n_tris,salt", [(4096, 11), (65536, 12)])
def test_closest_edge_rays(torch_cuda, n_tris, salt):
    print(f"case closest-edge n_tris={n_tris} salt={salt}", flush=True)
    verts, _ = rt_cases.soup(n_tris, 16, salt)
    rays = rt_cases.edge_rays(verts, salt=salt)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    h, s = _gpu_closest(torch_cuda, blas, rays)
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    _assert_same(h, s, hr, sr, "closest edge rays")
    assert (h["prim_id"] >= 0).sum() > len(rays) // 4


def test_closest_c0_full_scale(torch_cuda):
    """Config 0 at full size (65,536 triangles, 1,048,576 rays): bit-exact."""
    print("case closest c0 full", flush=True)
    verts, rays = W.c0()
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    h, s, vis = _gpu_closest(torch_cuda, blas, rays, stats=True)
    hr, sr, nv, tt = oracle_rt.trace_closest(nodes, tris, orig, rays, stats=True)
    _assert_same(h, s, hr, sr, "c0 full")
    # the kernel's own visit counters equal the oracle's
    assert np.array_equal(vis[:, 0], nv) and np.array_equal(vis[:, 1], tt)


def test_any_bit_exact(torch_cuda):
    print("case any-hit", flush=True)
    verts, rays = rt_cases.soup(16384, 65536, 21)
    rays = np.concatenate([rays, rt_cases.edge_rays(verts, salt=21)])
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, _ = blas.export()
    got = _gpu_any(torch_cuda, blas, rays)
    ref = oracle_rt.trace_any(nodes, tris, rays)
    assert np.array_equal(got, ref)
    # limited tmax: any-hit must agree with "closest t < tmax"
    rays2 = rays.copy()
    rays2["tmax"] = 2.0
    got2 = _gpu_any(torch_cuda, blas, rays2)
    assert np.array_equal(got2, oracle_rt.trace_any(nodes, tris, rays2))


def test_stack_spill_deep_chain(torch_cuda):
    """A tree 80 levels deep: the stack spills past the 32 LDS entries."""
    print("case deep chain", flush=True)
    nodes, tris, rays = rt_cases.chain_tree(80)
    blas = atlas_rt.Blas.create(nodes, tris)
    assert blas.info()[2] >= 80
    h, s = _gpu_closest(torch_cuda, blas, rays)
    hr, sr = oracle_rt.trace_closest(nodes, tris, np.arange(len(tris), dtype=np.uint32), rays)
    _assert_same(h, s, hr, sr, "deep chain")
    assert np.all(h["prim_id"] == 0)
    np.testing.assert_array_equal(h["t"], np.float32(105.0))
    assert np.array_equal(_gpu_any(torch_cuda, blas, rays), oracle_rt.trace_any(nodes, tris, rays))


def test_imported_oracle_blas(torch_cuda):
    """A BLAS built by the oracle and imported in the engine layout traces identically."""
    print("case imported", flush=True)
    verts, rays = rt_cases.soup(8192, 32768, 31)
    ob = oracle_rt.build_triangles(verts)
    blas = atlas_rt.Blas.create(ob.nodes, ob.tris, ob.ref_idx)
    h, s = _gpu_closest(torch_cuda, blas, rays)
    hr, sr = oracle_rt.trace_closest(ob.nodes, ob.tris, ob.ref_idx, rays)
    _assert_same(h, s, hr, sr, "imported")


def test_host_buffer_api(torch_cuda):
    """A device BLAS also answers the engine's CPU traversal on the host
    (BVH::GetIntersection semantics, the oracle's R-CPU), and it agrees with
    the GPU's HitClosest wherever the two semantics do (tmin = 0: same hit
    distance bit for bit)."""
    verts, rays = rt_cases.soup(2048, 8192, 41)
    rays["tmin"] = 0.0
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    h = blas.trace_closest_host(rays)
    hr, _, raw_t = oracle_rt.trace_closest(nodes, tris, orig, rays, mode=oracle_rt.MODE_RCPU, raw_t=True)
    assert np.array_equal(h["prim_id"], hr["prim_id"])
    hit = h["prim_id"] >= 0
    assert hit.mean() > 0.02 and np.array_equal(h["t"][hit].view(np.uint32), raw_t[hit].view(np.uint32))
    assert np.array_equal(blas.trace_any_host(rays), oracle_rt.trace_any(nodes, tris, rays, mode=oracle_rt.MODE_RCPU))
    _, _, gpu_t = oracle_rt.trace_closest(nodes, tris, orig, rays, raw_t=True)
    assert np.array_equal(raw_t.view(np.uint32), gpu_t.view(np.uint32))


def test_hits_match_brute_force_and_fp64(torch_cuda):
    """Independent of the tree: the GPU hit equals brute force over all
    triangles (ids differ only at equal float32 t), and the reported
    (t, u, v) are within 1e-5 relative of an FP64 recompute."""
    verts, rays = rt_cases.soup(4096, 65536, 51)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    h, s = _gpu_closest(torch_cuda, blas, rays)
    hg, sg, tg = oracle_rt.trace_closest(nodes, tris, orig, rays, raw_t=True)
    assert np.array_equal(s, sg)  # GPU == engine traversal (bit-exact slots)
    hb, _, tb = oracle_rt.trace_closest(nodes, tris, orig, rays, mode=oracle_rt.MODE_BRUTE, raw_t=True)
    assert np.array_equal(tg.view(np.int32), tb.view(np.int32))
    same = h["prim_id"] == hb["prim_id"]
    assert np.array_equal(h["t"][same].view(np.int32), hb["t"][same].view(np.int32))
    import fp64
    err = fp64.tuv_errors(verts, rays, h)
    print("fp64 errors", err)
    assert err["n_hits"] > 1000
    assert err["n_over"] == 0 and err["max_rel_t"] < 1e-6, err


def _tie_ok(rays, tris, s_gpu, s_ref, bad):
    """Rays whose GPU and oracle slots differ must be exact equal-t ties:
    both triangles hit at the same float32 t (the engine's own test)."""
    for i in bad:
        o, d = rays[i]["origin"], rays[i]["direction"]
        ts = []
        for s in (s_gpu[i], s_ref[i]):
            if s < 0:
                return False
            tr = tris[s]
            hit, sol = oracle_rt.triangle_test(o, d, tr["v0"][:3], tr["v1"][:3], tr["v2"][:3])
            ts.append(sol[0] if hit else None)
        if ts[0] is None or ts[0] != ts[1]:
            return False
    return True


VARIANTS = [("bvh2 grid, register stack", 1 | 0x10), ("bvh2 persistent, register stack", 2 | 0x10),
            ("bvh4", 3), ("bvh4, register stack", 3 | 0x10), ("bvh8", 4), ("bvh8, register stack", 4 | 0x10),
            ("bvh2 grid, float slab", 1 | 0x20), ("bvh2 persistent, float slab", 2 | 0x20),
            ("bvh2 grid, register stack, float slab", 1 | 0x30),
            ("bvh2 persistent, register stack, float slab", 2 | 0x30),
            ("bvh4, leaf batch", 3 | 0x40), ("bvh4, float slab", 3 | 0x20), ("bvh4, float slab, leaf batch", 3 | 0x60),
            ("bvh4, register stack, float slab, leaf batch", 3 | 0x70)]


@pytest.mark.parametrize("name,kernel", VARIANTS)
@pytest.mark.parametrize("n_tris,n_rays,salt", [(1, 4096, 71), (7, 8192, 72), (4096, 65536, 73), (65536, 262144, 74)])
def test_variant_kernels_match_engine_distance(torch_cuda, name, kernel, n_tris, n_rays, salt):
    """Wide-node, register-stack and float-slab variants: the same nearest
    distance and hit/miss as the engine's traversal for every ray; a
    different slot only at an exact equal-t tie (binary exact-slab variants:
    identical)."""
    print(f"case variant {name} n_tris={n_tris} n_rays={n_rays}", flush=True)
    verts, rays = rt_cases.soup(n_tris, n_rays, salt)
    rays[::101]["direction"] = np.nan
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    for _ in range(2):  # the persistent counter resets itself between launches
        h, s = _gpu_closest(torch_cuda, blas, rays, kernel=kernel)
        if (kernel & 0xF) <= 2 and not kernel & 0x20:
            _assert_same(h, s, hr, sr, name)
            continue
        # Another slot is allowed only for the same triangle (the root SBVH
        # split references a triangle crossing the plane from two slots) or
        # a different triangle at exactly the same float32 t.
        same_rec = np.all(rt_cases.hit_bits(h) == rt_cases.hit_bits(hr), axis=1)
        assert np.all(same_rec[s == sr]), name
        other = np.nonzero((s != sr) & ~same_rec)[0]
        dup = np.count_nonzero((s != sr) & same_rec)
        assert len(other) <= max(2, len(rays) // 10000) and _tie_ok(rays, tris, s, sr, other), (name, other[:5])
        assert np.all(orig[s[(s != sr) & same_rec]] == orig[sr[(s != sr) & same_rec]])
        print(f"{name}: {dup} hits on a duplicate slot of the same triangle, {len(other)} equal-t ties")


def _fp32_t(rays, tris, i, slot):
    tr = tris[slot]
    hit, sol = oracle_rt.triangle_test(rays[i]["origin"], rays[i]["direction"], tr["v0"][:3], tr["v1"][:3],
                                       tr["v2"][:3])
    return sol[0] if hit else np.inf


@pytest.mark.parametrize("kernel", [3, 4, 3 | 0x10, 4 | 0x10, 2 | 0x10, 2 | 0x20, 2 | 0x30, 1 | 0x30, 2 | 0x70,
                                    3 | 0x60, 3 | 0x70])
def test_variant_kernels_edge_rays_and_deep_chain(torch_cuda, kernel):
    """Vertex- and edge-aimed rays (non-watertight culling cases) and an
    80-deep tree whose stack spills past the register top and LDS rows.
    The widened float slab test enters every box the exact one enters."""
    print(f"case variant edge/deep kernel={kernel}", flush=True)
    verts, _ = rt_cases.soup(4096, 16, 75)
    rays = rt_cases.edge_rays(verts, 65536, 76)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    h, s = _gpu_closest(torch_cuda, blas, rays, kernel=kernel)
    # A wide traversal tests a subset of the engine's boxes on the way to each
    # triangle, and the widened float slab test enters a superset of the boxes
    # the exact one enters, so either may find a boundary hit the engine's
    # slab test culled but never loses one: every engine hit is matched by a
    # hit at a float32 t no larger (equal unless the engine culled the nearer
    # one).
    assert np.all(s[sr >= 0] >= 0), "variant lost a hit"
    diff = np.nonzero((s != sr) & (sr >= 0))[0]
    for i in diff:
        assert _fp32_t(rays, tris, i, s[i]) <= _fp32_t(rays, tris, i, sr[i]), i
    same = (s == sr) & (sr >= 0)
    assert np.array_equal(rt_cases.hit_bits(h[same]), rt_cases.hit_bits(hr[same]))
    print(f"edge rays: {len(diff)} of {(sr >= 0).sum()} hits on another slot, {(s[sr < 0] >= 0).sum()} extra hits")

    nodes_c, tris_c, rays_c = rt_cases.chain_tree(80)
    blas_c = atlas_rt.Blas.create(nodes_c, tris_c)
    h, s = _gpu_closest(torch_cuda, blas_c, rays_c, kernel=kernel)
    assert np.all(h["prim_id"] == 0) and np.all(s == 0)
    np.testing.assert_array_equal(h["t"], np.float32(105.0))
