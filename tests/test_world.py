"""Two-level (TLAS + instances) traversal: oracle pins on the CPU and GPU
parity of the world kernels.

Reference: RayTracingWorld::Update / UpdateForSoftwareRayTracing
(src/engine/raytracing/RayTracingWorld.cpp:32-162, 267-307) and HitClosest /
HitAny with CheckInstance (data/shader/raytracer/bvh.hsh:172-273, 359-441).
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W


def small_scene(n_meshes=3, n_inst=24, salt=0, tris_per_mesh=300):
    g = W.rng(2, 100 + salt)
    meshes = [W.triangle_soup(tris_per_mesh, g, extent=2.0, edge=(0.2, 0.8)) for _ in range(n_meshes)]
    xf = W.random_affine(g, n_inst, scale=(0.5, 2.0), extent=20.0, height=10.0)
    which = g.integers(0, n_meshes, n_inst)
    descs = atlas_rt.instance_descs(xf, which)
    # every third instance casts no shadow
    descs["mask"][::3] = _abi.MASK_ALL
    return meshes, descs


def world_aabbs(meshes, descs):
    """World AABBs: the 8 transformed corners of each mesh's bounds (float64)."""
    out = np.zeros((len(descs), 6), np.float32)
    for i, d in enumerate(descs):
        tri = meshes[d["blas"]].reshape(-1, 3)
        lo, hi = tri.min(axis=0), tri.max(axis=0)
        c = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float64)
        m = d["object_to_world"].astype(np.float64)
        p = c @ m[:, :3].T + m[:, 3]
        out[i, :3], out[i, 3:] = p.min(axis=0), p.max(axis=0)
    return out


def oracle_world(meshes, descs, boxes):
    """Oracle-only world: restated builders plus the engine's instance packing."""
    blases = []
    for v in meshes:
        ob = oracle_rt.build_triangles(v)
        blases.append((ob.nodes, ob.tris, ob.ref_idx))
    tl = oracle_rt.build_aabbs(boxes)
    inst = np.zeros(len(tl.ref_idx), dtype=oracle_rt.INSTANCE_DTYPE)
    for i, r in enumerate(tl.ref_idx):
        inst[i]["inverse_matrix"] = descs[r]["world_to_object"]
        inst[i]["mesh_offset"] = descs[r]["blas"]
        inst[i]["material_offset"] = descs[r]["material_offset"]
        inst[i]["next_instance"] = -1 if tl.end_of_node[i] else i + 1
        inst[i]["mask"] = descs[r]["mask"]
    return tl.nodes, inst, tl.ref_idx, blases


def scene_rays(n, salt, extent=25.0):
    g = W.rng(2, 900 + salt)
    return W.random_rays(n, g, extent=extent)


@pytest.mark.parametrize("salt", [0, 1])
def test_oracle_world_equals_brute_force(salt):
    meshes, descs = small_scene(salt=salt)
    boxes = world_aabbs(meshes, descs)
    tl, inst, order, blases = oracle_world(meshes, descs, boxes)
    rays = scene_rays(20000, salt)
    for mask in (_abi.MASK_ALL, _abi.MASK_SHADOW):
        h, s, i, rt = oracle_rt.trace_world(tl, inst, blases, rays, mask=mask, raw_t=True)
        hb, sb, ib, rtb = oracle_rt.trace_world(tl, inst, blases, rays, mask=mask, mode=oracle_rt.WORLD_BRUTE,
                                                raw_t=True)
        assert (h["prim_id"] >= 0).mean() > 0.05
        assert np.array_equal(rt.view(np.int32), rtb.view(np.int32))
        same = (s == sb) & (i == ib)
        assert (~same).sum() <= 2  # equal-t ties only (t equal above)
        assert np.array_equal(h["t"][same].view(np.int32), hb["t"][same].view(np.int32))
        # HitAnyTransparency's traversal (restore every TLAS step) == existence
        # of a closest hit
        a, _, _ = oracle_rt.trace_world(tl, inst, blases, rays, mask=mask, mode=oracle_rt.WORLD_ANY_RESTORE)
        assert np.array_equal(a["prim_id"], (hb["prim_id"] >= 0).astype(np.int32))
        # HitAny keeps the object-space ray of a masked-out instance
        # (bvh.hsh:384-387); it only agrees when the mask admits every instance
        e, _, _ = oracle_rt.trace_world(tl, inst, blases, rays, mask=mask, mode=oracle_rt.WORLD_ANY)
        if mask == _abi.MASK_ALL:
            assert np.array_equal(e["prim_id"], a["prim_id"])
        else:
            assert (e["prim_id"] != a["prim_id"]).sum() > 0


def test_oracle_world_scale_invariance():
    """t stays in world units under scaled instances: one mesh, identity vs
    uniform scale 2 around the origin gives exactly doubled hit distances for
    rays from the origin (powers of two keep the arithmetic exact)."""
    v = W.triangle_soup(200, W.rng(2, 5), extent=3.0, edge=(0.3, 1.0))
    ob = oracle_rt.build_triangles(v)
    rays = W.random_rays(4000, W.rng(2, 6), extent=0.0)
    ident = np.zeros(1, dtype=oracle_rt.INSTANCE_DTYPE)
    ident[0]["inverse_matrix"] = np.eye(3, 4)
    ident[0]["mask"] = 0xFF
    ident[0]["next_instance"] = -1
    scaled = ident.copy()
    scaled[0]["inverse_matrix"] = np.eye(3, 4) * 0.5
    t1 = oracle_rt.build_aabbs(np.array([[-5, -5, -5, 5, 5, 5]], np.float32))
    t2 = oracle_rt.build_aabbs(np.array([[-10, -10, -10, 10, 10, 10]], np.float32))
    blases = [(ob.nodes, ob.tris, ob.ref_idx)]
    inst1 = np.concatenate([ident, ident])  # single-instance TLAS lists it twice
    inst2 = np.concatenate([scaled, scaled])
    _, s1, _, r1 = oracle_rt.trace_world(t1.nodes, inst1, blases, rays, mask=0xFF, raw_t=True)
    _, s2, _, r2 = oracle_rt.trace_world(t2.nodes, inst2, blases, rays, mask=0xFF, raw_t=True)
    hit = s1 >= 0
    assert hit.mean() > 0.2 and np.array_equal(s1, s2)
    assert np.array_equal(r2[hit], 2 * r1[hit]) or np.allclose(r2[hit], 2 * r1[hit], rtol=1e-6)


def test_instance_mask_zero_hides_everything():
    meshes, descs = small_scene(salt=3)
    descs["mask"] = 0
    tl, inst, _, blases = oracle_world(meshes, descs, world_aabbs(meshes, descs))
    h, s, i = oracle_rt.trace_world(tl, inst, blases, scene_rays(2000, 3), mask=_abi.MASK_ALL)
    assert np.all(s == -1) and np.all(h["prim_id"] == -1)


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def _product_world(meshes, descs, boxes=None):
    blases = [atlas_rt.Blas.build(v) for v in meshes]
    w = atlas_rt.World()
    w.update(descs, blases, world_aabbs=boxes)
    return w, blases


def _oracle_from_product(w, blases):
    tl, inst, order = w.export()
    bl = [b.export() for b in blases]
    return tl, inst, order, bl


@pytest.mark.gpu
@pytest.mark.parametrize("salt,n_inst", [(0, 1), (1, 2), (2, 24), (3, 256)])
def test_gpu_world_closest_any_bit_exact(torch_cuda, salt, n_inst):
    print(f"case world salt={salt} n_inst={n_inst}", flush=True)
    torch = torch_cuda
    meshes, descs = small_scene(salt=salt, n_inst=n_inst)
    boxes = world_aabbs(meshes, descs)
    w, blases = _product_world(meshes, descs, boxes)
    tl, inst, order, bl = _oracle_from_product(w, blases)
    # the product TLAS equals the oracle's restated AABB builder
    ot = oracle_rt.build_aabbs(boxes)
    assert tl.tobytes() == ot.nodes.tobytes() and np.array_equal(order, ot.ref_idx)
    rays = np.concatenate([scene_rays(50000, salt), rt_cases.make_rays(np.zeros((2, 3)), [[np.nan, 1, 0], [1, 0, 0]])])
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    n = len(rays)
    for mask in (_abi.MASK_ALL, _abi.MASK_SHADOW):
        d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        d_inst = torch.empty(n, dtype=torch.int32, device="cuda")
        w.trace_closest_rays(d_rays, d_hits, d_inst, mask=mask)
        d_any = torch.empty(n, dtype=torch.int32, device="cuda")
        w.trace_any_rays(d_rays, d_any, mask=mask)
        d_any2 = torch.empty(n, dtype=torch.int32, device="cuda")
        w.trace_any_rays(d_rays, d_any2, mask=mask, flags=_abi.ANY_RESTORE_ALWAYS)
        torch.cuda.synchronize()
        hits = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
        gi = d_inst.cpu().numpy()
        hr, sr, ir = oracle_rt.trace_world(tl, inst, bl, rays, mask=mask)
        assert np.array_equal(rt_cases.hit_bits(hits), rt_cases.hit_bits(hr))
        expect_inst = np.where(ir >= 0, order[np.maximum(ir, 0)].astype(np.int32), -1)
        assert np.array_equal(gi, expect_inst)
        ar, _, _ = oracle_rt.trace_world(tl, inst, bl, rays, mask=mask, mode=oracle_rt.WORLD_ANY)
        assert np.array_equal(d_any.cpu().numpy(), ar["prim_id"])
        ar2, _, _ = oracle_rt.trace_world(tl, inst, bl, rays, mask=mask, mode=oracle_rt.WORLD_ANY_RESTORE)
        assert np.array_equal(d_any2.cpu().numpy(), ar2["prim_id"])
        if mask == _abi.MASK_ALL and n_inst >= 24:
            assert (hits["prim_id"] >= 0).sum() > n // 20


@pytest.mark.gpu
def test_gpu_world_packed_ray_interface(torch_cuda):
    """The engine-shaped PackedRay trace: count read on the device, ray.ID < 0
    skipped, hit = (engine float32 t, BVH slot, instance slot)."""
    torch = torch_cuda
    meshes, descs = small_scene(salt=7, n_inst=64)
    w, blases = _product_world(meshes, descs)
    tl, inst, order, bl = _oracle_from_product(w, blases)
    rays = scene_rays(30000, 7)
    n = len(rays)
    packed = np.zeros(n, dtype=atlas_rt.PACKED_RAY_DTYPE)
    packed["origin"][:, :3] = rays["origin"]
    packed["direction"][:, :3] = rays["direction"]
    ids = np.arange(n, dtype=np.int32)
    ids[::11] = -1
    packed["origin"][:, 3] = ids.view(np.float32)
    packed["hit"][:, 2] = np.full(n, 77, np.int32).view(np.float32)
    live = 25000
    d_in = torch.from_numpy(packed.view(np.float32).reshape(-1, 12).copy()).cuda()
    d_out = torch.zeros_like(d_in)
    d_count = torch.tensor([live], dtype=torch.int32, device="cuda")
    w.trace_closest_packed(d_in, d_out, d_count, n, tmin=0.0, tmax=1e12)
    d_occ = torch.zeros(n, dtype=torch.int32, device="cuda")
    w.trace_any_packed(d_in, d_occ, d_count, n, tmin=0.0, tmax=30.0)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy().view(atlas_rt.PACKED_RAY_DTYPE).reshape(-1)
    rr = rays.copy()
    rr["tmin"], rr["tmax"] = 0.0, 1e12
    _, sr, ir, rt = oracle_rt.trace_world(tl, inst, bl, rr, mask=_abi.MASK_ALL, raw_t=True)
    hit = out["hit"]
    ok = ids[:live] >= 0
    assert np.array_equal(hit[:live, 0][ok].view(np.int32), rt[:live][ok].view(np.int32))
    assert np.array_equal(hit[:live, 1][ok].view(np.int32), sr[:live][ok])
    expect_inst = np.where(sr[:live] >= 0, ir[:live], 77)
    assert np.array_equal(hit[:live, 2][ok].view(np.int32), expect_inst[ok])
    # skipped rays: t = 0, hitID = -1; rays past the count are untouched (zero)
    assert np.all(hit[:live, 0][~ok] == 0) and np.all(hit[:live, 1][~ok].view(np.int32) == -1)
    assert np.all(out[live:]["hit"] == 0)
    rr["tmax"] = 30.0
    ar, _, _ = oracle_rt.trace_world(tl, inst, bl, rr, mask=_abi.MASK_SHADOW, mode=oracle_rt.WORLD_ANY)
    occ = d_occ.cpu().numpy()
    assert np.array_equal(occ[:live], ar["prim_id"][:live]) and np.all(occ[live:] == 0)
