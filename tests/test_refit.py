"""Per-frame TLAS refit (atlas_rt_world_refit).

CPU: the oracle's refit restatement pinned three ways: refitting with
unchanged matrices reproduces the restated builder's TLAS byte for byte (its
node boxes come from the builder's own bounds, not from the refit's walk);
instance boxes agree with a float64 transform of the object box (exactly for
the identity); after moving the instances, every node's child box is the
union of the instance boxes beneath it, and the oracle's traversal of the
refitted tree finds the brute-force nearest hits.
GPU: the device refit equals the oracle's (nodes, instance records, transform
table) byte for byte, twice in a row; traces of the refitted world match the
oracle's traversal of the same arrays bit for bit; the path tracer follows the
moved instances.

Reference: RayTracingWorld::Update (src/engine/raytracing/RayTracingWorld.cpp:32-162,
267-307), AABB::Transform (src/engine/volume/AABB.cpp:35-60).
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W
from test_world import scene_rays, small_scene

ENV = (0.25, 0.3, 0.4)


def _boxes(descs, bl):
    """Instance world boxes in caller order, as the refit computes them."""
    n = len(descs)
    inst = np.zeros(n, dtype=oracle_rt.INSTANCE_DTYPE)
    inst["mesh_offset"] = descs["blas"]
    inst["next_instance"] = -1
    _, _, _, boxes = oracle_rt.world_refit(np.zeros(0, oracle_rt.NODE_DTYPE), inst, np.zeros((n, 12), np.float32),
                                           np.arange(n, dtype=np.uint32), descs, bl)
    return boxes


def _oracle_world(meshes, descs):
    """The world atlas_rt_world_update builds with world_aabbs NULL."""
    obs = [oracle_rt.build_triangles(v) for v in meshes]
    bl = [(ob.nodes, ob.tris) for ob in obs]
    tl = oracle_rt.build_aabbs(_boxes(descs, bl))
    inst = np.zeros(len(tl.ref_idx), dtype=oracle_rt.INSTANCE_DTYPE)
    xform = np.zeros((len(tl.ref_idx), 12), np.float32)
    for i, r in enumerate(tl.ref_idx):
        inst[i]["inverse_matrix"] = descs[r]["world_to_object"]
        inst[i]["mesh_offset"] = descs[r]["blas"]
        inst[i]["material_offset"] = descs[r]["material_offset"]
        inst[i]["next_instance"] = -1 if tl.end_of_node[i] else i + 1
        inst[i]["mask"] = descs[r]["mask"]
        xform[i] = descs[r]["object_to_world"].reshape(-1)
    return tl.nodes, inst, xform, tl.ref_idx, bl, [(ob.nodes, ob.tris, ob.ref_idx) for ob in obs]


def _moved(descs, salt, n_meshes):
    g = W.rng(2, 700 + salt)
    xf = W.random_affine(g, len(descs), scale=(0.5, 2.0), extent=20.0, height=10.0)
    moved = atlas_rt.instance_descs(xf, descs["blas"])
    moved["mask"] = descs["mask"]
    moved["material_offset"] = descs["material_offset"]
    return moved


@pytest.mark.parametrize("salt,n_inst", [(0, 1), (1, 24), (2, 200)])
def test_oracle_refit_without_motion_is_the_build(salt, n_inst):
    meshes, descs = small_scene(salt=salt, n_inst=n_inst)
    tl, inst, xform, order, bl, _ = _oracle_world(meshes, descs)
    t2, i2, x2, _ = oracle_rt.world_refit(tl, inst, xform, order, descs, bl)
    assert t2.tobytes() == tl.tobytes()
    assert i2.tobytes() == inst.tobytes()
    assert x2.tobytes() == xform.tobytes()


def test_oracle_instance_boxes_are_transformed_object_boxes():
    meshes, descs = small_scene(salt=4, n_inst=64)
    obs = [oracle_rt.build_triangles(v) for v in meshes]
    bl = [(ob.nodes, ob.tris) for ob in obs]
    boxes = _boxes(descs, bl)
    for i, d in enumerate(descs):
        tri = meshes[d["blas"]].reshape(-1, 3)
        lo, hi = tri.min(axis=0), tri.max(axis=0)
        c = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float64)
        m = d["object_to_world"].astype(np.float64)
        p = c @ m[:, :3].T + m[:, 3]
        scale = np.abs(p).max() + 1.0
        assert np.all(np.abs(boxes[i, :3] - p.min(axis=0)) <= 4e-7 * scale), i
        assert np.all(np.abs(boxes[i, 3:] - p.max(axis=0)) <= 4e-7 * scale), i
    ident = descs[:1].copy()
    ident["object_to_world"] = np.eye(3, 4, dtype=np.float32)
    ident["world_to_object"] = np.eye(3, 4, dtype=np.float32)
    tri = meshes[ident[0]["blas"]].reshape(-1, 3)
    assert np.array_equal(_boxes(ident, bl)[0], np.concatenate([tri.min(axis=0), tri.max(axis=0)]))


def _subtree_boxes_are_unions(tl, inst, boxes):
    def walk(ptr):
        if ptr >= 0:
            n = tl[ptr]
            lb, rb = walk(int(n["left_ptr"])), walk(int(n["right_ptr"]))
            assert np.array_equal(n["left_min"], lb[:3]) and np.array_equal(n["left_max"], lb[3:])
            assert np.array_equal(n["right_min"], rb[:3]) and np.array_equal(n["right_max"], rb[3:])
            return np.concatenate([np.minimum(lb[:3], rb[:3]), np.maximum(lb[3:], rb[3:])])
        s = ~ptr
        out = boxes[s].copy()
        while inst[s]["next_instance"] >= 0:
            s += 1
            out = np.concatenate([np.minimum(out[:3], boxes[s, :3]), np.maximum(out[3:], boxes[s, 3:])])
        return out
    if len(tl):
        walk(0)


@pytest.mark.parametrize("salt", [0, 1])
def test_oracle_refit_moves_and_still_finds_the_nearest_hits(salt):
    meshes, descs = small_scene(salt=salt, n_inst=48)
    tl, inst, xform, order, bl, blases = _oracle_world(meshes, descs)
    moved = _moved(descs, salt, len(meshes))
    t2, i2, x2, boxes = oracle_rt.world_refit(tl, inst, xform, order, moved, bl)
    assert t2.tobytes() != tl.tobytes()
    _subtree_boxes_are_unions(t2, i2, boxes)
    rays = scene_rays(20000, 10 + salt)
    h, s, i, rt = oracle_rt.trace_world(t2, i2, blases, rays, raw_t=True)
    hb, sb, ib, rtb = oracle_rt.trace_world(t2, i2, blases, rays, mode=oracle_rt.WORLD_BRUTE, raw_t=True)
    assert (h["prim_id"] >= 0).mean() > 0.05
    assert np.array_equal(rt.view(np.int32), rtb.view(np.int32))
    # and those are the hits of a world rebuilt at the new positions
    tn, inn, _, _, _, _ = _oracle_world(meshes, moved)
    hn, _, _, rtn = oracle_rt.trace_world(tn, inn, blases, rays, raw_t=True)
    assert np.array_equal(rt.view(np.int32), rtn.view(np.int32))


def _product(meshes, descs):
    blases = [atlas_rt.Blas.build(v) for v in meshes]
    w = atlas_rt.World()
    w.update(descs, blases)
    return w, blases


@pytest.mark.gpu
@pytest.mark.parametrize("salt,n_inst", [(0, 1), (1, 24), (2, 300)])
def test_gpu_refit_equals_oracle(torch_cuda, salt, n_inst):
    print(f"case refit salt={salt} n_inst={n_inst}", flush=True)
    torch = torch_cuda
    meshes, descs = small_scene(salt=salt, n_inst=n_inst)
    w, blases = _product(meshes, descs)
    tl, inst, xform, order, bl, oblases = _oracle_world(meshes, descs)
    gt, gi, go = w.export()
    assert gt.tobytes() == tl.tobytes() and gi.tobytes() == inst.tobytes() and np.array_equal(go, order)
    rays = scene_rays(40000, 20 + salt)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    n = len(rays)
    cur_t, cur_i, cur_x = tl, inst, xform
    for step in range(2):
        moved = _moved(descs, 10 * salt + step, len(meshes))
        keep = w.refit(moved)
        cur_t, cur_i, cur_x, _ = oracle_rt.world_refit(cur_t, cur_i, cur_x, order, moved, bl)
        gt, gi, go = w.export()
        assert gt.tobytes() == cur_t.tobytes(), step
        assert gi.tobytes() == cur_i.tobytes(), step
        d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        d_inst = torch.empty(n, dtype=torch.int32, device="cuda")
        w.trace_closest_rays(d_rays, d_hits, d_inst)
        torch.cuda.synchronize()
        del keep
        hits = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
        hr, sr, ir = oracle_rt.trace_world(cur_t, cur_i, oblases, rays)
        assert np.array_equal(rt_cases.hit_bits(hits), rt_cases.hit_bits(hr)), step
        expect_inst = np.where(ir >= 0, order[np.maximum(ir, 0)].astype(np.int32), -1)
        assert np.array_equal(d_inst.cpu().numpy(), expect_inst), step
        if n_inst > 1:
            assert (hr["prim_id"] >= 0).mean() > 0.02


@pytest.mark.gpu
def test_gpu_pt_follows_a_refit(torch_cuda):
    """A path tracer created before the refit renders the moved scene: per
    pixel against the oracle's path tracer on the refitted arrays."""
    print("case pt refit", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2()
    descs = atlas_rt.instance_descs(xf, which)
    descs["material_offset"] = which.astype(np.int32)
    blases = [atlas_rt.Blas.build(v) for v in meshes]
    for b, nrm in zip(blases, normals):
        b.set_shading(nrm)
    w = atlas_rt.World()
    w.update(descs, blases)
    W_, H_, bounces, frames = 96, 64, 3, 2
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    pt = atlas_rt.PathTracer(w, W_, H_, bounces, mats, W.sun_light(), env=ENV)
    g = W.rng(2, 901)
    xf2 = W.random_affine(g, len(xf), scale=(1.0, 4.0), extent=30.0, height=8.0)
    moved = atlas_rt.instance_descs(xf2, which)
    moved["material_offset"] = which.astype(np.int32)
    keep = w.refit(moved)
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    del keep
    tl, inst, order = w.export()
    obs = [oracle_rt.build_triangles(v) for v in meshes]
    xform = np.zeros((len(order), 12), np.float32)
    for i, r in enumerate(order):
        xform[i] = moved[r]["object_to_world"].reshape(-1)
    shading = [(n[ob.ref_idx], np.zeros(len(ob.ref_idx), np.int32)) for n, ob in zip(normals, obs)]
    ref = oracle_rt.pathtrace(tl, inst, xform, [(ob.nodes, ob.tris) for ob in obs], shading, mats,
                              W.sun_light(), np.array(ENV, np.float32), cam, W_, H_, bounces, frames, 0, 0)
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    print("pt after refit: fraction of matching pixels", close.mean())
    assert close.all()
    assert abs(a.mean() - r.mean()) <= 0.01 * r.mean()
