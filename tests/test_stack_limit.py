"""Traversal stack bound and its overflow status.

The engine's traversal stack is a fixed 32-entry array with no overflow check
(data/shader/raytracer/bvh.hsh:16).  Here the stack is sized from the tree by
default, so it never overflows; a caller may cap it (for instance at the
engine's 32), and then a ray that would push past the cap stops as a miss and
raises a flag that the status calls return as ATLAS_RT_E_STACK_OVERFLOW.

The hand-built 80-level chain needs one stack entry per level for rays along
+z, so a cap of 32 overflows on every ray and a cap of 81 on none.  Every
kernel still exits normally: the cap is a bound the kernel checks before each
push.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

pytestmark = pytest.mark.gpu

BINARY_KERNELS = [1, 1 | 0x10, 1 | 0x20, 2, 2 | 0x10, 2 | 0x40, 2 | 0x50]


def _dev(torch, rays):
    return torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()


def _closest(torch, blas, rays, kernel):
    n = len(rays)
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(n, dtype=torch.int32, device="cuda")
    blas.trace_closest(_dev(torch, rays), d_hits, d_slots, kernel=kernel)
    torch.cuda.synchronize()
    return d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1), d_slots.cpu().numpy()


@pytest.mark.parametrize("kernel", BINARY_KERNELS)
def test_blas_stack_limit_reports_overflow(torch_cuda, kernel):
    print(f"case blas stack limit kernel={kernel}", flush=True)
    torch = torch_cuda
    nodes, tris, rays = rt_cases.chain_tree(80)
    blas = atlas_rt.Blas.create(nodes, tris)
    assert blas.trace_status() == _abi.OK  # no trace yet
    blas.set_stack_limit(32)
    h, s = _closest(torch, blas, rays, kernel)
    assert np.all(h["prim_id"] == -1) and np.all(s == -1)
    assert np.array_equal(h["t"], rays["tmax"])
    assert blas.trace_status() == _abi.E_STACK_OVERFLOW
    assert blas.trace_status() == _abi.OK  # the flag was cleared
    blas.set_stack_limit(81)  # enough for the deepest push
    h, s = _closest(torch, blas, rays, kernel)
    assert np.all(h["prim_id"] == 0) and np.all(h["t"] == np.float32(105.0))
    assert blas.trace_status() == _abi.OK
    blas.set_stack_limit(0)  # sized from the tree
    h, s = _closest(torch, blas, rays, kernel)
    assert np.all(h["prim_id"] == 0)
    assert blas.trace_status() == _abi.OK
    blas.close()


def test_blas_any_hit_stack_limit(torch_cuda):
    print("case blas any-hit stack limit", flush=True)
    torch = torch_cuda
    nodes, tris, rays = rt_cases.chain_tree(80)
    blas = atlas_rt.Blas.create(nodes, tris)
    d = _dev(torch, rays)
    out = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    blas.set_stack_limit(32)
    blas.trace_any(d, out)
    assert np.all(out.cpu().numpy() == 0)
    assert blas.trace_status() == _abi.E_STACK_OVERFLOW
    blas.set_stack_limit(0)
    blas.trace_any(d, out)
    assert np.array_equal(out.cpu().numpy().astype(np.uint32), oracle_rt.trace_any(nodes, tris, rays))
    assert blas.trace_status() == _abi.OK
    blas.close()


def test_shallow_tree_under_the_engine_limit_is_unaffected(torch_cuda):
    """A 65,536-triangle tree stays well within 32 entries: with the engine's
    cap the results are the engine's and no overflow is reported."""
    print("case stack limit on a normal tree", flush=True)
    torch = torch_cuda
    verts, rays = rt_cases.soup(65536, 131072, 81)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    blas.set_stack_limit(32)
    hr, sr = oracle_rt.trace_closest(nodes, tris, orig, rays)
    for kernel in (1, 2 | 0x40):
        h, s = _closest(torch, blas, rays, kernel)
        assert np.array_equal(rt_cases.hit_bits(h), rt_cases.hit_bits(hr)) and np.array_equal(s, sr)
    assert blas.trace_status() == _abi.OK
    blas.close()


def _chain_world():
    nodes, tris, rays = rt_cases.chain_tree(80)
    blas = atlas_rt.Blas.create(nodes, tris)
    xf = np.zeros((1, 3, 4))
    xf[0, :, :3] = np.eye(3)
    descs = atlas_rt.instance_descs(xf, np.zeros(1, np.int64))
    w = atlas_rt.World()
    w.update(descs, [blas])
    return w, blas, rays


def test_world_stack_limit_reports_overflow(torch_cuda):
    print("case world stack limit", flush=True)
    torch = torch_cuda
    w, blas, rays = _chain_world()
    n = len(rays)
    d = _dev(torch, rays)
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    w.set_stack_limit(32)
    w.trace_closest_rays(d, hits)
    h = hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
    assert np.all(h["prim_id"] == -1)
    assert w.trace_status() == _abi.E_STACK_OVERFLOW
    assert w.trace_status() == _abi.OK
    w.set_stack_limit(0)
    w.trace_closest_rays(d, hits)
    h = hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
    assert np.all(h["prim_id"] == 0) and np.all(h["t"] == np.float32(105.0))
    assert w.trace_status() == _abi.OK
    w.close()
    blas.close()


def test_pathtracer_follows_the_world_stack_limit(torch_cuda):
    print("case path tracer stack limit", flush=True)
    w, blas, _ = _chain_world()
    blas.set_shading(np.tile(np.array([0, 0, -1] * 3, np.float32), (81, 1)))
    mats = W.c2_materials({"albedo": np.full((1, 3), 0.5, np.float32), "roughness": np.ones(1, np.float32)})
    # camera on the -z side looking along +z at the stacked quads
    cam = atlas_rt.camera((0.5, 0.5, -20.0), 0.0, 0.0, 10.0, 1.0)
    pt = atlas_rt.PathTracer(w, 16, 16, 1, mats, W.sun_light(), env=(0.2, 0.2, 0.2))
    pt.render(cam, 1)
    assert pt.status() == _abi.OK
    w.set_stack_limit(32)
    pt.render(cam, 1)
    assert pt.status() == _abi.E_STACK_OVERFLOW
    assert pt.status() == _abi.OK
    w.set_stack_limit(0)
    pt.render(cam, 1)
    assert pt.status() == _abi.OK
    pt.close()
    w.close()
    blas.close()
