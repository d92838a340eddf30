"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py
with the oracle): the oracle and the product's host builders reproduce them
exactly on the CPU, and the GPU kernels reproduce them on the device."""
import os

import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_oracle_reproduces_blas256():
    d = _load("blas256.npz")
    b = oracle_rt.build_triangles(d["verts"])
    assert b.nodes.tobytes() == d["nodes"].tobytes() and b.tris.tobytes() == d["tris"].tobytes()
    assert np.array_equal(b.ref_idx, d["ref_idx"])
    h, s = oracle_rt.trace_closest(d["nodes"], d["tris"], d["ref_idx"], d["rays"])
    assert h.tobytes() == d["hits"].tobytes() and np.array_equal(s, d["slots"])
    assert np.array_equal(oracle_rt.trace_any(d["nodes"], d["tris"], d["rays"]), d["any"])
    assert (d["slots"] >= 0).mean() > 0.01


def test_product_builder_reproduces_blas256_on_host():
    d = _load("blas256.npz")
    blas = atlas_rt.Blas.build(d["verts"], host_only=True)
    nodes, tris, orig = blas.export()
    assert nodes.tobytes() == d["nodes"].tobytes() and tris.tobytes() == d["tris"].tobytes()
    assert np.array_equal(orig, d["ref_idx"])


def _world_blases(d):
    return [(d[f"nodes{k}"], d[f"tris{k}"], d[f"ref{k}"]) for k in range(4)]


def test_oracle_and_host_tlas_reproduce_world4x64():
    d = _load("world4x64.npz")
    tl = oracle_rt.build_aabbs(d["boxes"])
    assert tl.nodes.tobytes() == d["tlas"].tobytes() and np.array_equal(tl.ref_idx, d["order"])
    nodes, order = atlas_rt.aabb_tree_build(d["boxes"])[:2]
    assert nodes.tobytes() == d["tlas"].tobytes() and np.array_equal(order, d["order"])
    for name, mask in (("all", _abi.MASK_ALL), ("shadow", _abi.MASK_SHADOW)):
        h, s, i = oracle_rt.trace_world(d["tlas"], d["inst"], _world_blases(d), d["rays"], mask=mask)
        assert h.tobytes() == d[f"hits_{name}"].tobytes() and np.array_equal(i, d[f"inst_{name}"])
        a, _, _ = oracle_rt.trace_world(d["tlas"], d["inst"], _world_blases(d), d["rays"], mask=mask,
                                        mode=oracle_rt.WORLD_ANY_RESTORE)
        assert np.array_equal(a["prim_id"], d[f"any_{name}"])
    assert (d["slots_all"] >= 0).mean() > 0.05


def _pt32_scene():
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=3, n_inst=12, salt=32, extent=6.0, height=2.0)
    qv, qn = pt_cases.quad_mesh(20.0)
    meshes.append(qv)
    normals.append(qn)
    xf = np.concatenate([xf, np.eye(3, 4)[None] + np.array([[[0, 0, 0, 0], [0, 0, 0, -4.0], [0, 0, 0, 0]]])])
    which = np.concatenate([which, [len(meshes) - 1]])
    mats = np.concatenate([mats, W.c2_materials({"albedo": np.full((1, 3), 0.6, np.float32),
                                                  "roughness": np.ones(1, np.float32)})])
    return meshes, normals, mats, xf, which


def test_oracle_reproduces_pt32_first_frames():
    d = _load("pt32.npz")
    meshes, normals, mats, xf, which = _pt32_scene()
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    cam = atlas_rt.camera((0.0, 6.0, -16.0), 0.0, 20.0, 50.0, 1.0)
    img = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], S["materials"],
                              W.sun_light(), np.array(W.C2_ENV, np.float32), cam, 32, 32, 4, 64)
    m64 = img[..., :3] / 64.0
    # 64 of the 1,024 samples: within 5 standard errors of the fixture's mean
    se64 = np.maximum(d["stderr"] * np.sqrt(1024 / 64), 1e-3)
    assert np.mean(np.abs(m64 - d["mean"]) <= 5 * se64) > 0.99
    assert abs(m64.mean() - d["mean"].mean()) < 3 * se64.mean() / np.sqrt(m64.size) + 1e-3


@pytest.mark.gpu
def test_gpu_reproduces_golden_traces(torch_cuda):
    print("case golden traces", flush=True)
    torch = torch_cuda
    d = _load("blas256.npz")
    blas = atlas_rt.Blas.create(d["nodes"], d["tris"], d["ref_idx"])
    rays = d["rays"]
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    for kernel in (1, 2):
        blas.trace_closest(d_rays, d_hits, d_slots, kernel=kernel)
        torch.cuda.synchronize()
        assert d_hits.cpu().numpy().tobytes() == d["hits"].tobytes()
        assert np.array_equal(d_slots.cpu().numpy(), d["slots"])
    d_any = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    blas.trace_any(d_rays, d_any)
    torch.cuda.synchronize()
    assert np.array_equal(d_any.cpu().numpy().astype(np.uint32), d["any"])

    w = _load("world4x64.npz")
    blases = [atlas_rt.Blas.build(w[f"mesh{k}"]) for k in range(4)]
    for k, b in enumerate(blases):
        assert b.export()[0].tobytes() == w[f"nodes{k}"].tobytes()
    world = atlas_rt.World()
    world.update(w["descs"], blases, world_aabbs=w["boxes"])
    rays = w["rays"]
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    for name, mask in (("all", _abi.MASK_ALL), ("shadow", _abi.MASK_SHADOW)):
        d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        d_inst = torch.empty(len(rays), dtype=torch.int32, device="cuda")
        world.trace_closest_rays(d_rays, d_hits, d_inst, mask=mask)
        d_any = torch.empty(len(rays), dtype=torch.int32, device="cuda")
        world.trace_any_rays(d_rays, d_any, mask=mask, flags=_abi.ANY_RESTORE_ALWAYS)
        torch.cuda.synchronize()
        assert d_hits.cpu().numpy().tobytes() == w[f"hits_{name}"].tobytes()
        ir = w[f"inst_{name}"]
        assert np.array_equal(d_inst.cpu().numpy(), np.where(ir >= 0, w["order"][np.maximum(ir, 0)].astype(np.int32),
                                                             -1))
        assert np.array_equal(d_any.cpu().numpy(), w[f"any_{name}"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_image(torch_cuda):
    print("case golden image", flush=True)
    d = _load("pt32.npz")
    meshes, normals, mats, xf, which = _pt32_scene()
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, _ = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    cam = atlas_rt.camera((0.0, 6.0, -16.0), 0.0, 20.0, 50.0, 1.0)
    pt = atlas_rt.PathTracer(w_, 32, 32, 4, mats, W.sun_light(), env=W.C2_ENV)
    pt.render(cam, 1024)
    img = pt.read().cpu().numpy()
    assert np.all(img[..., 3] == 1024)
    m = img[..., :3].astype(np.float64) / 1024.0
    ref = d["mean"].astype(np.float64)
    close = np.all(np.abs(m - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    print(f"golden image: {close.mean():.4f} of pixels match, mean {m.mean():.5f} vs {ref.mean():.5f}")
    assert close.all()
    assert abs(m.mean() - ref.mean()) <= 3 * float(np.sqrt(np.mean(d["stderr"] ** 2) / m.size)) + 1e-4
