"""Writes the golden fixtures of tests/test_golden.py with the oracle.

    python tests/golden/make_golden.py

blas256.npz   a 256-triangle soup (config-0 recipe in a 4-unit box, seed
              0xA71A5 + 0, salt 256), its BVH in the engine layout, 4,096
              rays from a 6-unit box and the
              engine-order closest hits (t, u, v, prim, slot) and any-hit flags.
world4x64.npz four 64-triangle meshes, 16 instances, the TLAS and packed
              instances, 4,096 rays and the closest (t, u, v, prim, slot,
              instance) and any-hit answers for both instance masks.
pt32.npz      a 32 x 32 path-traced image of 12 blobs over a quad, 4 bounces,
              sun + constant environment, 1,024 samples per pixel: the mean
              radiance and its per-pixel standard error.
The oracle is deterministic, so rerunning the script reproduces the files
bit for bit.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("atlas-engine_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import atlas_rt  # noqa: E402  (numpy-only helpers; no device needed)
import oracle_rt  # noqa: E402
import pt_cases  # noqa: E402
from atlas_rt import _abi  # noqa: E402
from atlas_rt import workloads as W  # noqa: E402


def blas256():
    g = W.rng(0, 256)
    verts = W.triangle_soup(256, g, extent=2.0, edge=(0.2, 0.8))
    rays = W.random_rays(4096, g, extent=3.0)
    b = oracle_rt.build_triangles(verts)
    h, s = oracle_rt.trace_closest(b.nodes, b.tris, b.ref_idx, rays)
    anyh = oracle_rt.trace_any(b.nodes, b.tris, rays)
    np.savez_compressed(os.path.join(HERE, "blas256.npz"), verts=verts, rays=rays, nodes=b.nodes, tris=b.tris,
                        ref_idx=b.ref_idx, hits=h, slots=s, any=anyh)


def world4x64():
    g = W.rng(2, 4064)
    meshes = [W.triangle_soup(64, g, extent=2.0, edge=(0.2, 0.8)) for _ in range(4)]
    xf = W.random_affine(g, 16, scale=(0.5, 2.0), extent=8.0, height=4.0)
    which = g.integers(0, 4, 16)
    descs = atlas_rt.instance_descs(xf, which)
    descs["mask"][::3] = _abi.MASK_ALL
    boxes = np.zeros((16, 6), np.float32)
    for i, d in enumerate(descs):
        tri = meshes[d["blas"]].reshape(-1, 3)
        lo, hi = tri.min(axis=0), tri.max(axis=0)
        c = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float64)
        m = d["object_to_world"].astype(np.float64)
        p = c @ m[:, :3].T + m[:, 3]
        boxes[i, :3], boxes[i, 3:] = p.min(axis=0), p.max(axis=0)
    bl = [oracle_rt.build_triangles(v) for v in meshes]
    tl = oracle_rt.build_aabbs(boxes)
    inst = np.zeros(len(tl.ref_idx), dtype=oracle_rt.INSTANCE_DTYPE)
    for i, r in enumerate(tl.ref_idx):
        inst[i]["inverse_matrix"] = descs[r]["world_to_object"]
        inst[i]["mesh_offset"] = descs[r]["blas"]
        inst[i]["material_offset"] = 0
        inst[i]["next_instance"] = -1 if tl.end_of_node[i] else i + 1
        inst[i]["mask"] = descs[r]["mask"]
    rays = W.random_rays(4096, g, extent=10.0)
    blases = [(b.nodes, b.tris, b.ref_idx) for b in bl]
    out = dict(rays=rays, boxes=boxes, descs=descs, tlas=tl.nodes, order=tl.ref_idx, inst=inst)
    for k, v in enumerate(meshes):
        out[f"mesh{k}"] = v
        out[f"nodes{k}"], out[f"tris{k}"], out[f"ref{k}"] = bl[k].nodes, bl[k].tris, bl[k].ref_idx
    for name, mask in (("all", _abi.MASK_ALL), ("shadow", _abi.MASK_SHADOW)):
        h, s, i = oracle_rt.trace_world(tl.nodes, inst, blases, rays, mask=mask)
        a, _, _ = oracle_rt.trace_world(tl.nodes, inst, blases, rays, mask=mask, mode=oracle_rt.WORLD_ANY_RESTORE)
        out[f"hits_{name}"], out[f"slots_{name}"], out[f"inst_{name}"], out[f"any_{name}"] = h, s, i, a["prim_id"]
    np.savez_compressed(os.path.join(HERE, "world4x64.npz"), **out)


PT32 = dict(width=32, height=32, bounces=4, spp=1024)


def pt32():
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=3, n_inst=12, salt=32, extent=6.0, height=2.0)
    qv, qn = pt_cases.quad_mesh(20.0)
    meshes.append(qv)
    normals.append(qn)
    xf = np.concatenate([xf, np.eye(3, 4)[None] + np.array([[[0, 0, 0, 0], [0, 0, 0, -4.0], [0, 0, 0, 0]]])])
    which = np.concatenate([which, [len(meshes) - 1]])
    mats = np.concatenate([mats, W.c2_materials({"albedo": np.full((1, 3), 0.6, np.float32),
                                                  "roughness": np.ones(1, np.float32)})])
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    cam = atlas_rt.camera((0.0, 6.0, -16.0), 0.0, 20.0, 50.0, 1.0)
    w, h, spp = PT32["width"], PT32["height"], PT32["spp"]
    sums = np.zeros((h, w, 3))
    sq = np.zeros((h, w, 3))
    for f in range(spp):
        img = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], S["materials"],
                                  W.sun_light(), np.array(W.C2_ENV, np.float32), cam, w, h, PT32["bounces"], 1,
                                  first_sample=f, pcg_draws=f * (PT32["bounces"] + 1))
        x = img[..., :3].astype(np.float64)
        sums += x
        sq += x * x
    mean = sums / spp
    var = np.maximum(sq / spp - mean * mean, 0.0)
    np.savez_compressed(os.path.join(HERE, "pt32.npz"), mean=mean.astype(np.float32),
                        stderr=np.sqrt(var / spp).astype(np.float32), meshes=np.array(len(meshes)))


if __name__ == "__main__":
    which = sys.argv[1:] or ["blas256", "world4x64", "pt32"]
    for name in which:
        globals()[name]()
        print("wrote", name)
