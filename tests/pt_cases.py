"""Scenes for the path-tracer tests (oracle-only arrays and product objects)."""
import numpy as np

import atlas_rt
import oracle_rt
from atlas_rt import _abi
from atlas_rt import workloads as W


def quad_mesh(size=50.0):
    """Two triangles spanning [-size, size]^2 in the y = 0 plane, geometric
    normal cross(v0 - v1, v0 - v2) = +y, vertex normals +y."""
    s = size
    a, b, c, d = [-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]
    verts = np.array([a + c + b, a + d + c], dtype=np.float32)
    t = verts.reshape(-1, 3, 3).astype(np.float64)
    ng = np.cross(t[:, 0] - t[:, 1], t[:, 0] - t[:, 2])
    assert np.all(ng[:, 1] > 0)
    normals = np.tile(np.array([0, 1, 0] * 3, np.float32), (2, 1))
    return verts, normals


def oracle_scene(meshes, normals, mats_per_mesh, xf, which, mask=_abi.MASK_ALL | _abi.MASK_SHADOW):
    """Oracle-only world + shading arrays (builders restated in the oracle)."""
    descs = atlas_rt.instance_descs(xf, which, mask=mask)
    descs["material_offset"] = which.astype(np.int32)
    blases, shading, refs = [], [], []
    for v, n in zip(meshes, normals):
        ob = oracle_rt.build_triangles(v)
        blases.append((ob.nodes, ob.tris))
        shading.append((n[ob.ref_idx], np.zeros(len(ob.ref_idx), np.int32)))
        refs.append(ob.ref_idx)
    boxes = np.zeros((len(descs), 6), np.float32)
    for i, dsc in enumerate(descs):
        tri = meshes[dsc["blas"]].reshape(-1, 3).astype(np.float64)
        lo, hi = tri.min(axis=0), tri.max(axis=0)
        c = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        m = dsc["object_to_world"].astype(np.float64)
        p = c @ m[:, :3].T + m[:, 3]
        boxes[i, :3], boxes[i, 3:] = p.min(axis=0) - 1e-3, p.max(axis=0) + 1e-3
    tl = oracle_rt.build_aabbs(boxes)
    inst = np.zeros(len(tl.ref_idx), dtype=oracle_rt.INSTANCE_DTYPE)
    xform = np.zeros((len(tl.ref_idx), 12), np.float32)
    for i, r in enumerate(tl.ref_idx):
        inst[i]["inverse_matrix"] = descs[r]["world_to_object"]
        inst[i]["mesh_offset"] = descs[r]["blas"]
        inst[i]["material_offset"] = descs[r]["material_offset"]
        inst[i]["next_instance"] = -1 if tl.end_of_node[i] else i + 1
        inst[i]["mask"] = descs[r]["mask"]
        xform[i] = descs[r]["object_to_world"].reshape(-1)
    return dict(tlas=tl.nodes, inst=inst, xform=xform, blases=blases, shading=shading, materials=mats_per_mesh,
                descs=descs, boxes=boxes, refs=refs, tlas_order=tl.ref_idx)


def small_c2(n_meshes=4, n_inst=48, salt=0, extent=30.0, height=8.0):
    g = W.rng(2, 500 + salt)
    meshes, normals = [], []
    for _ in range(n_meshes):
        v, n = W.blob_mesh(g, stacks=12, slices=24)
        meshes.append(v)
        normals.append(n)
    mats = W.c2_materials({"albedo": g.uniform(0.1, 0.9, (n_meshes, 3)).astype(np.float32),
                           "roughness": g.uniform(0.05, 1.0, n_meshes).astype(np.float32)})
    xf = W.random_affine(g, n_inst, scale=(1.0, 4.0), extent=extent, height=height)
    which = g.integers(0, n_meshes, n_inst)
    return meshes, normals, mats, xf, which


def product_scene(meshes, normals, xf, which, boxes):
    """Product world over the same meshes/instances (needs a GPU)."""
    blases = [atlas_rt.Blas.build(v) for v in meshes]
    for b, n in zip(blases, normals):
        b.set_shading(n)
    descs = atlas_rt.instance_descs(xf, which)
    descs["material_offset"] = which.astype(np.int32)
    w = atlas_rt.World()
    w.update(descs, blases, world_aabbs=boxes)
    return w, blases


def quad_xy(z, size=1.0, normal_z=-1.0):
    """Two triangles spanning [-size, size]^2 in the plane z, vertex normals
    (0, 0, normal_z)."""
    s = size
    a, b, c, d = [-s, -s, z], [s, -s, z], [s, s, z], [-s, s, z]
    verts = np.array([a + c + b, a + d + c], dtype=np.float32)
    normals = np.tile(np.array([0, 0, normal_z] * 3, np.float32), (2, 1))
    return verts, normals


def sphere_uvs(verts):
    """Per-triangle (uv0, uv1, uv2) from each vertex's direction (longitude,
    colatitude), (n, 6) float32."""
    p = verts.reshape(-1, 3).astype(np.float64)
    r = np.linalg.norm(p, axis=1) + 1e-12
    u = np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(p[:, 1] / r, -1, 1)) / np.pi
    return np.stack([u, v], axis=1).reshape(-1, 6).astype(np.float32)


def checker(n=8, lo=0, hi=255):
    """(n, n) uint8 checkerboard."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return np.where((i + j) % 2 == 0, hi, lo).astype(np.uint8)
