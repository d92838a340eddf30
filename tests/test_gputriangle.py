"""The engine's 96 B GPUTriangle shading records (RTStructures.h:14-21,
MeshData.cpp:172-240) as the path tracer's shading input.

CPU: the oracle's UnpackTriangle restatement (raytracer/common.hsh:17-44)
pinned independently of its code: unpackUnitVector maps field 0 to -1 and
1023 to +1 and inverts Packing::PackSignedVector3x10_1x2 (Packing.cpp:24-35)
to within its truncation (2/1023); unpackHalf2x16 equals numpy's binary16
conversion for every finite half; unpackUnorm4x8 is each byte over 255
correctly rounded; the material index and opacity come through as stored.
GPU: atlas_rt_blas_set_triangles refuses records of the wrong count or slot
order, and a scene shaded from packed records (vertex normals, texture
coordinates of an opacity map, vertex colours) renders per pixel as the
oracle renders the unpacked values; vertex colours change the image only for
materials that ask for them.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W


def _pack_unit(v):
    """Packing::PackSignedVector3x10_1x2 with w = 0 (MeshData.cpp:199-203)."""
    v = np.asarray(v, np.float32)
    f = lambda x: ((x * np.float32(0.5) + np.float32(0.5)) * np.float32(1023.0)).astype(np.int32)
    w = np.int32(int(np.float32(0.5) * np.float32(2.0)))
    return (f(v[..., 0]) | (f(v[..., 1]) << 10) | (f(v[..., 2]) << 20) | (w << 30)).astype(np.uint32)


def _pack_half2(a, b):
    return (np.asarray(a, np.float16).view(np.uint16).astype(np.uint32)
            | (np.asarray(b, np.float16).view(np.uint16).astype(np.uint32) << 16))


def _pack_unorm4x8(c):
    q = np.round(np.clip(np.asarray(c, np.float32), 0, 1) * 255.0).astype(np.uint32)
    return q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (q[..., 3] << 24)


def _bits(u):
    return np.asarray(u, np.uint32).view(np.float32)


def records_for(tris, orig, normals, uvs, tangents, colors, material, opacity):
    """GPUTriangle records in BVH slot order (tris: the BLAS's exported
    engine triangles; the other arrays per caller triangle)."""
    n = len(tris)
    r = np.zeros(n, dtype=_abi.GPU_TRIANGLE_DTYPE)
    nn = normals[orig].reshape(n, 3, 3)
    for k, f in enumerate(("v0", "v1", "v2")):
        r[f][:, :3] = tris[f][:, :3]
        r[f][:, 3] = _bits(_pack_unit(nn[:, k]))
    uv = uvs[orig]
    r["d0"][:, 0] = _bits(_pack_half2(uv[:, 0], uv[:, 1]))
    r["d0"][:, 1] = _bits(_pack_half2(uv[:, 2], uv[:, 3]))
    r["d0"][:, 2] = _bits(_pack_half2(uv[:, 4], uv[:, 5]))
    r["d0"][:, 3] = np.asarray(material, np.int32)[orig].view(np.float32)
    t = tangents[orig]
    r["d1"][:, 0] = _bits(_pack_unit(t[:, :3]))
    r["d1"][:, 1] = _bits(_pack_unit(t[:, 3:]))
    r["d1"][:, 2] = np.where(tris["v0"][:, 3] > 0, 1.0, -1.0)
    c = colors[orig].reshape(n, 3, 4)
    for k in range(3):
        r["d2"][:, k] = _bits(_pack_unorm4x8(c[:, k]))
    r["d2"][:, 3] = np.asarray(opacity, np.float32)[orig]
    return r


def _random_records(n, seed):
    g = np.random.default_rng(seed)
    r = np.zeros(n, dtype=_abi.GPU_TRIANGLE_DTYPE)
    for f in ("v0", "v1", "v2", "d0", "d1", "d2"):
        r[f] = _bits(g.integers(0, 2 ** 32, (n, 4), dtype=np.uint64).astype(np.uint32))
    r["d2"][:, 3] = g.uniform(0, 1, n).astype(np.float32)
    return r


def test_oracle_unpack_unit_vectors():
    b = np.array([0, 1023, 1023 << 10, 1023 << 20, (511 << 0) | (512 << 10) | (1 << 20)], np.uint32)
    r = np.zeros(len(b), dtype=_abi.GPU_TRIANGLE_DTYPE)
    r["v0"][:, 3] = _bits(b)
    n = oracle_rt.unpack_triangles(r)["normals"][:, :3]
    assert np.array_equal(n[0], [-1, -1, -1]) and np.array_equal(n[1], [1, -1, -1])
    assert np.array_equal(n[2], [-1, 1, -1]) and np.array_equal(n[3], [-1, -1, 1])
    ref = np.array([511, 512, 1], np.float64) / 1023.0 * 2.0 - 1.0
    assert np.allclose(n[4], ref, rtol=0, atol=2e-7)
    # inverts the engine's packing to within its truncation
    g = np.random.default_rng(3)
    v = g.normal(size=(20000, 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    r = np.zeros(len(v), dtype=_abi.GPU_TRIANGLE_DTYPE)
    r["v1"][:, 3] = _bits(_pack_unit(v))
    r["d1"][:, 0] = _bits(_pack_unit(-v))
    u = oracle_rt.unpack_triangles(r)
    for got, want in ((u["normals"][:, 3:6], v), (u["tangents"][:, :3], -v)):
        d = want.astype(np.float64) - got
        assert np.all(d >= -1e-6) and np.all(d <= 2.0 / 1023.0 + 1e-6)


def test_oracle_unpack_halves_colours_and_passthrough():
    h = np.arange(65536, dtype=np.uint32)
    finite = (h & 0x7C00) != 0x7C00
    h = h[finite]
    k = len(h) // 2
    lo, hi = h[:k], h[k:2 * k]
    r = np.zeros(k, dtype=_abi.GPU_TRIANGLE_DTYPE)
    r["d0"][:, 0] = _bits(lo | (hi << 16))
    r["d0"][:, 1] = _bits(hi | (lo << 16))
    u = oracle_rt.unpack_triangles(r)
    f16 = lambda x: x.astype(np.uint16).view(np.float16).astype(np.float32)
    assert np.array_equal(u["uvs"][:, 0], f16(lo)) and np.array_equal(u["uvs"][:, 1], f16(hi))
    assert np.array_equal(u["uvs"][:, 2], f16(hi)) and np.array_equal(u["uvs"][:, 3], f16(lo))
    byte = np.arange(256, dtype=np.uint32)
    r = np.zeros(256, dtype=_abi.GPU_TRIANGLE_DTYPE)
    r["d2"][:, 2] = _bits(byte | ((255 - byte) << 8) | (byte << 16) | (7 << 24))
    c = oracle_rt.unpack_triangles(r)["colors"][:, 8:12]
    assert np.array_equal(c[:, 0], (byte / 255.0).astype(np.float32))
    assert np.array_equal(c[:, 1], ((255 - byte) / 255.0).astype(np.float32))
    assert c[0, 0] == 0.0 and c[255, 0] == 1.0 and np.all(c[:, 3] == np.float32(7 / 255.0))
    rr = _random_records(64, 5)
    u = oracle_rt.unpack_triangles(rr)
    assert np.array_equal(u["material"], rr["d0"][:, 3].view(np.int32))
    assert np.array_equal(u["opacity"], rr["d2"][:, 3])


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _scene_inputs(salt):
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=4, n_inst=40, salt=salt)
    qv, qn = pt_cases.quad_mesh(60.0)
    meshes.append(qv)
    normals.append(qn)
    xf = np.concatenate([xf, np.eye(3, 4)[None] + np.array([[[0, 0, 0, 0], [0, 0, 0, -6.0], [0, 0, 0, 0]]])])
    which = np.concatenate([which, [len(meshes) - 1]])
    mats = np.concatenate([mats, W.c2_materials({"albedo": np.full((1, 3), 0.7, np.float32),
                                                  "roughness": np.ones(1, np.float32)})])
    g = np.random.default_rng(salt)
    per = []
    for k, v in enumerate(meshes):
        n = len(v)
        uvs = pt_cases.sphere_uvs(v)
        tangents = np.tile(np.array([1, 0, 0, 0, 0, 1], np.float32), (n, 1))
        colors = g.uniform(0.2, 1.0, (n, 12)).astype(np.float32)
        opacity = np.full(n, mats["opacity"][k], np.float32)
        if k == 2:
            opacity[:] = -1.0  # opacity map
        per.append(dict(uvs=uvs, tangents=tangents, colors=colors, material=np.zeros(n, np.int32),
                        opacity=opacity))
    mats["opacity_texture"][2] = 0
    mats["opacity"][2] = 0.9
    mats["use_vertex_colors"][[0, 3, 4]] = 1
    return meshes, normals, mats, xf, which, per


@pytest.mark.gpu
def test_set_triangles_refuses_wrong_records(torch_cuda):
    v, n = pt_cases.quad_mesh(5.0)
    blas = atlas_rt.Blas.build(np.concatenate([v, v + 1.0]))
    _, tris, orig = blas.export()
    nn = np.tile(n, (2, 1))
    rec = records_for(tris, orig, nn, np.zeros((4, 6), np.float32), np.zeros((4, 6), np.float32),
                      np.ones((4, 12), np.float32), np.zeros(4, np.int32), np.ones(4, np.float32))
    blas.set_triangles(rec)
    with pytest.raises(atlas_rt.AtlasRTError) as e:
        blas.set_triangles(rec[:-1])
    assert e.value.code == _abi.E_SIZE_MISMATCH
    with pytest.raises(atlas_rt.AtlasRTError) as e:
        blas.set_triangles(rec[::-1])
    assert e.value.code == _abi.E_INVALID_ARG
    blas.close()


@pytest.mark.gpu
def test_gpu_pt_from_packed_records_matches_oracle(torch_cuda):
    print("case pt packed triangles", flush=True)
    meshes, normals, mats, xf, which, per = _scene_inputs(21)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    unpacked = []
    for b, nrm, p in zip(blases, normals, per):
        _, tris, orig = b.export()
        rec = records_for(tris, orig, nrm, p["uvs"], p["tangents"], p["colors"], p["material"], p["opacity"])
        b.set_triangles(rec)
        unpacked.append(oracle_rt.unpack_triangles(rec))
    # the oracle shades from the unpacked values (slot order, as the records)
    S["shading"] = [(u["normals"], u["material"]) for u in unpacked]
    tex = [pt_cases.checker(8)]
    surf = {"opacity": [u["opacity"] for u in unpacked], "uvs": [u["uvs"] for u in unpacked],
            "tangents": [u["tangents"] for u in unpacked], "colors": [u["colors"] for u in unpacked],
            "textures": tex}
    W_, H_, bounces, frames = 96, 64, 3, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    sun = W.sun_light()
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=W.C2_ENV)
    pt.set_textures(tex)
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    ref = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                              np.array(W.C2_ENV, np.float32), cam, W_, H_, bounces, frames, surface=surf)
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    same = np.all(img.view(np.uint32) == ref.view(np.uint32), axis=-1)
    print(f"pt packed: {close.mean():.4f} of pixels within tolerance, {same.mean():.4f} bit for bit")
    assert close.all() and abs(a.mean() - r.mean()) <= 0.005 * r.mean()
    # vertex colours matter where the material asks for them
    mats2 = mats.copy()
    mats2["use_vertex_colors"] = 0
    ref2 = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats2, sun,
                               np.array(W.C2_ENV, np.float32), cam, W_, H_, bounces, frames, surface=surf)
    assert np.abs(ref2[..., :3] - ref[..., :3]).max() > 1e-2
    pt.close()
