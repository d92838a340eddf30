"""Material maps in the bounce shader: base colour, roughness, metalness, AO and
normal maps (GetSurfaceParameters, raytracer/surface.hsh:86-130; the
Sample*Bilinear functions of raytracer/texture.hsh).

CPU: the oracle pinned by reductions to the untextured shader: maps sampled
at texel centres multiply the material exactly as premultiplied values do
(the image is identical bit for bit), and a normal map with normal_scale 0
changes nothing.  GPU: the same identities on the device, and every map
together (RGBA base colour, R8 roughness / metalness / AO, RGB normal map with
a tangent frame, invert_uvs) per pixel against the oracle.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import workloads as W

ENV = W.C2_ENV


def _scene(n_meshes=4, n_inst=40, salt=13):
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=n_meshes, n_inst=n_inst, salt=salt)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    return meshes, normals, mats, xf, which, S


def _centre_uvs(n, tex_size, texel):
    """Every vertex at the centre of one texel.  The interpolated coordinate
    can miss the centre by an ulp, and then the bilinear footprint reaches a
    neighbour, so the maps below are 3 x 3 blocks of one value, sampled at a
    block's centre, holding only 0 and 255: any blend of equal 0 / 1 taps is
    exactly 0 / 1."""
    u = (texel[0] + 0.5) / tex_size
    v = (texel[1] + 0.5) / tex_size
    return np.tile(np.array([u, v] * 3, np.float32), (n, 1))


def _tangents(verts):
    """A tangent frame per triangle: its first edge and the in-plane
    perpendicular (any frame will do for parity)."""
    t = verts.reshape(-1, 3, 3).astype(np.float64)
    e0 = t[:, 1] - t[:, 0]
    n = np.cross(e0, t[:, 2] - t[:, 0])
    bt = np.cross(n, e0)
    e0 /= np.linalg.norm(e0, axis=1, keepdims=True) + 1e-30
    bt /= np.linalg.norm(bt, axis=1, keepdims=True) + 1e-30
    return np.concatenate([e0, bt], axis=1).astype(np.float32)


def _render(S, mats, surface, W_=64, H_=48, bounces=3, frames=2):
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    return oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, W.sun_light(),
                               np.array(ENV, np.float32), cam, W_, H_, bounces, frames, surface=surface)


def _premultiply_case():
    meshes, normals, mats, xf, which, S = _scene()
    size = 4
    g = np.random.default_rng(7)
    rgba = (g.integers(0, 2, (size, size, 4)) * 255).astype(np.uint8)
    rgba[..., 0] = 255  # every mesh keeps some red
    r8 = (g.integers(0, 2, (size, size)) * 255).astype(np.uint8)
    ones = np.full((size, size), 255, np.uint8)
    block = np.ones((3, 3), np.uint8)
    tex = [np.stack([np.kron(rgba[..., c], block) for c in range(4)], axis=-1), np.kron(r8, block),
           np.kron(ones, block)]
    # mesh k samples block (k, k + 1) of every map, at its centre texel
    texel = [(k % size, (k + 1) % size) for k in range(len(meshes))]
    uvs = [_centre_uvs(len(v), 3 * size, (3 * texel[k][0] + 1, 3 * texel[k][1] + 1)) for k, v in enumerate(meshes)]
    textured = mats.copy()
    textured["base_color_texture"] = 0
    textured["roughness_texture"] = 2
    textured["metalness_texture"] = 1
    textured["ao_texture"] = 1
    textured["metalness"] = 0.5
    plain = textured.copy()
    for f in ("base_color_texture", "roughness_texture", "metalness_texture", "ao_texture"):
        plain[f] = -1
    for k in range(len(meshes)):
        x, y = texel[k]
        c = rgba[y, x].astype(np.float32) / np.float32(255.0)
        plain["base_color"][k] = textured["base_color"][k] * c[:3]
        r = np.float32(r8[y, x]) / np.float32(255.0)
        plain["roughness"][k] = textured["roughness"][k] * np.float32(1.0)
        plain["metalness"][k] = textured["metalness"][k] * r
        plain["ao"][k] = textured["ao"][k] * r
    return meshes, normals, xf, which, S, tex, uvs, textured, plain


def test_oracle_maps_at_texel_centres_equal_premultiplied_materials():
    meshes, normals, xf, which, S, tex, uvs, textured, plain = _premultiply_case()
    surf = {"uvs": [uv[r] for uv, r in zip(uvs, S["refs"])], "textures": tex}
    a = _render(S, textured, surf)
    b = _render(S, plain, surf)
    assert a.tobytes() == b.tobytes()
    assert np.abs(a - _render(S, textured, {"uvs": surf["uvs"]})).max() > 0.01  # the maps matter


def test_oracle_normal_map_with_zero_scale_changes_nothing():
    meshes, normals, mats, xf, which, S = _scene()
    nm = np.random.default_rng(9).integers(0, 256, (8, 8, 3)).astype(np.uint8)
    uvs = [pt_cases.sphere_uvs(v) for v in meshes]
    tangents = [_tangents(v) for v in meshes]
    surf = {"uvs": [uv[r] for uv, r in zip(uvs, S["refs"])], "tangents": [t[r] for t, r in zip(tangents, S["refs"])],
            "textures": [nm]}
    m0 = mats.copy()
    m0["normal_texture"] = 0
    m0["normal_scale"] = 0.0
    assert _render(S, m0, surf).tobytes() == _render(S, mats, surf).tobytes()
    m1 = m0.copy()
    m1["normal_scale"] = 1.0
    assert np.abs(_render(S, m1, surf) - _render(S, mats, surf)).max() > 0.01


def _product(meshes, normals, xf, which, S, uvs=None, tangents=None):
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    for k, b in enumerate(blases):
        if uvs is not None:
            b.set_opacity(None, uvs[k])
        if tangents is not None:
            b.set_tangents(tangents[k])
    return w_, blases


@pytest.mark.gpu
def test_gpu_maps_at_texel_centres_equal_premultiplied_materials(torch_cuda):
    print("case pt maps premultiplied", flush=True)
    meshes, normals, xf, which, S, tex, uvs, textured, plain = _premultiply_case()
    w_, blases = _product(meshes, normals, xf, which, S, uvs=uvs)
    W_, H_ = 64, 48
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    imgs = []
    for mats in (textured, plain):
        pt = atlas_rt.PathTracer(w_, W_, H_, 3, mats, W.sun_light(), env=ENV)
        pt.set_textures(tex)
        pt.render(cam, 2)
        imgs.append(pt.read().cpu().numpy())
        pt.close()
    assert imgs[0].tobytes() == imgs[1].tobytes()


@pytest.mark.gpu
def test_gpu_all_maps_match_oracle_per_pixel(torch_cuda):
    print("case pt all maps", flush=True)
    meshes, normals, mats, xf, which, S = _scene(n_meshes=4, n_inst=48, salt=21)
    g = np.random.default_rng(10)
    tex = [g.integers(0, 256, (16, 16, 4)).astype(np.uint8),   # base colour
           g.integers(40, 256, (8, 8)).astype(np.uint8),        # roughness
           g.integers(0, 256, (8, 8)).astype(np.uint8),         # metalness
           g.integers(128, 256, (8, 8)).astype(np.uint8),       # AO
           g.integers(0, 256, (16, 16, 3)).astype(np.uint8)]    # normal map
    m = mats.copy()
    m["base_color_texture"] = 0
    m["roughness_texture"] = [1, 1, -1, 1]
    m["metalness_texture"] = [2, -1, 2, 2]
    m["metalness"] = 0.7
    m["ao_texture"] = 3
    m["normal_texture"] = [4, 4, 4, -1]
    m["normal_scale"] = [1.0, 0.5, 2.0, 1.0]
    m["invert_uvs"] = [0, 1, 0, 1]
    uvs = [pt_cases.sphere_uvs(v) for v in meshes]
    tangents = [_tangents(v) for v in meshes]
    w_, blases = _product(meshes, normals, xf, which, S, uvs=uvs, tangents=tangents)
    W_, H_, bounces, frames = 96, 64, 4, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, m, W.sun_light(), env=ENV)
    pt.set_textures(tex)
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    surf = {"uvs": [uv[r] for uv, r in zip(uvs, S["refs"])], "tangents": [t[r] for t, r in zip(tangents, S["refs"])],
            "textures": tex}
    ref = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], m, W.sun_light(),
                              np.array(ENV, np.float32), cam, W_, H_, bounces, frames, surface=surf)
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    print("pt all maps: fraction of matching pixels", close.mean(), "mean", a.mean(), r.mean())
    assert close.all()
    assert abs(a.mean() - r.mean()) <= 0.01 * r.mean()
    plain = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, W.sun_light(),
                                np.array(ENV, np.float32), cam, W_, H_, bounces, frames)
    assert np.abs(plain - ref).max() > 0.01
