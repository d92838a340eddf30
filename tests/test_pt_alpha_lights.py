"""Opacity-aware traversal and triangle lights.

CPU: closed-form pins of the oracle's HitAnyTransparency (including the
engine's per-leaf squaring), HitClosestTransparency, the bilinear opacity
sampler, the triangle-light list and direct lighting from a small area light.
GPU: the world-level opacity traces bit for bit against the oracle, and the
path tracer with cut-outs, translucent materials, an opacity map and
triangle lights per pixel against the oracle.

Reference: raytracer/bvh.hsh:106-170,275-357,443-524 (transparency leaf tests
and walks), raytracer/surface.hsh:147-160 (GetOpacity), raytracer/texture.hsh:18-25,
raytracer/direct.hsh:12-86 and lights.hsh:11-31 (GetLight, SampleLight),
RayTracingWorld.cpp:322-407 and RayTracingHelper.cpp:436-500 (light lists).
"""
import math

import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W


def _rays(origins, dirs, tmin=0.0, tmax=1e12):
    r = np.zeros((len(origins), 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = origins, tmin, dirs, tmax
    return r


def _plane_world(zs, opacities, shared=False):
    """Unit quads at the given z (identity transforms); one BLAS per plane, or
    one shared BLAS translated per plane.  Returns the oracle scene and the
    per-BLAS opacity in slot order."""
    if shared:
        v, n = pt_cases.quad_xy(0.0)
        meshes, normals = [v], [n]
        xf = np.zeros((len(zs), 3, 4))
        xf[:, :, :3] = np.eye(3)
        xf[:, 2, 3] = zs
        which = np.zeros(len(zs), np.int64)
    else:
        meshes, normals = zip(*[pt_cases.quad_xy(z) for z in zs])
        meshes, normals = list(meshes), list(normals)
        xf = np.zeros((len(zs), 3, 4))
        xf[:, :, :3] = np.eye(3)
        which = np.arange(len(zs))
    mats = W.c2_materials({"albedo": np.full((len(meshes), 3), 0.5, np.float32),
                           "roughness": np.ones(len(meshes), np.float32)})
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    ops = [np.full(len(r), o, np.float32)[r] for r, o in zip(S["refs"], opacities)]
    return S, ops


def test_oracle_transparency_squares_per_leaf():
    """HitAnyTransparency squares the transparency at every leaf it visits
    (CheckLeafTransparency returns the incoming value times (1 - opacity) of
    its hits, and the caller multiplies that in again, bvh.hsh:496), hit or
    not.  A unit quad's BLAS is a root with two one-triangle leaves, visited
    left (slot 0, the y < x triangle) first.  With n stacked quads of opacity
    a, f = 1 - a, a ray through the second triangle gets T <- T^4 f per quad
    (exponents 1, 5, 21), one through the first T <- T^4 f^2 (2, 10, 42),
    whatever order the TLAS visits the quads in."""
    second = _rays([[0.2, 0.3, 0.0]], [[0.0, 0.0, 1.0]])
    first = _rays([[0.3, 0.2, 0.0]], [[0.0, 0.0, 1.0]])
    for n, e_second, e_first in ((1, 1, 2), (2, 5, 10), (3, 21, 42)):
        S, ops = _plane_world(list(range(1, n + 1)), [0.3], shared=True)
        for ray, e in ((second, e_second), (first, e_first)):
            tr, slot, inst, t = oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops, ray)
            expect = 0.7 ** e if 0.7 ** e >= 1e-6 else 0.0  # below 1e-6 it becomes exactly 0
            assert tr[0] == pytest.approx(expect, rel=1e-5, abs=0.0), (n, e, tr[0])
            assert slot[0] >= 0 and t[0] == 1.0
    S, ops = _plane_world([1.0, 2.0, 3.0], [0.3], shared=True)
    # tmax keeps the walk out of the farther quads' subtrees
    short = _rays([[0.2, 0.3, 0.0]], [[0.0, 0.0, 1.0]], tmax=1.5)
    assert oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops, short)[0][0] == np.float32(0.7)
    # a miss is fully transparent, an opaque quad blocks
    miss = _rays([[5.0, 5.0, 0.0]], [[0.0, 0.0, 1.0]])
    assert oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops, miss)[0][0] == 1.0
    S, ops = _plane_world([1.0, 2.0], [0.3, 1.0])
    assert oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops, second)[0][0] == 0.0


def test_oracle_closest_skips_zero_opacity():
    """HitClosestTransparency drops candidates with opacity <= 0 and keeps any
    positive opacity (translucency is the shader's business)."""
    ray = _rays([[0.3, 0.2, 0.0]], [[0.0, 0.0, 1.0]])
    for ops_in, t_expect in (([0.0, 0.4, 1.0], 2.0), ([0.0, 0.0, 1.0], 3.0), ([1.0, 0.0, 0.0], 1.0),
                             ([0.0, 0.0, 0.0], 1e12)):
        S, ops = _plane_world([1.0, 2.0, 3.0], ops_in)
        tr, slot, inst, t = oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops, ray)
        assert t[0] == np.float32(t_expect), (ops_in, t[0])
        assert (slot[0] >= 0) == (t_expect < 1e12)


def test_oracle_opacity_sampler_closed_forms():
    """Bilinear filtering with repeat addressing at level 0: texel centres
    return the texel, a texture linear in x is reproduced between centres, the
    sampler is periodic, and the half-texel border blends the opposite edge."""
    tex = np.array([[0, 85, 170, 255]] * 3, np.uint8)  # 4 x 3, linear in x
    for i in range(4):
        assert oracle_rt.sample_opacity(tex, (i + 0.5) / 4, 0.5) == pytest.approx(tex[0, i] / 255.0, abs=1e-7)
    for x in np.linspace(0.125, 0.875, 13):
        assert oracle_rt.sample_opacity(tex, x, 0.3) == pytest.approx((4 * x - 0.5) / 3.0, abs=2e-7)
        assert oracle_rt.sample_opacity(tex, x + 3.0, 0.3 - 2.0) == pytest.approx(
            oracle_rt.sample_opacity(tex, x, 0.3), abs=1e-6)
    # x = 0 sits halfway between the last and the first column
    assert oracle_rt.sample_opacity(tex, 0.0, 0.5) == pytest.approx(0.5, abs=1e-7)
    checker = pt_cases.checker(2)
    assert oracle_rt.sample_opacity(checker, 0.5, 0.5) == pytest.approx(0.5, abs=1e-7)
    assert oracle_rt.sample_opacity(checker, 0.25, 0.25) == 1.0 and oracle_rt.sample_opacity(checker, 0.75, 0.25) == 0.0


def _lamp_scene(light_y=10.0, half=0.05, floor=50.0):
    """A floor quad (y = 0, normals up) and a small lamp quad facing down."""
    fv, fn = pt_cases.quad_mesh(floor)
    s = half
    a, b, c, d = [-s, light_y, -s], [s, light_y, -s], [s, light_y, s], [-s, light_y, s]
    lv = np.array([a + b + c, a + c + d], np.float32)
    ln = np.tile(np.array([0, -1, 0] * 3, np.float32), (2, 1))
    xf = np.zeros((2, 3, 4))
    xf[:, :, :3] = np.eye(3)
    mats = W.c2_materials({"albedo": np.ones((2, 3), np.float32), "roughness": np.ones(2, np.float32)})
    return [fv, lv], [fn, ln], mats, xf, np.array([0, 1])


def test_oracle_triangle_light_list():
    """Light list: the sun first, then one light per emissive triangle with
    its centroid, mean normal, Heron area, weight area * brightness of the
    linear colour and radiance linear(linear(colour)) * intensity; weights
    floored at 0.5% of the largest and normalised."""
    meshes, normals, mats, xf, which = _lamp_scene(half=0.5)
    xf[1, :, 3] = (1.0, 2.0, 3.0)  # move the lamp
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    srgb, inten = 0.8, 4.0
    emit = np.array([[0, 0, 0, 0], [srgb, srgb, srgb, inten]], np.float32)
    sun = W.sun_light(radiance=(3.0, 3.0, 3.0))
    surf = {"emitters": emit, "lights_cap": 16}
    cam = atlas_rt.camera((0.0, 5.0, 0.0), 0.0, 90.0, 1.0, 1.0)
    oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                        np.zeros(3, np.float32), cam, 1, 1, 0, 0, surface=surf)
    L = surf["lights"]
    assert len(L) == 3
    lin = np.float32(srgb) ** np.float32(2.2)
    bright_tri = 3 * np.float32(lin) * np.float32(0.3333)
    area = 0.5  # each triangle: half of the 1 x 1 lamp
    w = np.array([3 * 3.0 * 0.3333, area * bright_tri, area * bright_tri])
    w = np.maximum(w, 0.005 * w.max())
    assert np.allclose(L["data"][:, 1], w / w.sum(), rtol=1e-5)
    assert np.isclose(L["data"][:, 1].sum(), 1.0, rtol=1e-6)
    types = L["data"][:, 0].view(np.uint32) >> 28
    assert list(types) == [0, 1, 1]
    assert np.allclose(L["data"][1:, 2], area, rtol=1e-5)
    assert np.allclose(L["color"][1:, :3], inten * lin ** 2.2, rtol=1e-5)
    assert np.allclose(L["N"][1:, :3], [0, -1, 0])
    # centroids of the two triangles of the moved lamp
    ctr = L["P"][1:, :3] - np.array([1.0, 2.0, 3.0])
    assert np.allclose(ctr[:, 1], 10.0) and np.allclose(np.sort(np.abs(ctr[:, 0])), [1 / 6, 1 / 6], atol=1e-6)
    # the instance slot of both triangle lights is the lamp's TLAS slot
    lamp_slot = int(np.nonzero(S["tlas_order"] == 1)[0][0])
    assert np.all(L["data"][1:, 3].view(np.int32) == lamp_slot)


def test_oracle_triangle_light_direct_closed_form():
    """A small lamp 10 units above a white rough floor, seen head-on from
    below, no other light.  Each sample is BRDF * radiance * area_i / d^2
    over the engine's light pdf, which is the selection probability times the
    light's normalised list weight (direct.hsh:47), so with N equal-weight
    lights the pixel converges to N * sum_i BRDF * radiance * area_i / d^2.
    The engine makes one light per BVH slot (RayTracingWorld.cpp:331-333),
    and the root's spatial split puts every triangle of these lamps in two
    slots: a one-triangle lamp gives N = 2 (4x its irradiance), a two-triangle
    quad N = 4 (8x)."""
    brdf = 1.0 / 1.51 / math.pi + 0.04 * (1.0 / (1.0 + 1e-7)) / math.pi * 0.25
    inten = 1000.0
    emit = np.array([[0, 0, 0, 0], [1, 1, 1, inten]], np.float32)
    cam = atlas_rt.camera((0.0, 5.0, 0.0), 0.0, 90.0, 0.5, 1.0)
    for lamp in ("triangle", "quad"):
        meshes, normals, mats, xf, which = _lamp_scene()
        if lamp == "triangle":
            meshes[1] = np.array([[-0.05, 10.0, -0.05, -0.05, 10.0, 0.05, 0.05, 10.0, -0.05]], np.float32)
            normals[1] = normals[1][:1]
        S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
        surf = {"emitters": emit, "lights_cap": 8}
        img = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats,
                                  np.zeros(0, _abi.LIGHT_DTYPE), np.zeros(3, np.float32), cam, 8, 8, 0, 16,
                                  surface=surf)
        L = surf["lights"]
        n_slots = len(S["blases"][1][1])
        assert len(L) == n_slots == (2 if lamp == "triangle" else 4)
        expect = len(L) * sum(brdf * inten * 0.005 / 100.0 for _ in L)
        assert np.allclose(img[..., :3] / 16.0, expect, rtol=2e-3), (lamp, img[0, 0] / 16.0, expect)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _alpha_world(seed=7):
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=4, n_inst=40, salt=seed, extent=20.0, height=6.0)
    g = np.random.default_rng(seed)
    choices = np.array([0.0, 0.3, 0.7, 1.0], np.float32)
    op_caller = [choices[g.integers(0, 4, len(v))] for v in meshes]
    return meshes, normals, mats, xf, which, op_caller


@pytest.mark.gpu
def test_gpu_alpha_traces_match_oracle(torch_cuda):
    """atlas_rt_trace_transparency and atlas_rt_trace_closest_transparency
    against the oracle: transparency bits, slot, instance and t identical."""
    print("case alpha traces", flush=True)
    torch = torch_cuda
    meshes, normals, mats, xf, which, op_caller = _alpha_world()
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    for b, op in zip(blases, op_caller):
        b.set_opacity(op)
    ops_slot = [op[r] for op, r in zip(op_caller, S["refs"])]
    g = np.random.default_rng(1)
    n = 1 << 15
    o = g.uniform(-25, 25, (n, 3)).astype(np.float32)
    o[:, 1] = g.uniform(-5, 12, n)
    d = g.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tmin, tmax = 0.0, 1e12
    tr, slot, inst, t = oracle_rt.trace_world_alpha(S["tlas"], S["inst"], S["blases"], ops_slot,
                                                    _rays(o, d, tmin, tmax), mask=_abi.MASK_ALL)
    packed = np.zeros((n, 12), np.float32)
    packed[:, :3], packed[:, 4:7] = o, d
    packed[:, 3] = np.arange(n, dtype=np.int32).view(np.float32)
    d_in = torch.from_numpy(packed).cuda()
    d_out = torch.zeros_like(d_in)
    d_tr = torch.zeros(n, dtype=torch.float32, device="cuda")
    w_.trace_transparency_packed(d_in, d_tr, None, n, tmin, tmax, mask=_abi.MASK_ALL)
    w_.trace_closest_transparency_packed(d_in, d_out, None, n, tmin, tmax, mask=_abi.MASK_ALL)
    torch.cuda.synchronize()
    g_tr = d_tr.cpu().numpy()
    out = d_out.cpu().numpy()
    g_slot = out[:, 9].view(np.int32)
    g_inst = out[:, 10].view(np.int32)
    print(f"alpha traces: {np.mean(slot >= 0):.3f} hit, {np.mean((tr > 0) & (tr < 1)):.3f} partly transparent")
    assert np.mean(slot >= 0) > 0.2 and np.mean((tr > 0) & (tr < 1)) > 0.02
    assert g_tr.tobytes() == tr.tobytes()
    assert np.array_equal(g_slot, slot)
    assert out[:, 8].tobytes() == t.tobytes()
    hit = slot >= 0
    assert np.array_equal(g_inst[hit], inst[hit])


def _pt_compare(img, ref):
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    return close.mean(), a.mean(), r.mean()


@pytest.mark.gpu
def test_gpu_pt_opacity_matches_oracle_per_pixel(torch_cuda):
    """Cut-outs (opacity 0 triangles), translucent materials (opacity 0.5:
    refraction lobe and partial shadows) and an opacity-mapped material
    (checkerboard texture, -1 baked opacity) per pixel against the oracle."""
    print("case pt opacity", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=4, n_inst=40, salt=11)
    qv, qn = pt_cases.quad_mesh(60.0)
    meshes.append(qv)
    normals.append(qn)
    xf = np.concatenate([xf, np.eye(3, 4)[None] + np.array([[[0, 0, 0, 0], [0, 0, 0, -6.0], [0, 0, 0, 0]]])])
    which = np.concatenate([which, [len(meshes) - 1]])
    mats = np.concatenate([mats, W.c2_materials({"albedo": np.full((1, 3), 0.6, np.float32),
                                                  "roughness": np.ones(1, np.float32)})])
    mats["opacity"][1] = 0.5          # translucent
    mats["opacity_texture"][2] = 0    # checkerboard map
    mats["opacity"][2] = 0.9
    mats["invert_uvs"][2] = 1
    g = np.random.default_rng(3)
    baked, uvs = [], []
    for k, v in enumerate(meshes):
        op = np.full(len(v), mats["opacity"][k], np.float32)
        if k == 0:
            op[g.random(len(v)) < 0.3] = 0.0  # cut-outs
        if k == 2:
            op[:] = -1.0
        baked.append(op)
        uvs.append(pt_cases.sphere_uvs(v))
    tex = [pt_cases.checker(8)]
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    for b, op, uv in zip(blases, baked, uvs):
        b.set_opacity(op, uv)
    W_, H_, bounces, frames = 96, 64, 4, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    sun = W.sun_light()
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=W.C2_ENV)
    pt.set_textures(tex)
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    surf = {"opacity": [op[r] for op, r in zip(baked, S["refs"])], "uvs": [uv[r] for uv, r in zip(uvs, S["refs"])],
            "textures": tex}
    ref = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                              np.array(W.C2_ENV, np.float32), cam, W_, H_, bounces, frames, surface=surf)
    frac, am, rm = _pt_compare(img, ref)
    print("pt opacity parity: fraction of matching pixels", frac, "mean", am, rm)
    assert frac == 1.0 and abs(am - rm) <= 0.01 * rm
    # the cut-outs and the map matter: the same scene all opaque differs
    ref_opaque = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                                     np.array(W.C2_ENV, np.float32), cam, W_, H_, bounces, frames)
    assert _pt_compare(ref_opaque, ref)[0] < 0.95


@pytest.mark.gpu
def test_gpu_pt_triangle_lights_match_oracle(torch_cuda):
    """sampleEmissives: emissive lamp quads become triangle lights; the GPU's
    light list equals the oracle's bit for bit and the images agree per pixel."""
    print("case pt triangle lights", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=3, n_inst=30, salt=5)
    lv, ln = pt_cases.quad_xy(0.0, size=2.0, normal_z=-1.0)
    meshes.append(lv)
    normals.append(ln)
    n_lamp = 4
    lamp_xf = np.zeros((n_lamp, 3, 4))
    for i in range(n_lamp):
        ang = 2 * np.pi * i / n_lamp
        # rotate the quad's -z normal to point down and place it above the blobs
        lamp_xf[i, :, :3] = [[1, 0, 0], [0, 0, -1], [0, 1, 0]]
        lamp_xf[i, :, 3] = (18 * np.cos(ang), 14.0, 18 * np.sin(ang))
    xf = np.concatenate([xf, lamp_xf])
    which = np.concatenate([which, np.full(n_lamp, len(meshes) - 1)])
    mats = np.concatenate([mats, W.c2_materials({"albedo": np.full((1, 3), 0.8, np.float32),
                                                  "roughness": np.ones(1, np.float32)})])
    emit = np.zeros((len(mats), 4), np.float32)
    emit[-1] = (1.0, 0.9, 0.7, 30.0)
    lin = emit[-1, :3] ** 2.2
    mats["emissive"][-1] = lin * emit[-1, 3]
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_, bounces, frames = 96, 64, 3, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    sun = W.sun_light(radiance=(0.5, 0.5, 0.5))
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=(0.02, 0.02, 0.03))
    pt.set_emissive_lights(emit)
    lights = pt.lights()
    surf = {"emitters": emit, "lights_cap": 256}
    ref = oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                              np.array((0.02, 0.02, 0.03), np.float32), cam, W_, H_, bounces, frames, surface=surf)
    assert len(lights) == 1 + len(S["blases"][-1][1]) * n_lamp  # one light per lamp BVH slot and lamp
    assert lights.tobytes() == surf["lights"].tobytes()
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    frac, am, rm = _pt_compare(img, ref)
    print("pt triangle lights parity: fraction of matching pixels", frac, "mean", am, rm)
    assert frac == 1.0 and abs(am - rm) <= 0.01 * rm
    # switching the lamps off darkens the image
    pt.set_emissive_lights(None)
    assert len(pt.lights()) == 1
    pt.reset()
    pt.render(cam, frames)
    dark = pt.read().cpu().numpy()
    assert dark[..., :3].mean() < 0.9 * img[..., :3].mean()
