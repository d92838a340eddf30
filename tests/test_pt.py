"""Path tracer: closed-form pins of the oracle's restatement (CPU) and GPU
parity with it (per pixel, same RNG streams).

Reference: PathTracingRenderer::Render (src/engine/renderer/PathTracingRenderer.cpp:29-292),
rayGen.csh, rayHit.csh, brdf/*.hsh.
"""
import math

import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

ENV = (0.25, 0.3, 0.4)


def _cam(eye, yaw, pitch, w, h, fov=50.0):
    return atlas_rt.camera(eye, yaw, pitch, fov, w / h)


def _oracle_render(S, cam, w, h, bounces, frames, lights, env=ENV, first_sample=0, pcg_draws=0):
    return oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], S["materials"], lights,
                               np.array(env, np.float32), cam, w, h, bounces, frames, first_sample, pcg_draws)


def test_oracle_pt_env_only_is_exact():
    """Rays that miss everything return min(env * 1, 10) exactly."""
    v, n = pt_cases.quad_mesh()
    mats = W.c2_materials({"albedo": np.full((1, 3), 0.5, np.float32), "roughness": np.ones(1, np.float32)})
    xf = np.zeros((1, 3, 4))
    xf[0, :, :3] = np.eye(3)
    S = pt_cases.oracle_scene([v], [n], mats, xf, np.zeros(1, np.int64))
    cam = _cam((0.0, 10.0, 0.0), 0.0, -60.0, 16, 12)  # looking up, the quad is below
    img = _oracle_render(S, cam, 16, 12, 4, 3, W.sun_light())
    assert np.all(img[..., 3] == 3.0)
    for k in range(3):
        assert np.all(img[..., k] == np.float32(3 * ENV[k]) + 0.0) or np.allclose(img[..., k], 3 * ENV[k], rtol=1e-6)


def test_oracle_pt_direct_light_head_on_closed_form():
    """Sun straight down onto a white rough quad seen straight from above,
    no bounces: N = V = L = H, so the renormalized Disney diffuse is
    1/1.51 / pi and the GGX term (alpha = 1, F0 = 0.16 * 0.5^2 = 0.04) is
    0.04 * (1/pi) * 0.25; L = sun * (diffuse + specular)."""
    v, n = pt_cases.quad_mesh()
    mats = W.c2_materials({"albedo": np.ones((1, 3), np.float32), "roughness": np.ones(1, np.float32)})
    xf = np.zeros((1, 3, 4))
    xf[0, :, :3] = np.eye(3)
    S = pt_cases.oracle_scene([v], [n], mats, xf, np.zeros(1, np.int64))
    sun = W.sun_light(direction=(0.0, -1.0, 0.0), radiance=(2.0, 2.0, 2.0))
    # a narrow camera looking straight down: V is within ~0.01 rad of N
    cam = atlas_rt.camera((0.0, 10.0, 0.0), 0.0, 90.0, 0.5, 1.0)
    img = _oracle_render(S, cam, 8, 8, 0, 1, sun)
    expect = 2.0 * (1.0 / 1.51 / math.pi + 0.04 * (1.0 / (1.0 + 1e-7)) / math.pi * 0.25)
    assert np.allclose(img[..., :3], expect, rtol=2e-3), (img[0, 0], expect)


@pytest.mark.gpu
def test_gpu_pt_matches_oracle_per_pixel(torch_cuda):
    """Same scene, camera, seeds and RNG streams: the GPU's wavefront path
    tracer and the oracle's per-pixel restatement agree pixel by pixel.  Both
    sides compute sin, cos and the Schlick power with the same restatement,
    so no sampling decision can flip: every pixel agrees to float rounding
    (sums of the same terms), with no allowance for outliers."""
    print("case pt parity", flush=True)
    torch = torch_cuda
    meshes, normals, mats, xf, which = pt_cases.small_c2()
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    tl, inst, order = w_.export()
    assert tl.tobytes() == S["tlas"].tobytes() and inst.tobytes() == S["inst"].tobytes()
    for b, (bn, bt) in zip(blases, S["blases"]):
        nn, tt, _ = b.export()
        assert nn.tobytes() == bn.tobytes() and tt.tobytes() == bt.tobytes()
    W_, H_, bounces, frames = 96, 64, 4, 3
    cam = _cam((0.0, 15.0, -70.0), 0.0, 12.0, W_, H_)
    sun = W.sun_light()
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
    pt.render(cam, frames)
    img = pt.read().cpu().numpy()
    ref = _oracle_render(S, cam, W_, H_, bounces, frames, sun)
    assert np.all(img[..., 3] == frames) and np.all(ref[..., 3] == frames)
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    print("pt parity: fraction of matching pixels", close.mean(), "mean", a.mean(), r.mean())
    assert close.all()
    assert abs(a.mean() - r.mean()) <= 0.01 * r.mean()
    # the scene is actually lit: some sun, some env, some dark
    assert (r > 0).any() and r.max() > 0.2


@pytest.mark.gpu
def test_gpu_pt_progressive_and_reset(torch_cuda):
    torch = torch_cuda
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=3)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    cam = _cam((0.0, 15.0, -70.0), 0.0, 12.0, 48, 32)
    pt = atlas_rt.PathTracer(w_, 48, 32, 3, mats, W.sun_light(), env=ENV)
    pt.render(cam, 2)
    pt.render(cam, 3)
    img = pt.read().cpu().numpy()
    assert pt.sample_count() == 5 and np.all(img[..., 3] == 5)
    # the 5 frames continue the same seed stream as one call of 5 frames
    ref = _oracle_render(S, cam, 48, 32, 3, 5, W.sun_light())
    close = np.all(np.abs(img[..., :3] - ref[..., :3]) <= 1e-4 + 1e-3 * np.abs(ref[..., :3]), axis=-1)
    assert close.all()
    pt.reset()
    img2 = pt.read().cpu().numpy()
    assert np.all(img2 == 0) and pt.sample_count() == 0
    pt.render(cam, 1)
    # after a reset the sample index restarts at 0 but the PCG stream (the
    # engine's global Random) continues: 5 frames x 4 seeds were drawn
    ref1 = _oracle_render(S, cam, 48, 32, 3, 1, W.sun_light(), first_sample=0, pcg_draws=5 * 4)
    img3 = pt.read().cpu().numpy()
    close = np.all(np.abs(img3[..., :3] - ref1[..., :3]) <= 1e-4 + 1e-3 * np.abs(ref1[..., :3]), axis=-1)
    assert close.all() and np.all(img3[..., 3] == 1)


@pytest.mark.gpu
def test_gpu_pt_renders_the_world_after_an_update(torch_cuda):
    """RayTracingWorld::Update then PathTracingRenderer::Render: an update
    replaces the world's device arrays (here with fewer BLASes and fewer
    instances), and the next render traces the new world, per pixel against
    the oracle's render of it."""
    print("case pt after world update", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=5)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_, bounces = 64, 48, 3
    cam = _cam((0.0, 15.0, -70.0), 0.0, 12.0, W_, H_)
    sun = W.sun_light()
    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
    pt.render(cam, 2)
    before = pt.read().cpu().numpy()
    g = W.rng(2, 905)
    xf2 = W.random_affine(g, 30, scale=(1.0, 4.0), extent=30.0, height=8.0)
    which2 = g.integers(0, 3, 30)
    S2 = pt_cases.oracle_scene(meshes[:3], normals[:3], mats, xf2, which2)
    descs2 = atlas_rt.instance_descs(xf2, which2)
    descs2["material_offset"] = which2.astype(np.int32)
    w_.update(descs2, blases[:3], world_aabbs=S2["boxes"])
    tl, inst, _ = w_.export()
    assert tl.tobytes() == S2["tlas"].tobytes() and inst.tobytes() == S2["inst"].tobytes()
    pt.reset()
    pt.render(cam, 2)
    img = pt.read().cpu().numpy()
    ref = _oracle_render(S2, cam, W_, H_, bounces, 2, sun, pcg_draws=2 * (bounces + 1))
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    print("pt after update: fraction of matching pixels", close.mean())
    assert close.all() and abs(a.mean() - r.mean()) <= 0.01 * r.mean()
    assert not np.allclose(before[..., :3], img[..., :3])  # the scene did change
    pt.close()
