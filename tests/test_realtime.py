"""Realtime path tracing: G-buffer, per-frame radiance and the temporal resolve.

CPU: the oracle's binary16 conversion against numpy's; the temporal restatement
pinned in closed form: a constant image with a matching history keeps its
value and grows its history length; with no valid history every pixel becomes
its 11 x 11 depth-weighted neighbourhood mean (one bright pixel spreads as
(120 c + b) / 121); a history reprojected one pixel to the right blends in
the right neighbour with the clip-limited weight, and pixels whose history
falls off the screen restart.
GPU: the G-buffer equals the oracle's bit for bit (depth, velocity, normal,
material, after a camera move and an instance refit); the frame radiance
agrees per pixel with the oracle's realtime path trace (1 and 4 samples per
pixel, the quantised sum); the resolve equals the oracle's resolve of the same
inputs to 1e-5.

Reference: PathTracingRenderer.cpp:29-292, rayGen.csh:35-45, rayHit.csh:76-149,
pathtracer/temporal.csh, common/normalencode.hsh, common/ycocg.hsh.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

ENV = (0.25, 0.3, 0.4)


def test_oracle_binary16_is_numpys():
    g = np.random.default_rng(1)
    v = np.concatenate([g.normal(0, 1, 2000), g.normal(0, 1e-5, 500), g.normal(0, 3e4, 200),
                        [0.0, -0.0, 1.0 / 64, 65504.0, 65520.0, 1e6, -1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                         np.inf, -np.inf, 5.9604645e-08, 6.1035156e-05, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11]])
    v = v.astype(np.float32)
    with np.errstate(over="ignore"):
        ref = v.astype(np.float16).view(np.uint16)
    assert np.array_equal(oracle_rt.f2h(v), ref)
    bits = np.arange(0, 65536, 7, dtype=np.uint16)
    finite = (bits & 0x7C00) != 0x7C00
    assert np.array_equal(oracle_rt.h2f(bits[finite]), bits[finite].view(np.float16).astype(np.float32))


def _pack(a, b):
    return (oracle_rt.f2h(a).astype(np.uint32) | (oracle_rt.f2h(b).astype(np.uint32) << 16)).reshape(np.shape(a))


def _buffers(h, w, c, depth=5.0, material=7, hist_len=0.0, hist_material=None, velocity=(0.0, 0.0)):
    frame = np.zeros((h, w, 4), np.float32)
    frame[..., :3] = c
    d = np.full((h, w), depth, np.float32)
    vel = _pack(np.full((h, w), velocity[0], np.float32), np.full((h, w), velocity[1], np.float32))
    nrm = _pack(np.zeros((h, w), np.float32), np.zeros((h, w), np.float32))  # +y
    mat = np.full((h, w), material, np.uint16)
    hrad = np.zeros((h, w, 4), np.float32)
    hrad[..., :3] = c
    hrad[..., 3] = hist_len
    hmat = mat if hist_material is None else np.full((h, w), hist_material, np.uint16)
    return dict(frame=frame, depth=d, velocity=vel, normal=nrm, material=mat, hdepth=d, hnormal=nrm, hmaterial=hmat,
                hradiance=hrad)


def test_oracle_temporal_constant_image_keeps_value_and_grows_history():
    c = np.array([0.5, 0.25, 1.0], np.float32)  # dyadic: YCoCg round trips exactly
    B = _buffers(20, 24, c, hist_len=40.0)
    out = oracle_rt.temporal(**B)
    assert np.array_equal(out[..., :3], np.broadcast_to(c, (20, 24, 3)))
    assert np.all(out[..., 3] == 41.0)


def test_oracle_temporal_without_history_is_the_neighbourhood_mean():
    h, w = 30, 32
    c = np.array([0.5, 0.25, 1.0], np.float32)
    B = _buffers(h, w, c, hist_material=3)  # no pixel's history matches
    bright = np.array([8.0, 4.0, 2.0], np.float32)
    B["frame"][15, 16, :3] = bright
    out = oracle_rt.temporal(**B)
    assert np.all(out[..., 3] == 1.0)
    ref = np.zeros((h, w, 3))
    yy, xx = np.mgrid[0:h, 0:w]
    for y in range(h):
        for x in range(w):
            ys = np.clip(np.arange(y - 5, y + 6), 0, h - 1)
            xs = np.clip(np.arange(x - 5, x + 6), 0, w - 1)
            ref[y, x] = B["frame"][ys][:, xs, :3].astype(np.float64).mean(axis=(0, 1))
    assert np.allclose(out[..., :3], ref, rtol=2e-6, atol=1e-7)
    assert np.allclose(out[15, 16, :3], (120 * c.astype(np.float64) + bright) / 121, rtol=2e-6)
    assert np.allclose(out[0, 0, :3], c, rtol=2e-6)


def test_oracle_temporal_reprojects_one_pixel_right():
    """velocity = +1 pixel in x: the history of pixel x comes from x + 1.  With
    a constant current image (zero variance) the history is clipped entirely
    (clip = 1, limited to historyClipMax = 0.2), so the weight is
    (31 / 32) * 0.8; the last column's history lies off the screen."""
    h, w = 16, 64
    c = np.array([0.5, 0.5, 0.5], np.float32)
    B = _buffers(h, w, c, hist_len=40.0, velocity=(1.0 / w, 0.0))
    ramp = (np.arange(w, dtype=np.float32) / w + 1.0)[None, :, None]
    B["hradiance"][..., :3] = np.broadcast_to(ramp, (h, w, 3))
    out = oracle_rt.temporal(**B)
    f = np.float32(31.0 / 32.0) * np.float32(0.8)
    expect = c[0] * (1 - f) + ramp[0, 1:, 0] * f
    assert np.allclose(out[:, :-1, 0], np.broadcast_to(expect, (h, w - 1)), rtol=1e-6)
    assert np.all(out[:, :-1, 3] == 41.0)
    assert np.allclose(out[:, -1, :3], 0.5, rtol=1e-6) and np.all(out[:, -1, 3] == 1.0)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _scene():
    meshes, normals, mats, xf, which = pt_cases.small_c2()
    mats = mats.copy()
    mats["id"] = np.arange(len(mats), dtype=np.int32) * 1009 + 17
    descs = atlas_rt.instance_descs(xf, which)
    descs["material_offset"] = which.astype(np.int32)
    blases = [atlas_rt.Blas.build(v) for v in meshes]
    for b, nrm in zip(blases, normals):
        b.set_shading(nrm)
    w = atlas_rt.World()
    w.update(descs, blases)
    obs = [oracle_rt.build_triangles(v) for v in meshes]
    shading = [(n[ob.ref_idx], np.zeros(len(ob.ref_idx), np.int32)) for n, ob in zip(normals, obs)]
    return dict(meshes=meshes, normals=normals, mats=mats, xf=xf, which=which, descs=descs, blases=blases, world=w,
                obs=obs, shading=shading)


def _xform(S):
    """The world's object-to-world rows in instance-slot order."""
    _, _, order = S["world"].export()
    return np.stack([S["cur_descs"][r]["object_to_world"].reshape(-1) for r in order]).astype(np.float32)


def _oracle_frame(S, cam, fm, W_, H_, bounces, spf, last_xform, frame_index):
    tl, inst, _ = S["world"].export()
    return oracle_rt.pathtrace_realtime(tl, inst, _xform(S), [(ob.nodes, ob.tris) for ob in S["obs"]], S["shading"],
                                        S["mats"], W.sun_light(), np.array(ENV, np.float32), cam, W_, H_, bounces,
                                        spf, fm, last_xform, pcg_draws=frame_index * (bounces + 1))


@pytest.mark.gpu
@pytest.mark.parametrize("spf", [1, 4])
def test_gpu_realtime_matches_oracle(torch_cuda, spf):
    print(f"case realtime spf={spf}", flush=True)
    torch = torch_cuda
    S = _scene()
    W_, H_, bounces = 96, 64, 3
    pt = atlas_rt.PathTracer(S["world"], W_, H_, bounces, S["mats"], W.sun_light(), env=ENV)
    pt.set_realtime(samples_per_frame=spf)
    proj = atlas_rt.projection_matrix(50.0, W_ / H_)
    poses = [((0.0, 15.0, -70.0), 0.0, 12.0), ((1.5, 15.5, -69.0), 2.0, 11.0), ((3.0, 16.0, -68.0), 4.0, 10.0)]
    g = W.rng(2, 77)
    S["cur_descs"] = S["descs"]
    last_xform = None
    last_vp = None
    hist = None
    for f, (eye, yaw, pitch) in enumerate(poses):
        if f == 2:  # move the instances too: velocity from the last matrices
            xf2 = S["xf"].copy()
            xf2[:, :, 3] += g.uniform(-0.5, 0.5, (len(xf2), 3))
            moved = atlas_rt.instance_descs(xf2, S["which"])
            moved["material_offset"] = S["which"].astype(np.int32)
            S["keep"] = S["world"].refit(moved)
            S["cur_descs"] = moved
        view = atlas_rt.view_matrix(eye, yaw, pitch)
        vp = proj @ view
        fm = atlas_rt.frame_matrices(view, proj, vp if last_vp is None else last_vp,
                                     jitter=(0.3 * (f % 2) - 0.1, 0.2 - 0.25 * (f % 2)))
        cam = atlas_rt.camera(eye, yaw, pitch, 50.0, W_ / H_)
        if last_xform is None:
            last_xform = _xform(S)
        pt.render_realtime(cam, fm)
        got = {k: pt.read_realtime(k).cpu().numpy() for k in range(6)}
        frame, depth, vel, nrm, mat = _oracle_frame(S, cam, fm, W_, H_, bounces, spf, last_xform, f)
        assert np.array_equal(got[_abi.RT_DEPTH].view(np.uint32), depth.view(np.uint32)), f
        assert np.array_equal(got[_abi.RT_MATERIAL].view(np.uint16), mat), f
        assert np.array_equal(got[_abi.RT_NORMAL].view(np.uint32), nrm), f
        assert np.array_equal(got[_abi.RT_VELOCITY].view(np.uint32), vel), f
        if f == 2:
            v = oracle_rt.h2f((vel & 0xFFFF).astype(np.uint16))
            assert np.abs(v).max() > 0
        a, r = got[_abi.RT_FRAME][..., :3].astype(np.float64), frame[..., :3].astype(np.float64)
        close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
        print(f"frame {f}: matching pixels {close.mean():.4f}")
        assert close.all(), f
        # the resolve of this frame's inputs against the last frame's
        if hist is None:
            hist = dict(hdepth=np.zeros_like(depth), hnormal=np.zeros_like(nrm), hmaterial=np.zeros_like(mat),
                        hradiance=np.zeros((H_, W_, 4), np.float32))
        ref = oracle_rt.temporal(got[_abi.RT_FRAME], got[_abi.RT_DEPTH], got[_abi.RT_VELOCITY].view(np.uint32),
                                 got[_abi.RT_NORMAL].view(np.uint32), got[_abi.RT_MATERIAL].view(np.uint16), **hist)
        res = got[_abi.RT_RESOLVED]
        ok = np.all(np.abs(res - ref) <= 1e-6 + 1e-5 * np.abs(ref), axis=-1)
        print(f"frame {f}: resolve within 1e-5: {ok.mean():.5f}, history lengths {np.unique(res[..., 3])[:6]}")
        assert ok.mean() > 0.999, f
        if f >= 1:
            assert (res[..., 3] > 1).mean() > 0.5  # most pixels found their history
        hist = dict(hdepth=got[_abi.RT_DEPTH], hnormal=got[_abi.RT_NORMAL].view(np.uint32),
                    hmaterial=got[_abi.RT_MATERIAL].view(np.uint16), hradiance=res)
        last_xform = _xform(S)
        last_vp = vp
    pt.close()


@pytest.mark.gpu
@pytest.mark.parametrize("history", [False, True])
def test_gpu_realtime_after_world_update(torch_cuda, history):
    """An update between realtime frames re-packs the instances.  Without an
    object history the update drops the last 8 instances and moves the others:
    each kept instance's velocity comes from its own previous-frame matrix,
    matched by its index in the update's array.  With one
    (atlas_rt_world_set_history, the engine's includeObjectHistory) the update
    also appends 4 new instances whose last matrices are their current ones,
    and every instance's velocity comes from the matrix the caller gave.  The
    G-buffer equals the oracle's bit for bit."""
    print(f"case realtime after update history={history}", flush=True)
    S = _scene()
    W_, H_, bounces = 96, 64, 2
    pt = atlas_rt.PathTracer(S["world"], W_, H_, bounces, S["mats"], W.sun_light(), env=ENV)
    pt.set_realtime(samples_per_frame=1)
    proj = atlas_rt.projection_matrix(50.0, W_ / H_)
    eye, yaw, pitch = (0.0, 15.0, -70.0), 0.0, 12.0
    view = atlas_rt.view_matrix(eye, yaw, pitch)
    vp = proj @ view
    cam = atlas_rt.camera(eye, yaw, pitch, 50.0, W_ / H_)
    pt.render_realtime(cam, atlas_rt.frame_matrices(view, proj, vp))
    g = W.rng(2, 906)
    n_old = len(S["descs"])
    keep = n_old - 8
    added = 4 if history else 0
    xf2 = np.concatenate([S["xf"][:keep], W.random_affine(g, 4, scale=(1.0, 4.0), extent=30.0, height=8.0)[:added]])
    xf2[:keep, :, 3] += g.uniform(-0.5, 0.5, (keep, 3))
    which2 = np.concatenate([S["which"][:keep], g.integers(0, len(S["meshes"]), 4)[:added]])
    descs2 = atlas_rt.instance_descs(xf2, which2)
    descs2["material_offset"] = which2.astype(np.int32)
    S["world"].update(descs2, S["blases"])
    last_caller = np.stack([(S["descs"][r] if r < keep else descs2[r])["object_to_world"].reshape(12)
                            for r in range(len(descs2))]).astype(np.float32)
    if history:
        S["world"].set_history(last_caller)
    S["cur_descs"] = descs2
    _, _, order = S["world"].export()
    last_xform = last_caller[order]
    pt.render_realtime(cam, atlas_rt.frame_matrices(view, proj, vp))
    got = {k: pt.read_realtime(k).cpu().numpy() for k in range(6)}
    frame, depth, vel, nrm, mat = _oracle_frame(S, cam, atlas_rt.frame_matrices(view, proj, vp), W_, H_, bounces, 1,
                                                last_xform, 1)
    assert np.array_equal(got[_abi.RT_DEPTH].view(np.uint32), depth.view(np.uint32))
    assert np.array_equal(got[_abi.RT_MATERIAL].view(np.uint16), mat)
    assert np.array_equal(got[_abi.RT_NORMAL].view(np.uint32), nrm)
    assert np.array_equal(got[_abi.RT_VELOCITY].view(np.uint32), vel)
    v = oracle_rt.h2f((vel & 0xFFFF).astype(np.uint16))
    assert np.abs(v).max() > 0  # the kept instances moved
    a, r = got[_abi.RT_FRAME][..., :3].astype(np.float64), frame[..., :3].astype(np.float64)
    close = np.all(np.abs(a - r) <= 1e-4 + 1e-3 * np.abs(r), axis=-1)
    assert close.all()
    pt.close()
