"""The resident frame kernel against the bounce-by-bounce launches.

By default progressive frames run as resident kernels that walk every path
from its camera ray to its last bounce in one lane (trace, shade, shadow ray,
next bounce), several frames per launch.  ATLAS_RT_PT_WAVEFRONT runs the engine's dispatch shape instead: one
closest-hit launch, one shade launch and one shadow-ray launch per bounce over
compacted path queues.  Every path computes the same arithmetic in the same
order either way and writes only its own pixel, so the accumulated image, the
realtime frame, its G-buffer and the resolve must be equal bit for bit: over
an opaque world (the closest-hit and shadow walks share one step body), over
cut-outs, translucency and an opacity map (the transparency walks), with
triangle lights, and in realtime mode with four samples per pixel.  The
bounce-by-bounce image is itself pinned per pixel against the oracle by
test_pt.py, test_pt_alpha_lights.py and test_realtime.py.
"""
import numpy as np
import pytest

import atlas_rt
import pt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

pytestmark = pytest.mark.gpu

ENV = (0.25, 0.3, 0.4)


def _both(make, render, read):
    out = []
    for wavefront in (False, True):
        pt = make(wavefront)
        render(pt)
        out.append(read(pt))
        pt.close()
    return out


def _plain(salt=3, n_inst=48):
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_inst=n_inst, salt=salt)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    return w_, blases, mats


@pytest.mark.parametrize("fif", [1, 3])
def test_resident_equals_wavefront_opaque(torch_cuda, fif):
    print(f"case resident opaque fif={fif}", flush=True)
    w_, blases, mats = _plain()
    W_, H_ = 160, 96
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)

    def render(pt):
        pt.render(cam, 3)
        pt.render(cam, 2)

    a, b = _both(lambda wf: atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, frames_in_flight=fif,
                                                wavefront=wf),
                 render, lambda pt: pt.read().cpu().numpy())
    assert np.all(a[..., 3] == 5) and a[..., :3].max() > 0
    diff = np.nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1))
    assert len(diff[0]) == 0, f"{len(diff[0])} pixels differ, first {list(zip(*diff))[:5]}"


def test_resident_batches_equal_wavefront(torch_cuda):
    """19 frames in one call run as resident launches of 8, 8 and 3 frames
    (each frame its own seeds and radiance buffer, summed per pixel in frame
    order): the image equals 19 bounce-by-bounce frames bit for bit."""
    print("case resident batches", flush=True)
    w_, blases, mats = _plain(salt=6)
    W_, H_ = 64, 48
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    a, b = _both(lambda wf: atlas_rt.PathTracer(w_, W_, H_, 3, mats, W.sun_light(), env=ENV, wavefront=wf),
                 lambda pt: pt.render(cam, 19), lambda pt: pt.read().cpu().numpy())
    assert np.all(a[..., 3] == 19)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("fif", [1, 3])
def test_resident_grid_shape_trials_keep_the_image(torch_cuda, fif):
    """After an untimed first launch, four plain launches alternate four and
    three resident groups per CU (timed, each after the previous one has
    ended); the sixth waits for them and every later launch takes the faster
    shape: nine one-frame calls give the bounce-by-bounce image bit for bit."""
    print(f"case resident grid shape trials fif={fif}", flush=True)
    w_, blases, mats = _plain(salt=4)
    W_, H_ = 96, 64
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)

    def render(pt):
        for _ in range(9):
            pt.render(cam, 1)

    a, b = _both(lambda wf: atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, frames_in_flight=fif,
                                                wavefront=wf),
                 render, lambda pt: pt.read().cpu().numpy())
    assert np.all(a[..., 3] == 9) and a[..., :3].max() > 0
    assert a.tobytes() == b.tobytes()


def test_resident_path_order_with_frames_in_flight(torch_cuda):
    """Batches of frames pull the paths of last launch's hit pixels first
    (a per-range stable partition built before each launch).  Three slots in
    flight write the shared hit map while another slot builds its order;
    the order must stay a permutation, so six 4-frame calls give the
    bounce-by-bounce image bit for bit (ragged size: paths in row order)."""
    print("case resident path order fif=3", flush=True)
    w_, blases, mats = _plain(salt=9)
    W_, H_ = 90, 61
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)

    def render(pt):
        for _ in range(6):
            pt.render(cam, 4)

    a, b = _both(lambda wf: atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, frames_in_flight=3,
                                                wavefront=wf),
                 render, lambda pt: pt.read().cpu().numpy())
    assert np.all(a[..., 3] == 24) and a[..., :3].max() > 0
    assert a.tobytes() == b.tobytes()


def test_resident_equals_wavefront_ragged_and_deep(torch_cuda):
    """A frame size that is not a multiple of the 8x8 tiles (paths in row
    order) and eight bounces."""
    print("case resident ragged", flush=True)
    w_, blases, mats = _plain(salt=7, n_inst=64)
    W_, H_ = 101, 67
    cam = atlas_rt.camera((0.0, 12.0, -60.0), 8.0, 10.0, 60.0, W_ / H_)
    a, b = _both(lambda wf: atlas_rt.PathTracer(w_, W_, H_, 8, mats, W.sun_light(), env=ENV, wavefront=wf),
                 lambda pt: pt.render(cam, 3), lambda pt: pt.read().cpu().numpy())
    assert a.tobytes() == b.tobytes()


def test_resident_equals_wavefront_opacity_and_lamps(torch_cuda):
    """Cut-outs, a translucent material, an opacity map and emissive lamp
    quads (triangle lights): the transparency walks."""
    print("case resident opacity lamps", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=4, n_inst=40, salt=11)
    qv, qn = pt_cases.quad_mesh(60.0)
    lv, ln = pt_cases.quad_xy(0.0, size=2.0, normal_z=-1.0)
    meshes += [qv, lv]
    normals += [qn, ln]
    floor = np.eye(3, 4)[None] + np.array([[[0, 0, 0, 0], [0, 0, 0, -6.0], [0, 0, 0, 0]]])
    lamp = np.zeros((3, 3, 4))
    for i in range(3):
        ang = 2 * np.pi * i / 3
        lamp[i, :, :3] = [[1, 0, 0], [0, 0, -1], [0, 1, 0]]
        lamp[i, :, 3] = (18 * np.cos(ang), 14.0, 18 * np.sin(ang))
    xf = np.concatenate([xf, floor, lamp])
    which = np.concatenate([which, [len(meshes) - 2], np.full(3, len(meshes) - 1)])
    extra = W.c2_materials({"albedo": np.full((2, 3), 0.6, np.float32), "roughness": np.ones(2, np.float32)})
    mats = np.concatenate([mats, extra])
    mats["opacity"][1] = 0.5
    mats["opacity_texture"][2] = 0
    mats["opacity"][2] = 0.9
    emit = np.zeros((len(mats), 4), np.float32)
    emit[-1] = (1.0, 0.9, 0.7, 30.0)
    mats["emissive"][-1] = emit[-1, :3] ** 2.2 * emit[-1, 3]
    g = np.random.default_rng(3)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    for k, (b, v) in enumerate(zip(blases, meshes)):
        op = np.full(len(v), mats["opacity"][k], np.float32)
        if k == 0:
            op[g.random(len(v)) < 0.3] = 0.0
        if k == 2:
            op[:] = -1.0
        b.set_opacity(op, pt_cases.sphere_uvs(v))
    W_, H_ = 96, 64
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)

    def make(wf):
        pt = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(radiance=(0.5, 0.5, 0.5)), env=W.C2_ENV,
                                 wavefront=wf)
        pt.set_textures([pt_cases.checker(8)])
        pt.set_emissive_lights(emit)
        return pt

    a, b = _both(make, lambda pt: pt.render(cam, 3), lambda pt: pt.read().cpu().numpy())
    assert a[..., :3].max() > 0
    assert a.tobytes() == b.tobytes()


def test_resident_equals_wavefront_realtime(torch_cuda):
    print("case resident realtime", flush=True)
    w_, blases, mats = _plain(salt=4)
    W_, H_ = 96, 64
    proj = atlas_rt.projection_matrix(50.0, W_ / H_)
    poses = [((0.0, 15.0, -70.0), 0.0, 12.0), ((1.5, 15.5, -69.0), 2.0, 11.0)]

    def make(wf):
        pt = atlas_rt.PathTracer(w_, W_, H_, 3, mats, W.sun_light(), env=ENV, wavefront=wf)
        pt.set_realtime(samples_per_frame=4)
        return pt

    def render(pt):
        last = None
        for f, (eye, yaw, pitch) in enumerate(poses):
            view = atlas_rt.view_matrix(eye, yaw, pitch)
            vp = proj @ view
            fm = atlas_rt.frame_matrices(view, proj, vp if last is None else last, jitter=(0.3 * f - 0.1, 0.2))
            pt.render_realtime(atlas_rt.camera(eye, yaw, pitch, 50.0, W_ / H_), fm)
            last = vp

    a, b = _both(make, render, lambda pt: [pt.read_realtime(k).cpu().numpy() for k in range(6)])
    for k in range(6):
        assert a[k].tobytes() == b[k].tobytes(), k
    assert a[_abi.RT_FRAME][..., :3].max() > 0


def test_work_counts_equal_the_oracles(torch_cuda):
    """The counting instance of the resident kernel: its closest-hit and
    shadow rays, inner nodes, triangle tests and instance entries over three
    frames equal the oracle's counts of the same paths (its traversal
    restatement counts each fetch as it makes it), and the image equals the
    one the non-counting kernel renders."""
    import oracle_rt
    print("case resident work counts", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_inst=48, salt=3)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_, bounces, frames = 64, 48, 4, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    sun = W.sun_light()
    imgs = []
    for counting in (True, False):
        pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
        if counting:
            pt.count_work()
            assert set(pt.work_counts().values()) == {0}
        pt.render(cam, frames)
        imgs.append(pt.read().cpu().numpy())
        if counting:
            g = pt.work_counts()
            pt.count_work(False)
        pt.close()
    assert imgs[0].tobytes() == imgs[1].tobytes()
    oracle_rt.pathtrace(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"], mats, sun,
                        np.array(ENV, np.float32), cam, W_, H_, bounces, frames)
    o = oracle_rt.pathtrace_counts()
    print("work counts gpu", g, "oracle", o)
    assert g["closest_rays"] >= frames * W_ * H_ and g["hits"] <= g["closest_rays"]
    assert g["shadow_rays"] <= g["hits"]
    assert g["closest_rays"] + g["shadow_rays"] == o["rays"]
    assert (g["nodes"], g["triangles"], g["instances"]) == (o["node_visits"], o["tri_tests"], o["instance_visits"])


def test_work_counts_refuse_uncounted_launches(torch_cuda):
    """Bounce-by-bounce frames do not count: reading the counters after one
    is an error, not a partial sum."""
    w_, blases, mats = _plain(salt=5, n_inst=16)
    W_, H_ = 32, 16
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    pt = atlas_rt.PathTracer(w_, W_, H_, 2, mats, W.sun_light(), env=ENV, wavefront=True)
    pt.count_work()
    pt.render(cam, 1)
    with pytest.raises(atlas_rt.AtlasRTError):
        pt.work_counts()
    pt.close()


def test_graph_replays_equal_eager_frames(torch_cuda):
    """A captured batch of 3 progressive frames, replayed between eager
    render calls, leaves after every step the image (and sample count) that
    the same sequence of eager calls leaves; each replay is read back with
    eager work and synchronisations in between.  A replay after the world's
    stack limit changed is refused."""
    print("case graph replay", flush=True)
    torch = torch_cuda
    w_, blases, mats = _plain(salt=8)
    W_, H_ = 80, 48
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    g = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV)
    e = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV)
    g.render(cam, 1)
    e.render(cam, 1)
    g.capture(cam, 3)
    assert g.sample_count() == 1
    assert g.read().cpu().numpy().tobytes() == e.read().cpu().numpy().tobytes()
    side = torch.cuda.Stream()
    for step, (kind, n) in enumerate([("replay", 3), ("render", 2), ("replay", 3), ("replay", 3), ("render", 5),
                                      ("replay", 3)]):
        if kind == "replay":
            g.replay(stream=side if step == 3 else None)
        else:
            g.render(cam, n)
        e.render(cam, n)
        torch.cuda.synchronize()
        a, b = g.read().cpu().numpy(), e.read().cpu().numpy()
        assert g.sample_count() == e.sample_count(), step
        assert a.tobytes() == b.tobytes(), step
    assert np.all(a[..., 3] == 20) and a[..., :3].max() > 0
    w_.set_stack_limit(8)
    with pytest.raises(atlas_rt.AtlasRTError):
        g.replay()
    w_.set_stack_limit(0)
    g.capture(cam, 2)
    g.replay()
    e.render(cam, 2)
    assert g.read().cpu().numpy().tobytes() == e.read().cpu().numpy().tobytes()
    g.close()
    e.close()


def test_capture_refuses_bounce_by_bounce_frames(torch_cuda):
    w_, blases, mats = _plain(salt=5, n_inst=16)
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, 2.0)
    pt = atlas_rt.PathTracer(w_, 32, 16, 2, mats, W.sun_light(), env=ENV, wavefront=True)
    with pytest.raises(atlas_rt.AtlasRTError):
        pt.capture(cam, 1)
    with pytest.raises(atlas_rt.AtlasRTError):
        pt.replay()
    pt.close()


def test_resident_cache_frames_equal_wavefront(torch_cuda):
    """With the radiance cache and one frame at a time the resident kernel
    ends render paths at their second vertex with a query and records the
    training paths' vertices, as the bounce-by-bounce kernels do.  One frame: the image (direct part plus
    the cache term) is the same bit for bit, and so are the numbers of
    queries and training samples and the multiset of training samples (their
    order in the buffer follows the lanes that wrote them).  Three more
    frames with training in between: every frame trains on samples in
    another order, so the images agree to float rounding."""
    print("case resident cache", flush=True)
    torch = torch_cuda
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=0, extent=8.0, height=8.0)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_, bounces = 96, 64, 6
    cam = atlas_rt.camera((0.0, 0.0, -8.8), 0.0, 0.0, 55.0, W_ / H_)
    out = []
    for wavefront in (False, True):
        rc = atlas_rt.RadianceCache((-12.0,) * 3, (12.0,) * 3, levels=8, log2_table=14, base_resolution=4,
                                    max_resolution=128, seed=5)
        pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, W.sun_light(), env=ENV, wavefront=wavefront,
                                 frames_in_flight=1)
        pt.set_cache(rc, train_stride=4, train_steps=1)
        pt.render(cam, 1)
        img1 = pt.read().cpu().numpy()
        stats = pt.cache_stats()
        q, t = pt.cache_samples(W_ * H_ * (bounces + 1))
        samples = np.concatenate([q.cpu().numpy(), t.cpu().numpy()], axis=1)
        samples = samples[np.lexsort(samples.T[::-1])]
        pt.render(cam, 3)
        img4 = pt.read().cpu().numpy()
        out.append((img1, stats, samples, img4))
        pt.close()
        rc.close()
    (a1, sa, qa, a4), (b1, sb, qb, b4) = out
    print("cache stats resident", sa, "bounce by bounce", sb)
    assert sa[0] == sb[0] and sa[1] == sb[1] and sa[0] > 100 and sa[1] > 100
    assert a1.tobytes() == b1.tobytes()
    assert qa.tobytes() == qb.tobytes()
    assert np.all(a4[..., 3] == 4) and np.all(b4[..., 3] == 4)
    x, y = a4[..., :3].astype(np.float64), b4[..., :3].astype(np.float64)
    assert np.all(np.abs(x - y) <= 1e-3 + 1e-3 * np.abs(y)), np.abs(x - y).max()
