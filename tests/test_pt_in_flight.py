"""Progressive frames in flight (ATLAS_RT_PT_FRAMES_IN_FLIGHT).

Frames run on the path tracer's own streams, several at a time, and events
order what they share: the PCG seed stream, the accumulation image and the
radiance cache's training.  The image therefore does not depend on how many
frames are in flight: 1 (frames one after another, the engine's order) and
2, 3, 4 must give the same accumulation image bit for bit, and with the cache
the same network weights and hash-table entries after training every frame.
The serial setting is itself pinned against the oracle by test_pt.py and
test_pt_cache.py.
"""
import numpy as np
import pytest

import atlas_rt
import pt_cases
from atlas_rt import workloads as W

ENV = (0.25, 0.3, 0.4)


def _plain_scene(salt=3):
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=salt)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    return w_, blases, mats


def test_frames_in_flight_flag_bits():
    from atlas_rt import _abi
    assert _abi.PT_FRAMES_IN_FLIGHT_SHIFT == 8
    assert (4 << _abi.PT_FRAMES_IN_FLIGHT_SHIFT) & _abi.PT_RAY_BINNING == 0


@pytest.mark.gpu
@pytest.mark.parametrize("binning", [False, True])
def test_gpu_frames_in_flight_same_image(torch_cuda, binning):
    print(f"case pt frames in flight binning={binning}", flush=True)
    w_, blases, mats = _plain_scene()
    W_, H_ = 96, 64
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    imgs = {}
    for n in (1, 2, 3, 4):
        pt = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, binning=binning, frames_in_flight=n)
        pt.render(cam, 3)
        pt.render(cam, 4)   # a second call continues the seed stream and the slot rotation
        imgs[n] = pt.read().cpu().numpy()
        assert pt.sample_count() == 7
        pt.close()
    assert np.all(imgs[1][..., 3] == 7)
    assert imgs[1][..., :3].max() > 0
    for n in (2, 3, 4):
        assert imgs[n].tobytes() == imgs[1].tobytes(), n


@pytest.mark.gpu
def test_gpu_frames_in_flight_reset_continues_seed_stream(torch_cuda):
    print("case pt frames in flight reset", flush=True)
    w_, blases, mats = _plain_scene(salt=5)
    W_, H_ = 64, 48
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    imgs = {}
    for n in (1, 3):
        pt = atlas_rt.PathTracer(w_, W_, H_, 3, mats, W.sun_light(), env=ENV, frames_in_flight=n)
        pt.render(cam, 5)
        pt.reset()
        pt.render(cam, 2)
        imgs[n] = pt.read().cpu().numpy()
        pt.close()
    assert np.all(imgs[1][..., 3] == 2)
    assert imgs[3].tobytes() == imgs[1].tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("query_lag", [0, 1])
def test_gpu_frames_in_flight_same_cache_training(torch_cuda, query_lag):
    """With the cache, frame f's queries must see the network after frame
    f - 1's training (query_lag 0) or frame f - 2's (query_lag 1), and the
    frames train in order.  The table gradients are summed with float
    atomics, and Adam's first steps move every weight by the learning rate
    in the direction of its gradient's sign, so two serial runs already end
    with different weights (a near-zero gradient's sign follows the
    summation order).  What does not depend on the network is exact: the
    last frame's training-sample and query counts and the sample counts;
    the image and the weights of 3 frames in flight must differ from a
    serial run no more than a few times what a second serial run does."""
    print(f"case pt cache frames in flight query_lag={query_lag}", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=0, extent=8.0, height=8.0)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_ = 80, 60
    cam = atlas_rt.camera((0.0, 0.0, -8.8), 0.0, 0.0, 55.0, W_ / H_)
    bounds = ((-12.0, -12.0, -12.0), (12.0, 12.0, 12.0))

    def run(n, lag=query_lag):
        rc = atlas_rt.RadianceCache(bounds[0], bounds[1], seed=5, levels=8, log2_table=14, base_resolution=4,
                                    max_resolution=128, learning_rate=1e-2)
        pt = atlas_rt.PathTracer(w_, W_, H_, 5, mats, W.sun_light(), env=ENV, frames_in_flight=n)
        pt.set_cache(rc, train_stride=4, train_steps=2, query_lag=lag)
        pt.render(cam, 4)
        pt.render(cam, 2)  # a second call restarts the lag from the network as it finds it
        img = pt.read().cpu().numpy().astype(np.float64)
        stats = pt.cache_stats()
        mlp = rc.get_params()[0].cpu().numpy().astype(np.float64)
        pt.close()
        rc.close()
        return img, stats, mlp

    a, b, c = run(1), run(1), run(3)
    assert a[1][0] > 0 and a[1][1] > 0
    assert c[1][:2] == a[1][:2]   # training samples and queries of the last frame
    assert np.all(c[0][..., 3] == 6)
    for k in (0, 2):
        noise = np.abs(b[k] - a[k]).mean()
        diff = np.abs(c[k] - a[k]).mean()
        print(f"{'image' if k == 0 else 'weights'}: serial vs serial {noise:.3e}, 3 in flight vs serial {diff:.3e}")
        assert diff <= 4.0 * noise + 1e-9 * np.abs(a[k]).mean()


@pytest.mark.gpu
def test_gpu_frames_in_flight_alpha_and_triangle_lights(torch_cuda):
    """The opacity-aware trace and shadow kernels and triangle lights: the
    same image bit for bit for 1 and 3 frames in flight."""
    print("case pt frames in flight alpha + triangle lights", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2(n_meshes=4, n_inst=40, salt=7, extent=20.0, height=6.0)
    g = np.random.default_rng(7)
    choices = np.array([0.0, 0.3, 0.7, 1.0], np.float32)
    ops = [choices[g.integers(0, 4, len(v))] for v in meshes]
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    for b, op in zip(blases, ops):
        b.set_opacity(op)
    emit = np.zeros((len(mats), 4), np.float32)
    emit[0] = (1.0, 0.9, 0.7, 5.0)   # every triangle of mesh 0's instances emits
    W_, H_ = 96, 64
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    imgs = {}
    for n in (1, 3):
        pt = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(radiance=(0.5, 0.5, 0.5)), env=(0.02, 0.02, 0.03),
                                 frames_in_flight=n)
        pt.set_emissive_lights(emit)
        assert len(pt.lights()) > 1
        pt.render(cam, 5)
        imgs[n] = pt.read().cpu().numpy()
        pt.close()
    assert imgs[1][..., :3].max() > 0 and np.all(imgs[1][..., 3] == 5)
    assert imgs[3].tobytes() == imgs[1].tobytes()


@pytest.mark.gpu
def test_gpu_frames_in_flight_calls_on_different_streams(torch_cuda):
    """Render calls queued on two different streams without any
    synchronisation between them: the second call's first frames still
    follow the first call's frames (seeds, accumulation), so the image equals
    one serial call on the default stream."""
    print("case pt frames in flight two streams", flush=True)
    torch = torch_cuda
    w_, blases, mats = _plain_scene(salt=11)
    W_, H_ = 96, 64
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    ref = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, frames_in_flight=1)
    ref.render(cam, 7)
    img_ref = ref.read().cpu().numpy()
    ref.close()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    pt = atlas_rt.PathTracer(w_, W_, H_, 4, mats, W.sun_light(), env=ENV, frames_in_flight=3)
    pt.render(cam, 3, stream=s1)
    pt.render(cam, 4, stream=s2)
    img = pt.read(stream=s2).cpu().numpy()
    pt.close()
    assert img.tobytes() == img_ref.tobytes()
