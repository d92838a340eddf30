"""Path tracing with the radiance cache.

Parity: with a frozen cache, one frame of the GPU's cache mode against the
oracle's restatement of the same bounce loop (render paths stop at their
second vertex; about one pixel in train_stride traces the full path) with
the cache evaluated by the FP64 cache oracle; the GPU's training samples
against the suffix sums of the oracle's per-vertex records.

Gain: at equal samples per pixel the trained cache beats both the plain
path tracer and the same renderer with an untrained cache, measured by
rel-MSE against a long plain render.
"""
import numpy as np
import pytest

import atlas_rt
import nrc_oracle as O
import oracle_rt
import pt_cases
from atlas_rt import workloads as W

ENV = (0.25, 0.3, 0.4)
BOUNDS = ((-12.0, -12.0, -12.0), (12.0, 12.0, 12.0))
CACHE = dict(levels=8, log2_table=14, base_resolution=4, max_resolution=128)


def _scene(salt=0):
    """A dense cluster (48 blobs over a 16-unit cube, seen from its edge) so
    that most second bounces land on another blob and the cache term
    matters."""
    meshes, normals, mats, xf, which = pt_cases.small_c2(salt=salt, extent=8.0, height=8.0)
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    return S, w_, blases, mats


def _cam(w, h):
    return atlas_rt.camera((0.0, 0.0, -8.8), 0.0, 0.0, 55.0, w / h)


def rel_mse(img, ref):
    x = img[..., :3] / np.maximum(img[..., 3:4], 1)
    r = ref[..., :3] / np.maximum(ref[..., 3:4], 1)
    return float(np.mean((x - r) ** 2 / (r * r + 0.01)))


def _targets_from_records(train, rec):
    """Training samples the oracle's records imply: for each surface vertex
    k, (query, (sum of the radiance added at bounces >= k) / throughput)."""
    out_q, out_t = [], []
    for p in np.argwhere(train):
        r = rec[tuple(p)]
        s = np.zeros(3, np.float32)
        for k in range(r.shape[0] - 1, -1, -1):
            s = (s + r[k, 16:19]).astype(np.float32)
            if r[k, 15] == 0:
                continue
            beta = r[k, 12:15]
            t = np.where(beta > 0, s / np.where(beta > 0, beta, 1), 0).astype(np.float32)
            out_q.append(r[k, :12])
            out_t.append(t)
    return np.array(out_q, np.float32).reshape(-1, 12), np.array(out_t, np.float32).reshape(-1, 3)


@pytest.mark.gpu
def test_gpu_pt_cache_frame_matches_oracle(torch_cuda):
    print("case pt cache parity", flush=True)
    torch = torch_cuda
    S, w_, blases, mats = _scene()
    W_, H_, bounces, stride = 96, 64, 6, 4
    cam = _cam(W_, H_)
    sun = W.sun_light()
    rc = atlas_rt.RadianceCache(BOUNDS[0], BOUNDS[1], seed=5, **CACHE)
    P = O.Params(BOUNDS[0], BOUNDS[1], CACHE["levels"], CACHE["log2_table"], CACHE["base_resolution"],
                 CACHE["max_resolution"])
    g = np.random.default_rng(21)
    m = np.zeros(O.MASTER_SIZE, np.float32)
    for l in range(6):
        lim = np.sqrt(6.0 / (O.K[l] + O.N[l]))
        m[O.MASTER_OFFSETS[l]:O.MASTER_OFFSETS[l + 1]] = g.uniform(-lim, lim, O.K[l] * O.N[l])
    t16 = g.uniform(-1, 1, (P.levels, P.T, 2)).astype(np.float16)
    rc.set_params(torch.from_numpy(m).cuda(), torch.from_numpy(t16.view(np.int16).reshape(-1).copy()).cuda())

    pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
    pt.set_cache(rc, train_stride=stride, train_steps=0)
    pt.render(cam, 1)
    img = pt.read().cpu().numpy()
    n_samples, n_queries, _ = pt.cache_stats()
    gq, gt = pt.cache_samples(W_ * H_ * (bounces + 1))
    gq, gt = gq.cpu().numpy(), gt.cpu().numpy()

    rad, query, train, rec = oracle_rt.pathtrace_cache(S["tlas"], S["inst"], S["xform"], S["blases"], S["shading"],
                                                       S["materials"], sun, np.array(ENV, np.float32), cam, W_, H_,
                                                       bounces, stride)
    has_q = query[..., 15] > 0
    pred = np.zeros((H_, W_, 3))
    pred[has_q] = O.predict(P, m, t16.astype(np.float64), query[has_q][:, :12], quantize=True)
    cache_term = query[..., 12:15] * np.maximum(pred, 0)
    expect = rad + cache_term
    a = img[..., :3].astype(np.float64)
    tol = 1e-4 + 1e-3 * np.abs(expect) + 3e-3 * np.abs(cache_term)
    close = np.all(np.abs(a - expect) <= tol, axis=-1)
    print(f"pt cache parity: {close.mean():.4f} of pixels match; queries gpu {n_queries} oracle {has_q.sum()}; "
          f"training pixels {train.sum()}, cache term mean {cache_term.mean():.4f}, image mean {a.mean():.4f}")
    assert np.all(img[..., 3] == 1)
    assert close.all()
    assert abs(int(n_queries) - int(has_q.sum())) <= 0.02 * has_q.sum()
    assert cache_term.mean() > 0.01 * expect.mean()  # the cache term is really exercised

    oq, ot = _targets_from_records(train, rec)
    print(f"pt cache samples: gpu {len(gq)} (stats {n_samples}), oracle {len(oq)}")
    assert len(gq) == n_samples and abs(len(gq) - len(oq)) <= 0.03 * len(oq) and len(oq) > 200
    # Training samples are written in a different order on the two sides:
    # pair each GPU sample with the nearest oracle sample in (position,
    # direction, normal); every one must have a partner and match its target.
    from scipy.spatial import cKDTree
    key = lambda q: np.concatenate([q[:, 0:3], 10.0 * q[:, 4:7], 10.0 * q[:, 8:11]], axis=1)
    dist, j = cKDTree(key(oq)).query(key(gq))
    found = dist < 1e-3
    exact = np.all(gq == oq[j], axis=1)
    tmatch = np.all(np.abs(gt[:, :3] - ot[j]) <= 1e-5 + 1e-3 * np.abs(ot[j]), axis=1) & found
    print(f"pt cache samples: {found.mean():.4f} of GPU samples have an oracle vertex within 1e-3 "
          f"({exact.mean():.4f} bit-exact), {tmatch.sum() / max(found.sum(), 1):.4f} of those match the target")
    assert found.all() and tmatch.sum() == found.sum()
    assert np.all(gt[:, 3] == 0) and np.all(np.isfinite(gt))


@pytest.mark.gpu
def test_gpu_pt_cache_gain_at_equal_samples(torch_cuda):
    """The trained cache lowers rel-MSE against a long plain render at equal
    samples per pixel, and an untrained cache does not: what the network
    adds over the plain tracer and over the same renderer without learning."""
    print("case pt cache gain", flush=True)
    torch = torch_cuda
    S, w_, blases, mats = _scene(salt=1)
    W_, H_, bounces, spp = 96, 64, 6, 16
    cam = _cam(W_, H_)
    sun = W.sun_light()

    ref_pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
    ref_pt.render(cam, 2048)
    ref = ref_pt.read().cpu().numpy()

    plain = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
    plain.render(cam, spp)
    e_plain = rel_mse(plain.read().cpu().numpy(), ref)

    def cached(train_steps, warmup):
        rc = atlas_rt.RadianceCache(BOUNDS[0], BOUNDS[1], seed=9, learning_rate=1e-2, **CACHE)
        pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, sun, env=ENV)
        pt.set_cache(rc, train_stride=4, train_steps=train_steps)
        if warmup:
            pt.render(cam, warmup)
            pt.reset()
        pt.render(cam, spp)
        stats = pt.cache_stats()
        return rel_mse(pt.read().cpu().numpy(), ref), stats

    e_trained, st = cached(4, 128)
    e_untrained, _ = cached(0, 0)
    print(f"pt cache gain at {spp} spp: rel-MSE plain {e_plain:.5f}, trained cache {e_trained:.5f} "
          f"(last loss {st[2]:.4f}, {st[0]} samples/frame), untrained cache {e_untrained:.5f}")
    assert e_trained < 0.8 * e_plain
    assert e_trained < 0.8 * e_untrained
