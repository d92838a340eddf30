"""The configurations at their full sizes, through properties that do not
need the oracle to run at that size.

- C2 (256 instances, 1,015,808 instanced triangles, 1280x720, 4 bounces):
  the resident frame kernel and the bounce-by-bounce launches render the
  same two frames bit for bit (the bounce-by-bounce image is pinned per
  pixel against the oracle on reduced scenes); the work counters see at
  least one closest-hit ray per path.
- C3 (512 BLAS, 7.9M triangles, 31.9M rays): one batched launch equals the
  512 per-BLAS launches of the same engine-order kernel on every ray (the
  per-BLAS kernel is pinned bit for bit against the oracle, C0 at full
  size included).
- C4's cache grid (12 levels x 2^19 entries, 16..2048): inference and
  gradients at that size against the FP64 cache oracle on a query sample.
"""
import math

import numpy as np
import pytest

import atlas_rt
import nrc_oracle as O
from atlas_rt import _abi
from atlas_rt import workloads as W
from atlas_rt.bench_suite import _world

pytestmark = pytest.mark.gpu


def test_c2_full_frame_resident_equals_bounce_by_bounce(torch_cuda):
    print("case full-size c2", flush=True)
    meshes, mats_d, xf, which = W.c2_scene()
    w, blases = _world(meshes, xf, which)
    cc = W.c2_camera()
    cam = atlas_rt.camera(cc["eye"], cc["yaw_deg"], cc["pitch_deg"], cc["fov_y_deg"], cc["width"] / cc["height"])
    mats = W.c2_materials(mats_d)
    imgs = []
    for wavefront in (False, True):
        pt = atlas_rt.PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV,
                                 wavefront=wavefront)
        if not wavefront:
            pt.count_work()
        pt.render(cam, 2)
        imgs.append(pt.read().cpu().numpy())
        if not wavefront:
            work = pt.work_counts()
        pt.close()
    px = cc["width"] * cc["height"]
    print("c2 work over 2 frames", work)
    assert np.all(imgs[0][..., 3] == 2) and imgs[0][..., :3].max() > 0
    assert imgs[0].tobytes() == imgs[1].tobytes()
    assert work["closest_rays"] >= 2 * px and work["hits"] <= work["closest_rays"]
    assert work["nodes"] > work["closest_rays"] and work["instances"] > 0


def test_c3_full_batched_launch_equals_per_blas_launches(torch_cuda):
    print("case full-size c3", flush=True)
    torch = torch_cuda
    meshes, rays, table = W.c3()
    assert len(table) == 512 and len(rays) > 30_000_000
    blases = atlas_rt.build_many(meshes)
    bs = atlas_rt.BatchSet(blases, [(b, o, c, 0) for b, o, c in table])
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    del rays
    d_a = torch.full((d_rays.shape[0], 4), -7.0, dtype=torch.float32, device="cuda")
    d_b = torch.full_like(d_a, -9.0)
    bs.trace_closest(d_rays, d_a, kernel=_abi.KERNEL_BVH2_EXACT, block=256, refill=0)
    for b, o, c in table:
        blases[b].trace_closest(d_rays[o:o + c], d_b[o:o + c], kernel=_abi.KERNEL_BVH2_EXACT)
    torch.cuda.synchronize()
    a, b_ = d_a.view(torch.int32), d_b.view(torch.int32)
    n_diff = int((a != b_).any(dim=1).sum().item())
    hits = int((a[:, 3] >= 0).sum().item())
    print(f"c3 full size: {d_rays.shape[0]} rays, {hits} hits, {n_diff} differ")
    assert n_diff == 0 and hits > 0.1 * d_rays.shape[0]


def test_c4_full_grid_cache_matches_oracle(torch_cuda):
    """The C4 cache grid at its size: 12 levels of 2^19 entries, resolutions
    16 to 2048 (four dense levels, eight hashed), on random parameters."""
    print("case full-size cache grid", flush=True)
    torch = torch_cuda
    bmin, bmax = (-110.0, -30.0, -110.0), (110.0, 30.0, 110.0)
    rc = atlas_rt.RadianceCache(bmin, bmax, levels=12, log2_table=19, base_resolution=16, max_resolution=2048, seed=3)
    P = O.Params(bmin, bmax, 12, 19, 16, 2048)
    g = np.random.default_rng(31)
    m = np.zeros(O.MASTER_SIZE, np.float32)
    for l in range(6):
        lim = math.sqrt(6.0 / (O.K[l] + O.N[l]))
        m[O.MASTER_OFFSETS[l]:O.MASTER_OFFSETS[l + 1]] = g.uniform(-lim, lim, O.K[l] * O.N[l])
    t16 = g.uniform(-1, 1, (P.levels, P.T, 2)).astype(np.float16)
    rc.set_params(torch.from_numpy(m).cuda(), torch.from_numpy(t16.view(np.int16).reshape(-1).copy()).cuda())
    tab = t16.astype(np.float64)
    n = 2048
    q = np.zeros((n, 12), np.float32)
    q[:, 0:3] = g.uniform(np.array(bmin) - 5, np.array(bmax) + 5, (n, 3))
    q[:, 3] = g.uniform(0, 1, n)
    for k in (4, 8):
        d = g.normal(size=(n, 3))
        q[:, k:k + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    out = rc.infer(torch.from_numpy(q).cuda()).cpu().numpy()
    yq = O.predict(P, m, tab, q, quantize=True)
    err = np.abs(out[:, :3] - yq).max() / np.abs(yq).max()
    print(f"full grid infer: max err vs bf16-operand oracle {err:.2e}")
    assert err < 2e-3
    t = np.abs(g.normal(size=(n, 4))).astype(np.float32)
    gm, gt = rc.gradients(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda())
    gm, gt = gm.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64).reshape(P.levels, P.T, 2)
    _, gmq, gtq, _ = O.loss_and_grads(P, m, tab, q, t, quantize=True)
    em = np.linalg.norm(gm - gmq) / np.linalg.norm(gmq)
    et = np.linalg.norm(gt - gtq) / np.linalg.norm(gtq)
    print(f"full grid gradients: rel L2 vs bf16-operand oracle: weights {em:.2e}, table {et:.2e}")
    assert em < 1e-2 and et < 1e-2
    rc.close()
