"""Secondary measurements of the benchmark (all on one GPU, inputs resident
in device memory before each timed region):

  c1           coherent primary rays over a 1024 x 1024 heightfield
               (2,093,058 triangles) at 1920 x 1080, then one any-hit shadow
               ray per hit toward a directional light: Mrays/s over both.
  c2           two-level path trace, 256 instances of 8 meshes, 1280 x 720,
               4 bounces, sun + constant environment: ms per frame (one
               sample per pixel), without and with ray binning, and a
               realtime frame (4 samples per pixel and the temporal resolve).
  tlas_update  per-frame TLAS update of C2's meshes at 256 and 65,536
               instances: the host rebuild (atlas_rt_world_update) against
               the device refit (atlas_rt_world_refit, matrices resident), and
               1M random rays traced after a small motion on the refitted and
               on the rebuilt tree.
  c3           512 independent BLAS (1k-64k triangles, log-uniform) each with
               its own random ray batch (4k-256k rays, log-uniform), traced
               by one autotuned persistent launch and by its captured graph.
  cache_study  time to equal rel-MSE against a long plain render, plain
               path tracer vs the radiance-cache path tracer, 1280 x 720,
               8 bounces, sun + constant environment; on the config-2 layout
               and on the dense radiance-cache layout.
"""
import math
import time

import numpy as np

from . import (HIT_DTYPE, Blas, BatchSet, PathTracer, RadianceCache, World, build_many, build_many_device, camera,
               instance_descs)
from . import workloads as W


def _sync_time(torch, fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t


EXACT_KERNELS = (1, 2, 2 | 0x40)  # engine order and arithmetic: the engine's slots bit for bit


def c1(torch, log):
    """Primary rays timed on every kernel; the headline is the fastest
    engine-exact one.  The shadow batch holds one ray per primary hit.  The
    returned "_check" entry carries the arrays the caller's CPU leg compares
    (it is removed before printing)."""
    verts, prim = W.c1()
    t = time.perf_counter()
    blas = Blas.build(verts)
    build_s = time.perf_counter() - t
    n = len(prim)
    d_prim = torch.from_numpy(prim.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(n, dtype=torch.int32, device="cuda")
    res = {}
    for kernel in (1, 2, 2 | 0x40, 3, 4):
        for _ in range(2):
            blas.trace_closest(d_prim, d_hits, kernel=kernel)
        steps = 10
        tp = _sync_time(torch, lambda: [blas.trace_closest(d_prim, d_hits, kernel=kernel) for _ in range(steps)])
        res[kernel] = tp / steps
    best = min(EXACT_KERNELS, key=res.get)
    blas.trace_closest(d_prim, d_hits, d_slots, kernel=best)
    torch.cuda.synchronize()
    hits = d_hits.cpu().numpy().view(HIT_DTYPE).reshape(-1).copy()
    slots = d_slots.cpu().numpy().copy()
    hit = hits["prim_id"] >= 0
    shadow = np.ascontiguousarray(W.shadow_rays(prim, hits, verts)[hit])
    n_shadow = len(shadow)
    d_sh = torch.from_numpy(shadow.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_occ = torch.empty(n_shadow, dtype=torch.int32, device="cuda")
    for _ in range(2):
        blas.trace_any(d_sh, d_occ)
    steps = 10
    ts = _sync_time(torch, lambda: [blas.trace_any(d_sh, d_occ) for _ in range(steps)]) / steps
    occ = d_occ.cpu().numpy().astype(np.uint32)
    total = res[best] + ts
    # algorithmic bytes of the primary batch (a statistics pass outside the
    # timed runs): 48 B per ray + 64 B per node visit + 48 B per triangle test
    d_vis = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    blas.trace_closest_stats(d_prim, d_hits, d_vis)
    torch.cuda.synchronize()
    vis = d_vis.cpu().numpy().astype(np.int64)
    prim_bytes = 48 * n + 64 * int(vis[:, 0].sum()) + 48 * int(vis[:, 1].sum())
    prim_gbps = prim_bytes / res[best] / 1e9
    log(f"c1: primary {n / res[best] / 1e6:.0f} Mrays/s (kernel {best}), shadow {n_shadow / ts / 1e6:.0f} Mrays/s; "
        f"primary {prim_bytes / 1e9:.2f} GB algorithmic, {prim_gbps:.0f} GB/s")
    nodes, tris, orig = blas.export()
    return {
        "workload": "c1: 1024x1024 value-noise heightfield (2,093,058 tris), 1920x1080 pinhole at 30 deg pitch, "
                    "plus one any-hit shadow ray per primary hit toward a directional light",
        "triangles": int(len(verts)), "primary_rays": n, "primary_hits": int(hit.sum()), "shadow_rays": n_shadow,
        "build_s": round(build_s, 3),
        "primary_ms": round(res[best] * 1e3, 4), "primary_kernel": best,
        "primary_kernel_note": "fastest engine-exact kernel (1 grid, 2 persistent, 66 persistent leaf-batched); "
                               "3 and 4 (4-/8-wide nodes, same nearest t, other slot at ties) are timed beside it",
        "primary_ms_by_kernel": {str(k): round(v * 1e3, 4) for k, v in res.items()},
        "shadow_ms": round(ts * 1e3, 4),
        "mrays_per_s": round((n + n_shadow) / total / 1e6, 1),
        "primary_roofline": {"bound": "hbm", "achieved": round(prim_gbps, 1), "peak": 8000.0, "unit": "GB/s",
                             "frac": round(prim_gbps / 8000.0, 4), "alg_bytes": int(prim_bytes),
                             "node_visits_per_ray": round(float(vis[:, 0].mean()), 3),
                             "tri_tests_per_ray": round(float(vis[:, 1].mean()), 3),
                             "note": "48 B per ray + 64 B per node visit + 48 B per triangle test over the "
                                     "primary batch's kernel time; coherent camera rays share most node fetches "
                                     "in the L2"},
        "_check": {"nodes": nodes, "tris": tris, "orig": orig, "primary": prim, "hits": hits, "slots": slots,
                   "shadow": shadow, "occluded": occ},
    }


def c2(torch, log, frames=16):
    meshes, mats_d, xf, which = W.c2_scene()
    w, blases = _world(meshes, xf, which)
    cc = W.c2_camera()
    cam = camera(cc["eye"], cc["yaw_deg"], cc["pitch_deg"], cc["fov_y_deg"], cc["width"] / cc["height"])
    mats = W.c2_materials(mats_d)
    # bounce-0 rays alone: one camera ray per pixel centre through the
    # two-level world (TLAS, instance transform, BLAS), the trace the CPU
    # leg times and checks
    prim = W.pinhole_rays(cc["width"], cc["height"], cc["eye"], cc["yaw_deg"], cc["pitch_deg"], cc["fov_y_deg"])
    npx = len(prim)
    d_prim = torch.from_numpy(prim.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_ph = torch.empty((npx, 4), dtype=torch.float32, device="cuda")
    d_pi = torch.empty(npx, dtype=torch.int32, device="cuda")
    for _ in range(2):
        w.trace_closest_rays(d_prim, d_ph, d_pi)
    t_prim = _sync_time(torch, lambda: [w.trace_closest_rays(d_prim, d_ph, d_pi) for _ in range(10)]) / 10
    tl, inst, order = w.export()
    check = {"tlas": tl, "instances": inst, "order": order, "blases": [b.export() for b in blases],
             "primary": prim, "hits": d_ph.cpu().numpy().view(HIT_DTYPE).reshape(-1).copy(),
             "instance": d_pi.cpu().numpy().copy()}
    res = {}
    imgs = {}
    # (ray binning, frames in flight, bounce-by-bounce launches): the
    # library's default (resident frame kernel), binned, one frame at a time,
    # and the engine's per-bounce dispatch shape
    shapes = ((False, None, False), (True, None, False), (False, 1, False), (False, None, True), (False, 1, True))
    def warm(pt):
        # the first six resident launches time the two grid shapes and pick one
        for _ in range(6):
            pt.render(cam, 1)

    for binning, fif, wavefront in shapes:
        pt = PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV, binning=binning,
                        frames_in_flight=fif, wavefront=wavefront)
        warm(pt)
        pt.reset()
        t = _sync_time(torch, lambda: pt.render(cam, frames))
        res[(binning, fif, wavefront)] = t / frames
        imgs[(binning, fif, wavefront)] = pt.read().cpu().numpy()
        del pt
    # the engine's own pattern: one Render call (one sample per pixel) per
    # displayed frame, the library's frames in flight overlapping the calls
    pt = PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV)
    warm(pt)
    pt.reset()
    t_calls = _sync_time(torch, lambda: [pt.render(cam, 1) for _ in range(frames)]) / frames
    same_calls = bool(pt.read().cpu().numpy().tobytes() == imgs[(False, None, False)].tobytes())
    del pt
    base = imgs[(False, None, False)].tobytes()
    same = bool(base == imgs[(True, None, False)].tobytes())
    same_serial = bool(base == imgs[(False, 1, False)].tobytes())
    same_wavefront = all(base == imgs[(False, f, True)].tobytes() for f in (None, 1))
    res[False], res[True] = res[(False, None, False)], res[(True, None, False)]
    px = cc["width"] * cc["height"]
    # traversal work of one frame (a counting run outside the timed ones):
    # algorithmic bytes = 64 B per inner node (TLAS or BLAS), 48 B per
    # triangle test, 64 B per instance record, plus per closest hit the
    # 96 B shading record and 96 B material the shading reads
    pt = PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV)
    pt.count_work()
    pt.render(cam, 1)
    work = pt.work_counts()
    del pt
    # a batch of 8 frames captured once as a HIP graph and replayed
    pt = PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV)
    warm(pt)
    pt.capture(cam, 8)
    pt.replay()
    t_graph = _sync_time(torch, lambda: [pt.replay() for _ in range(frames // 8)]) / (8 * (frames // 8))
    del pt
    trav_bytes = 64 * work["nodes"] + 48 * work["triangles"] + 64 * work["instances"]
    frame_bytes = trav_bytes + 192 * work["hits"]
    achieved = frame_bytes / res[False] / 1e9
    # realtime mode at the engine's defaults (4 samples per pixel, temporal resolve)
    from . import frame_matrices, projection_matrix, view_matrix
    pt = PathTracer(w, cc["width"], cc["height"], 4, mats, W.sun_light(), env=W.C2_ENV)
    pt.set_realtime(samples_per_frame=4)
    view = view_matrix(cc["eye"], cc["yaw_deg"], cc["pitch_deg"])
    proj = projection_matrix(cc["fov_y_deg"], cc["width"] / cc["height"])
    fm = frame_matrices(view, proj, proj @ view)
    for _ in range(6):
        pt.render_realtime(cam, fm)
    t_rt = _sync_time(torch, lambda: [pt.render_realtime(cam, fm) for _ in range(frames)]) / frames
    del pt
    log(f"c2 realtime: {t_rt * 1e3:.3f} ms/frame at 4 samples per pixel with the temporal resolve")
    log(f"c2: {res[False] * 1e3:.3f} ms/frame plain, {res[True] * 1e3:.3f} ms/frame binned (same image: {same}), "
        f"{res[(False, 1, False)] * 1e3:.3f} ms/frame one frame at a time (same image: {same_serial}); "
        f"bounce-by-bounce {res[(False, None, True)] * 1e3:.3f} / {res[(False, 1, True)] * 1e3:.3f} "
        f"(same image: {same_wavefront}); one frame per call {t_calls * 1e3:.3f} (same image: {same_calls}); "
        f"{frame_bytes / 1e6:.0f} MB/frame algorithmic, {achieved:.0f} GB/s")
    return {
        "workload": "c2: 256 instances of 8 random-blob meshes, 1280x720, 4 bounces, sun + constant env, "
                    "one sample per pixel per frame",
        "triangles": int(sum(len(m[0]) for m in meshes)),
        "instanced_triangles": int(sum(len(meshes[k][0]) for k in which)), "instances": int(len(which)),
        "ms_per_frame": round(res[False] * 1e3, 3),
        "mpaths_per_s": round(px / res[False] / 1e6, 1),
        "frames_in_flight": 3,
        "launch_shape": "resident frame kernel, up to 8 frames per launch",
        "graph_replay_ms_per_frame": round(t_graph * 1e3, 3),
        "graph_replay_note": "8 frames captured once as a HIP graph (atlas_rt_pathtracer_capture), replayed",
        "one_frame_at_a_time_ms_per_frame": round(res[(False, 1, False)] * 1e3, 3),
        "one_frame_per_call_ms_per_frame": round(t_calls * 1e3, 3),
        "one_frame_per_call_note": "render(camera, 1) per call, as PathTracingRenderer::Render is called once per "
                                   "displayed frame; three frames in flight; same image as one 16-frame call: "
                                   + str(same_calls),
        "bounce_by_bounce_ms_per_frame": round(res[(False, None, True)] * 1e3, 3),
        "bounce_by_bounce_one_frame_at_a_time_ms_per_frame": round(res[(False, 1, True)] * 1e3, 3),
        "bounce_by_bounce_image_identical": same_wavefront,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "alg_bytes_per_frame": int(frame_bytes),
                     "traversal_bytes_per_frame": int(trav_bytes), "work_per_frame": work,
                     "note": "algorithmic bytes of one frame from the kernel's own work counters (64 B per node, "
                             "48 B per triangle test, 64 B per instance, 192 B of shading reads per hit) over "
                             "the measured ms/frame"},
        "in_flight_image_identical": same_serial,
        "binned_ms_per_frame": round(res[True] * 1e3, 3),
        "binning_speedup": round(res[False] / res[True], 3),
        "binned_image_identical": same,
        "realtime_ms_per_frame": round(t_rt * 1e3, 3),
        "realtime_note": "realtime mode: 4 samples per pixel, G-buffer, temporal resolve",
        "world_primary_ms": round(t_prim * 1e3, 4),
        "world_primary_mrays_per_s": round(npx / t_prim / 1e6, 1),
        "world_primary_note": "one closest-hit camera ray per pixel centre through TLAS + instances + BLAS "
                              "(atlas_rt_world_trace_closest_rays)",
        "_check": check,
    }


def tlas_update(torch, log, sizes=(256, 65536), reps=10):
    meshes, _, _, _ = W.c2_scene()
    blases = []
    for v, nrm in meshes:
        b = Blas.build(v)
        blases.append(b)
    out = {"workload": "C2's 8 blob meshes; instances random affine over [-100, 100] x [-20, 20] x [-100, 100] "
                       "at 256 instances, each side scaled by (n / 256)^(1/3) above; motion = every instance "
                       "translated by U[-0.5, 0.5]^3",
           "sizes": []}
    for n in sizes:
        g = W.rng(2, 4000 + n)
        grow = (n / 256) ** (1.0 / 3.0)
        xf = W.random_affine(g, n, extent=100.0 * grow, height=20.0 * grow)
        which = g.integers(0, len(meshes), n)
        descs = instance_descs(xf, which)
        w = World()
        w.update(descs, blases)
        t_host = min(_sync_time(torch, lambda: w.update(descs, blases)) for _ in range(3))
        xf2 = xf.copy()
        xf2[:, :, 3] += g.uniform(-0.5, 0.5, (n, 3))
        moved = instance_descs(xf2, which)
        d_moved = torch.from_numpy(moved.view(np.uint8).copy()).cuda()
        w.refit(d_moved)
        t_refit = _sync_time(torch, lambda: [w.refit(d_moved) for _ in range(reps)]) / reps
        rays = W.random_rays(1 << 20, g, extent=100.0 * grow)
        d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
        d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        w.trace_closest_rays(d_rays, d_hits)
        t_tr_refit = _sync_time(torch, lambda: [w.trace_closest_rays(d_rays, d_hits) for _ in range(3)]) / 3
        h_refit = d_hits.cpu().numpy().view(HIT_DTYPE).reshape(-1).copy()
        w2 = World()
        w2.update(moved, blases)
        w2.trace_closest_rays(d_rays, d_hits)
        t_tr_build = _sync_time(torch, lambda: [w2.trace_closest_rays(d_rays, d_hits) for _ in range(3)]) / 3
        h_build = d_hits.cpu().numpy().view(HIT_DTYPE).reshape(-1)
        same_t = bool(np.array_equal(h_refit["t"].view(np.int32), h_build["t"].view(np.int32)))
        log(f"tlas {n}: host rebuild {t_host * 1e3:.2f} ms, device refit {t_refit * 1e3:.3f} ms; trace "
            f"{t_tr_refit * 1e3:.2f} ms refit vs {t_tr_build * 1e3:.2f} ms rebuilt (same t: {same_t})")
        out["sizes"].append({
            "instances": n, "host_rebuild_ms": round(t_host * 1e3, 3), "device_refit_ms": round(t_refit * 1e3, 4),
            "trace_1m_rays_after_motion_ms": {"refit": round(t_tr_refit * 1e3, 3),
                                             "rebuilt": round(t_tr_build * 1e3, 3)},
            "same_hit_distances": same_t,
        })
        del w, w2
    return out


def c3(torch, log):
    meshes, rays, table = W.c3()
    t = time.perf_counter()
    blases = build_many(meshes)
    build_s = time.perf_counter() - t
    bs = BatchSet(blases, [(b, o, c, 0) for b, o, c in table])
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    t = time.perf_counter()
    bs.trace_closest(d_rays, d_hits)  # first AUTO launch runs the autotuner
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t
    tuned, trials = bs.tuning()
    # headline: the fastest engine-exact candidate (binary nodes, exact slab);
    # the tuner's overall choice is reported beside it
    exact = [(c, ms) for c, ms in trials if (c["kernel"] & 0xF) in (1, 2) and not c["kernel"] & 0x20]
    best = min(exact, key=lambda x: x[1])[0]
    cfg = dict(kernel=best["kernel"], block=best["block"], refill=best["refill"])
    steps = 5
    bs.trace_closest(d_rays, d_hits, **cfg)
    tl = _sync_time(torch, lambda: [bs.trace_closest(d_rays, d_hits, **cfg) for _ in range(steps)]) / steps
    tt = _sync_time(torch, lambda: [bs.trace_closest(d_rays, d_hits) for _ in range(steps)]) / steps
    bs.graph_capture(d_rays, d_hits, **cfg)
    tg = _sync_time(torch, lambda: [bs.graph_launch() for _ in range(steps)]) / steps
    torch.cuda.synchronize()
    bs.trace_closest(d_rays, d_hits, **cfg)
    torch.cuda.synchronize()
    # the CPU leg checks and times the first 24 batches
    n_check = 24
    hits_all = d_hits.cpu().numpy().view(HIT_DTYPE).reshape(-1)
    check = {"batches": []}
    for b, o, c in table[:n_check]:
        check["batches"].append((blases[b].export(), rays[o:o + c], hits_all[o:o + c].copy()))
    # statistics pass (outside the timed launches): every batch's inner
    # nodes and triangle tests through the engine-order kernel, for the
    # algorithmic bytes of a launch (48 B per ray, 64 B per node, 48 B per
    # triangle test)
    d_vis = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    d_sh = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    for b, o, c in table:
        blases[b].trace_closest_stats(d_rays[o:o + c], d_sh[o:o + c], d_vis[o:o + c])
    torch.cuda.synchronize()
    vis = d_vis.cpu().numpy().astype(np.int64)
    # compare bits: a miss's prim_id -1 reads as a float NaN
    same = bool(torch.equal(d_sh.view(torch.int32), d_hits.view(torch.int32)))
    del d_vis, d_sh
    nodes, tests = int(vis[:, 0].sum()), int(vis[:, 1].sum())
    alg = 48 * n + 64 * nodes + 48 * tests
    achieved = alg / tl / 1e9
    log(f"c3: {n / tl / 1e6:.0f} Mrays/s launch (engine-exact {best}), {n / tg / 1e6:.0f} Mrays/s graph; "
        f"tuner's pick {tuned}: {n / tt / 1e6:.0f} Mrays/s; {alg / 1e9:.2f} GB algorithmic, {achieved:.0f} GB/s")
    # the same meshes built on the GPU (linear BVH), traced by an autotuned launch
    del bs
    t = time.perf_counter()
    dev_blases, dev_ms = build_many_device(meshes)
    dev_build_s = time.perf_counter() - t
    bsd = BatchSet(dev_blases, [(b, o, c, 0) for b, o, c in table])
    bsd.trace_closest(d_rays, d_hits)
    torch.cuda.synchronize()
    best_d, _ = bsd.tuning()
    tld = _sync_time(torch, lambda: [bsd.trace_closest(d_rays, d_hits) for _ in range(steps)]) / steps
    log(f"c3 device build: {dev_ms:.1f} ms GPU, {dev_build_s:.2f} s wall; trace {n / tld / 1e6:.0f} Mrays/s")
    del bsd, dev_blases
    return {
        "workload": "c3: 512 BLAS of 1k-64k triangles (log-uniform), each with its own 4k-256k random rays "
                    "(log-uniform), one persistent launch",
        "blas": len(meshes), "triangles": int(sum(len(m) for m in meshes)), "rays": n,
        "build_s": round(build_s, 2), "autotune_s": round(tune_s, 2),
        "best": best, "best_note": "fastest engine-exact candidate of the autotuner's trials (binary nodes, "
                                   "exact slab: the engine's slots bit for bit)",
        "tuner_pick": tuned, "tuner_pick_ms": round(tt * 1e3, 3),
        "trials": [{"config": c, "ms": round(ms, 3)} for c, ms in trials],
        "ms": round(tl * 1e3, 3), "graph_ms": round(tg * 1e3, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "alg_bytes_per_launch": alg,
                     "node_visits_per_ray": round(nodes / n, 3), "tri_tests_per_ray": round(tests / n, 3),
                     "stats_hits_equal_launch": same,
                     "note": "48 B per ray + 64 B per inner node + 48 B per triangle test, counted per batch by "
                             "the engine-order statistics kernel, over the engine-exact launch's measured time"},
        "mrays_per_s": round(n / tl / 1e6, 1), "graph_mrays_per_s": round(n / tg / 1e6, 1),
        "_check": check,
        "device_build": {
            "builder": "GPU linear BVH (Morton order, radix tree), engine layout",
            "gpu_ms": round(dev_ms, 2), "wall_s": round(dev_build_s, 2),
            "note": "wall time includes the triangle uploads, the host mirrors and the wide-node collapse",
            "trace_ms": round(tld * 1e3, 3), "trace_mrays_per_s": round(n / tld / 1e6, 1), "best": best_d,
        },
    }


# ---------------------------------------------------------------------------
# radiance cache vs plain path tracing
# ---------------------------------------------------------------------------

def _world(meshes, xf, which):
    blases = []
    for v, nrm in meshes:
        b = Blas.build(v)
        b.set_shading(nrm)
        blases.append(b)
    descs = instance_descs(xf, which)
    descs["material_offset"] = which.astype(np.int32)
    w = World()
    w.update(descs, blases)
    return w, blases


def rel_mse(torch, img, ref):
    """mean over pixels and channels of (x - r)^2 / (r^2 + 0.01), linear HDR."""
    x = img[..., :3] / img[..., 3:4].clamp(min=1)
    r = ref[..., :3] / ref[..., 3:4].clamp(min=1)
    return float(torch.mean((x - r) ** 2 / (r * r + 0.01)).item())


def _curve(torch, pt, cam, ref, checkpoints, t0=0.0):
    """(spp, seconds, rel-MSE) after each checkpoint; rendering time only."""
    out, t, done = [], t0, 0
    for spp in checkpoints:
        t += _sync_time(torch, lambda: pt.render(cam, spp - done))
        done = spp
        out.append((spp, t, rel_mse(torch, pt.read(), ref)))
    return out


def _time_to(curve, target):
    """Seconds for a (spp, seconds, rel-MSE) curve to reach `target`,
    interpolated in log-log between checkpoints; None if never reached."""
    prev = None
    for spp, t, e in curve:
        if e <= target:
            if prev is None or prev[2] <= e:
                return t
            lt = math.log(prev[1]) + (math.log(t) - math.log(prev[1])) * \
                (math.log(prev[2]) - math.log(target)) / (math.log(prev[2]) - math.log(e))
            return math.exp(lt)
        prev = (spp, t, e)
    return None


# cache settings compared: (training-pixel stride, Adam steps per frame, warm-up frames, query lag)
CACHE_SETTINGS = ((128, 1, 8, 1), (512, 1, 8, 1), (2048, 1, 8, 1), (2048, 1, 2, 1), (2048, 2, 2, 1),
                  (4096, 1, 8, 1), (128, 1, 8, 0))


def _study(torch, log, name, meshes, mats, xf, which, cc, bounds, ref_spp, bounces=8,
           checkpoints=(1, 2, 4, 8, 16, 32, 64, 128, 256)):
    w, blases = _world(meshes, xf, which)
    mats = W.c2_materials(mats)
    sun = W.sun_light()
    width, height = cc["width"], cc["height"]
    cam = camera(cc["eye"], cc["yaw_deg"], cc["pitch_deg"], cc["fov_y_deg"], width / height)
    ref_pt = PathTracer(w, width, height, bounces, mats, sun, env=W.C2_ENV)
    t_ref = 0.0
    chunk = 256
    for k in range(0, ref_spp, chunk):
        t_ref += _sync_time(torch, lambda: ref_pt.render(cam, min(chunk, ref_spp - k)))
        log(f"{name}: reference {k + chunk} / {ref_spp} spp")
    ref = ref_pt.read().clone()
    del ref_pt
    plain = _curve(torch, PathTracer(w, width, height, bounces, mats, sun, env=W.C2_ENV), cam, ref, checkpoints)
    runs = []
    for train_stride, train_steps, warmup, lag in CACHE_SETTINGS:
        rc = RadianceCache(bounds[0], bounds[1], levels=12, log2_table=19, base_resolution=16, max_resolution=2048,
                           learning_rate=1e-2)
        pt = PathTracer(w, width, height, bounces, mats, sun, env=W.C2_ENV)
        pt.set_cache(rc, train_stride=train_stride, train_steps=train_steps, query_lag=lag)
        t_warm = _sync_time(torch, lambda: pt.render(cam, warmup))
        samples, queries, loss = pt.cache_stats()
        pt.reset()
        cache = _curve(torch, pt, cam, ref, checkpoints, t0=t_warm)
        speedups = {}
        for spp, t_p, e_p in plain:
            if spp < 4:
                continue
            t_c = _time_to(cache, e_p)
            speedups[str(spp)] = None if t_c is None else round(t_p / t_c, 3)
        log(f"{name}: stride {train_stride} steps {train_steps} warm-up {warmup} lag {lag}: speedups {speedups}")
        runs.append({
            "train_stride": train_stride, "train_steps_per_frame": train_steps, "warmup_frames": warmup,
            "query_lag": lag,
            "warmup_ms": round(t_warm * 1e3, 1), "samples_per_frame": samples, "queries_per_frame": queries,
            "last_loss": round(loss, 4),
            "cache_curve": [[s, round(t * 1e3, 2), round(e, 6)] for s, t, e in cache],
            "speedup_to_plain_rel_mse_at_spp": speedups,
        })
        del pt, rc
    head = max((r["speedup_to_plain_rel_mse_at_spp"].get("64") or 0.0) for r in runs)
    return {
        "workload": name,
        "resolution": [width, height], "bounces": bounces, "reference_spp": ref_spp,
        "reference_ms_per_spp": round(t_ref / ref_spp * 1e3, 3),
        "cache_grid": {"levels": 12, "log2_table": 19, "base_resolution": 16, "max_resolution": 2048},
        "plain_curve": [[s, round(t * 1e3, 2), round(e, 6)] for s, t, e in plain],
        "cache_runs": runs,
        "speedup_at_plain_64spp": head,
    }


def cache_study(torch, log, ref_spp):
    out = {"definition": "plain path tracer's rel-MSE vs the reference after N spp and its render time T_plain; "
                         "the cache path tracer's render time T_cache (including its warm-up frames) to reach that "
                         "rel-MSE, log-log interpolated between checkpoints; speedup = T_plain / T_cache (None: "
                         "the cache's bias floor stays above the target)"}
    meshes, mats, xf, which = W.c4_scene()
    out["dense"] = _study(torch, log, "c4: config-2 meshes and materials, 256 instances in a 30-unit cube, camera at "
                                      "its edge", meshes, mats, xf, which, W.c4_camera(),
                          ((-20.0, -20.0, -20.0), (20.0, 20.0, 20.0)), ref_spp)
    meshes, mats, xf, which = W.c2_scene()
    out["c2_layout"] = _study(torch, log, "c2 layout: 256 instances over 200 x 40 x 200, camera outside", meshes, mats,
                              xf, which, W.c2_camera(), ((-110.0, -30.0, -110.0), (110.0, 30.0, 110.0)), ref_spp)
    out["speedup_at_plain_64spp"] = out["dense"]["speedup_at_plain_64spp"]
    return out
