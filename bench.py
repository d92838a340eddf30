"""Benchmark: Mrays/s, incoherent closest-hit, one MI355X.

Workload (one step = one closest-hit trace of the whole ray batch):
  c0: 65,536 random triangles (centroids U[-10,10]^3, edges 0.05-0.5, random
      orientation), the engine's SBVH/SAH BVH (one triangle per leaf),
      1,048,576 rays with origins U[-10,10]^3 and directions uniform on the
      sphere, tmin 1e-4; outputs (t, u, v, primID) per ray.
Rays and the BVH are resident in HBM before the timed region.

Prints one JSON line.  `roofline` prices the traversal kernel with the
algorithmic bytes of the engine layout (64 B per node visit, 48 B per
triangle test, 32 B ray in, 16 B hit out; visit counts from the kernel's own
statistics pass) over its average launch time measured with HIP events on the
launch stream.  `cpu_baseline` times the oracle's restatement of the engine
traversal on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atlas-engine_amd"))

HBM_PEAK_GBPS = 8000.0
REG = 0x10
FSLAB = 0x20
LEAFB = 0x40
EXACT_KERNELS = (1, 2, 1 | REG, 2 | REG, 2 | LEAFB, 2 | REG | LEAFB)  # engine order and arithmetic: bit-exact
KERNEL_NAMES = {1: "bvh2-exact grid (engine order, bit-exact)",
                2: "bvh2-exact persistent wave-refill (engine order, bit-exact)",
                3: "bvh4 persistent (wide nodes, same nearest hit)",
                4: "bvh8 persistent (wide nodes, same nearest hit)",
                1 | REG: "bvh2-exact grid, register stack top",
                2 | REG: "bvh2-exact persistent, register stack top",
                3 | REG: "bvh4 persistent, register stack top",
                4 | REG: "bvh8 persistent, register stack top",
                1 | FSLAB: "bvh2 grid, widened float slab (same nearest hit)",
                2 | FSLAB: "bvh2 persistent, widened float slab (same nearest hit)",
                1 | REG | FSLAB: "bvh2 grid, register stack top, widened float slab (same nearest hit)",
                2 | REG | FSLAB: "bvh2 persistent, register stack top, widened float slab (same nearest hit)",
                2 | LEAFB: "bvh2-exact persistent, leaf-batched waves",
                2 | REG | LEAFB: "bvh2-exact persistent, register stack top, leaf-batched waves",
                2 | FSLAB | LEAFB: "bvh2 persistent, widened float slab, leaf-batched waves (same nearest hit)",
                2 | REG | FSLAB | LEAFB: "bvh2 persistent, register stack top, widened float slab, leaf-batched "
                                         "waves (same nearest hit)"}
METRIC = "Mrays/s incoherent closest-hit; path-trace time to equal rel-MSE, cache vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kernel", type=int, default=0, help="atlas_rt kernel variant (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work for the baseline sample")
    ap.add_argument("--no-extra", action="store_true", help="only the c0 headline (skip c1, c3 and the cache study)")
    ap.add_argument("--ref-spp", type=int, default=4096, help="samples per pixel of the plain reference render")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Hardware-counter summary of one traversal variant (kernel id -> file): the
# traffic it reports belongs to that kernel only.
PMC_SUMMARIES = {18: os.path.join(ROOT, "profiles", "k_blas_closest_persistent_regstack_c0_pmc.md"),
                 50: os.path.join(ROOT, "profiles", "k_blas_closest_persistent_regstack_fslab_c0_pmc.md"),
                 66: os.path.join(ROOT, "profiles", "k_blas_closest_persistent_leafbatch_c0_pmc.md")}


def pmc_traffic(kernel):
    """HBM traffic per launch of the selected C0 kernel from the committed
    FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy kernel), plus the
    L2 hit rate and VALU busy fraction of the same runs; None if that kernel
    has no counter summary."""
    path = PMC_SUMMARIES.get(kernel)
    if path is None:
        return None
    try:
        text = open(path).read()
    except OSError:
        return None
    import re
    out = {}
    m = re.search(r"traffic_bytes_per_launch: (\d+)", text)
    if not m:
        return None
    out["traffic"] = int(m.group(1))
    out["source"] = os.path.relpath(path, ROOT) + " (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    m = re.search(r"L2 \(TCC\) hit rate: ([0-9.]+)%", text)
    out["l2_hit_rate"] = float(m.group(1)) / 100 if m else None
    m = re.search(r"VALU busy: .*= ([0-9.]+)%", text)
    out["valu_busy"] = float(m.group(1)) / 100 if m else None
    return out


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(nodes, tris, orig, rays, gpu_hits, budget_s):
    """Oracle (restated engine traversal) on the host cores, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_rt

    threads = host_threads()
    probe = rays[: min(len(rays), 32768)]
    t0 = time.perf_counter()
    oracle_rt.trace_closest(nodes, tris, orig, probe, threads=threads)
    rate = len(probe) / max(time.perf_counter() - t0, 1e-6)
    n = int(min(len(rays), max(65536, rate * budget_s / 6.0)))
    sample = rays[:n]
    times = []
    ref = None
    for i in range(6):  # 1 warm-up + median of 5
        t0 = time.perf_counter()
        ref, _ = oracle_rt.trace_closest(nodes, tris, orig, sample, threads=threads)
        if i:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    mism = int(np.sum(ref["prim_id"] != gpu_hits["prim_id"][:n]))
    tbits = int(np.sum(ref["t"].view(np.int32) != gpu_hits["t"].view(np.int32)[:n]))
    return {
        "value": round(n / med / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} of the {len(rays)} c0 rays, R-GPU traversal restated in C++ "
                  f"(-O3, no FMA contraction), {threads} threads, median of 5 after 1 warm-up",
    }, {"checked_rays": n, "prim_id_mismatches": mism, "t_bit_mismatches": tbits}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is design-only "
                                   "(rays and BLAS batches are independent), run with --gpus 1"}))
        sys.exit(2)
    import torch

    import atlas_rt
    from atlas_rt import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    verts, rays = W.c0()
    t0 = time.perf_counter()
    blas = atlas_rt.Blas.build(verts, parallel=True)
    build_s = time.perf_counter() - t0
    nodes, tris, orig = blas.export()
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    d_vis = torch.empty((n, 2), dtype=torch.int32, device="cuda")

    # statistics pass (outside the timed region): algorithmic bytes per launch
    blas.trace_closest_stats(d_rays, d_hits, d_vis)
    torch.cuda.synchronize()
    vis = d_vis.cpu().numpy().astype(np.int64)
    node_visits = int(vis[:, 0].sum())
    tri_tests = int(vis[:, 1].sum())
    alg_bytes = 48 * n + 64 * node_visits + 48 * tri_tests

    stream = torch.cuda.current_stream()
    kernel = args.kernel
    variants = {}
    if kernel == 0:
        # pick the fastest traversal variant (outside the timed region)
        for k in (1, 2, 3, 4, 1 | REG, 2 | REG, 3 | REG, 4 | REG, 1 | FSLAB, 2 | FSLAB, 1 | REG | FSLAB,
                  2 | REG | FSLAB, 2 | LEAFB, 2 | REG | LEAFB, 2 | FSLAB | LEAFB, 2 | REG | FSLAB | LEAFB):
            for _ in range(3):
                blas.trace_closest(d_rays, d_hits, kernel=k, stream=stream)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(5):
                blas.trace_closest(d_rays, d_hits, kernel=k, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            variants[k] = e0.elapsed_time(e1) / 5
        # the engine-exact kernels win unless another variant (same nearest
        # hit, other traversal order) is more than 3% faster
        exact = min((k for k in variants if k in EXACT_KERNELS), key=variants.get)
        fastest = min(variants, key=variants.get)
        kernel = fastest if variants[fastest] < 0.97 * variants[exact] else exact
    for _ in range(args.warmup):
        blas.trace_closest(d_rays, d_hits, kernel=kernel, stream=stream)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        blas.trace_closest(d_rays, d_hits, kernel=kernel, stream=stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1000.0 / args.steps
    value = n / (ms_per_step / 1000.0) / 1e6

    hits = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
    achieved = alg_bytes / (kernel_ms / 1000.0) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c0: incoherent closest-hit, 65536 random triangles, engine SBVH (1 tri/leaf), "
                        "1048576 rays U[-10,10]^3 x uniform sphere, tmin 1e-4",
            "triangles": int(len(verts)),
            "rays": n,
            "bvh_nodes": int(len(nodes)),
            "kernel": KERNEL_NAMES.get(kernel, str(kernel)),
            "variant_ms": {KERNEL_NAMES.get(k, str(k)): round(v, 4) for k, v in variants.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "traffic_source": None,
            "kernel_ms": round(kernel_ms, 4),
            "alg_bytes_per_launch": alg_bytes,
            "node_visits_per_ray": round(node_visits / n, 3),
            "tri_tests_per_ray": round(tri_tests / n, 3),
        },
        "build_s": round(build_s, 3),
        "hit_fraction": round(float((hits["prim_id"] >= 0).mean()), 4),
    }
    pmc = pmc_traffic(kernel)
    if pmc:
        out["roofline"]["traffic"] = pmc["traffic"]
        out["roofline"]["traffic_source"] = pmc["source"]
        out["roofline"]["traffic_GBps"] = round(pmc["traffic"] / (kernel_ms / 1000.0) / 1e9, 1)
        out["roofline"]["l2_hit_rate"] = pmc["l2_hit_rate"]
        out["roofline"]["valu_busy"] = pmc["valu_busy"]
    if not args.no_cpu_baseline:
        cb, parity = cpu_baseline(nodes, tris, orig, rays, hits, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["parity"] = parity
        out["gpu_vs_cpu"] = round(value / cb["value"], 2)
    log(f"c0 done: {value:.1f} Mrays/s")
    if not args.no_extra:
        from atlas_rt import bench_suite
        del d_rays, d_hits, d_vis, blas
        torch.cuda.empty_cache()
        out["c1"] = bench_suite.c1(torch, log)
        out["c2"] = bench_suite.c2(torch, log)
        out["tlas_update"] = bench_suite.tlas_update(torch, log)
        out["c3"] = bench_suite.c3(torch, log)
        out["radiance_cache"] = bench_suite.cache_study(torch, log, args.ref_spp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
