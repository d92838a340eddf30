"""Batched traces (many BLAS / ray-batch pairs in one launch): every launch
shape the autotuner may pick, and graph replays, equal the oracle's engine
traversal per batch bit for bit."""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

pytestmark = pytest.mark.gpu


def _small_c3(salt=0):
    return W.c3(n_batches=24, tri_range=(64, 8192), ray_range=(100, 20000), seed_salt=salt)


def _expected(meshes, rays, table, blases):
    exp = np.zeros(len(rays), dtype=atlas_rt.HIT_DTYPE)
    for b, off, cnt in table:
        nodes, tris, orig = blases[b].export()
        h, _ = oracle_rt.trace_closest(nodes, tris, orig, rays[off:off + cnt])
        exp[off:off + cnt] = h
    return exp


def _run(torch, bs, rays, **kw):
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.full((len(rays), 4), -7.0, dtype=torch.float32, device="cuda")
    bs.trace_closest(d_rays, d_hits, **kw)
    torch.cuda.synchronize()
    return d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)


def test_batched_all_launch_shapes_bit_exact(torch_cuda):
    print("case batched shapes", flush=True)
    torch = torch_cuda
    meshes, rays, table = _small_c3(1)
    rays[::101]["direction"] = np.nan
    blases = atlas_rt.build_many(meshes)
    bs = atlas_rt.BatchSet(blases, [(b, o, c, 0) for b, o, c in table])
    exp = _expected(meshes, rays, table, blases)
    reg = _abi.KERNEL_STACK_REGISTER
    lb = _abi.KERNEL_LEAF_BATCH
    for kernel in (_abi.KERNEL_BVH2_EXACT, _abi.KERNEL_BVH2_PERSISTENT, _abi.KERNEL_BVH2_EXACT | reg,
                   _abi.KERNEL_BVH2_PERSISTENT | reg, _abi.KERNEL_BVH2_PERSISTENT | lb,
                   _abi.KERNEL_BVH2_PERSISTENT | reg | lb):
        for block in (64, 128, 256):
            got = _run(torch, bs, rays, kernel=kernel, block=block, refill=24)
            assert np.array_equal(rt_cases.hit_bits(got), rt_cases.hit_bits(exp)), (kernel, block)
    # wide nodes and the float slab test: the engine's nearest hit; another triangle only at an exact
    # equal-t tie (then the refined t of the other triangle may differ in the
    # last bits, so compare hit/miss and the float32 t through the oracle)
    fs = _abi.KERNEL_SLAB_FLOAT
    for kernel in (_abi.KERNEL_BVH4, _abi.KERNEL_BVH8, _abi.KERNEL_BVH4 | reg, _abi.KERNEL_BVH8 | reg,
                   _abi.KERNEL_BVH2_EXACT | reg | fs, _abi.KERNEL_BVH2_PERSISTENT | reg | fs,
                   _abi.KERNEL_BVH2_PERSISTENT | fs | lb):
        for block in (64, 256):
            got = _run(torch, bs, rays, kernel=kernel, block=block, refill=32)
            diff = np.nonzero(np.any(rt_cases.hit_bits(got) != rt_cases.hit_bits(exp), axis=1))[0]
            assert len(diff) <= max(2, len(rays) // 10000), (kernel, block, len(diff))
            assert np.array_equal(got["prim_id"] >= 0, exp["prim_id"] >= 0)
    # autotuned launch (any candidate: hit/miss and distances agree)
    got = _run(torch, bs, rays)
    assert np.array_equal(got["prim_id"] >= 0, exp["prim_id"] >= 0)
    assert np.allclose(got["t"], exp["t"], rtol=1e-6, atol=0)
    best, trials = bs.tuning()
    assert len(trials) >= 45 and best["block"] in (64, 128, 256)
    kinds = {int(t[0]["kernel"]) for t in trials}
    assert {1, 2, 3, 4, 1 | reg, 3 | reg, 4 | reg, 2 | lb, 2 | reg | lb} <= kinds


def test_batched_device_count_and_graph_replay(torch_cuda):
    """A captured graph replays correctly several times with host work and
    syncs in between, and a device-side count caps the work."""
    print("case batched graph", flush=True)
    torch = torch_cuda
    meshes, rays, table = _small_c3(2)
    blases = atlas_rt.build_many(meshes)
    # two batches share a BLAS
    table[1, 0] = table[0, 0]
    bs = atlas_rt.BatchSet(blases, [(b, o, c, 0) for b, o, c in table])
    exp = _expected(meshes, rays, table, blases)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    d_count = torch.tensor([len(rays)], dtype=torch.int32, device="cuda")
    bs.graph_capture(d_rays, d_hits, d_count, kernel=_abi.KERNEL_BVH2_PERSISTENT, block=256, refill=32)
    for rep in range(3):
        d_hits.fill_(-3.0)
        torch.cuda.synchronize()
        cap = len(rays) if rep != 1 else len(rays) // 3
        d_count.fill_(cap)
        bs.graph_launch()
        torch.cuda.synchronize()
        got = d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1)
        # work items are taken in batch order; the first `cap` of them are done
        order = np.concatenate([np.arange(o, o + c) for _, o, c in table])
        done, untouched = order[:cap], order[cap:]
        assert np.array_equal(rt_cases.hit_bits(got[done]), rt_cases.hit_bits(exp[done])), rep
        assert np.all(got["t"][untouched] == -3.0)
