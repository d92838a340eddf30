"""The closest-hit entry point that allocates nothing
(atlas_rt_trace_closest_rays_noalloc) against atlas_rt_trace_closest_rays with
the same kernel: bit-identical hits and slots, two calls in a row (the second
with fewer rays, on the scratch the first one left), for each persistent
variant; kernels it does not run are refused before anything is launched."""
import numpy as np
import pytest

import atlas_rt
import rt_cases
from atlas_rt import _abi

pytestmark = pytest.mark.gpu

P = _abi.KERNEL_BVH2_PERSISTENT
VARIANTS = [_abi.KERNEL_AUTO, P, P | _abi.KERNEL_LEAF_BATCH, P | _abi.KERNEL_STACK_REGISTER,
            P | _abi.KERNEL_SLAB_FLOAT | _abi.KERNEL_LEAF_BATCH]


def _trace(torch, blas, rays, kernel, noalloc):
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    n = len(rays)
    d_hits = torch.full((n, 4), 7.0, dtype=torch.float32, device="cuda")
    d_slots = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    if noalloc:
        blas.trace_closest_noalloc(d_rays, d_hits, d_slots, kernel=kernel)
    else:
        blas.trace_closest(d_rays, d_hits, d_slots,
                           kernel=(P | _abi.KERNEL_LEAF_BATCH) if kernel == _abi.KERNEL_AUTO else kernel)
    torch.cuda.synchronize()
    return d_hits.cpu().numpy().view(np.uint32).copy(), d_slots.cpu().numpy().copy()


@pytest.mark.parametrize("kernel", VARIANTS)
def test_noalloc_matches_the_allocating_trace(torch_cuda, kernel):
    torch = torch_cuda
    verts, rays = rt_cases.soup(16384, 200000, 31)
    blas = atlas_rt.Blas.build(verts)
    for n in (200000, 37001):  # the second call reuses the scratch with fewer rays
        print(f"case noalloc kernel={kernel:#x} rays={n}", flush=True)
        h_ref, s_ref = _trace(torch, blas, rays[:n], kernel, noalloc=False)
        h, s = _trace(torch, blas, rays[:n], kernel, noalloc=True)
        assert np.array_equal(h, h_ref), f"hits differ for kernel {kernel:#x}, {n} rays"
        assert np.array_equal(s, s_ref), f"slots differ for kernel {kernel:#x}, {n} rays"
        assert (s >= 0).sum() > n // 20  # the soup is hit
    blas.close()


def test_noalloc_refuses_other_kernels(torch_cuda):
    torch = torch_cuda
    verts, rays = rt_cases.soup(64, 128, 32)
    blas = atlas_rt.Blas.build(verts)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    for k in (_abi.KERNEL_BVH2_EXACT, _abi.KERNEL_BVH4, _abi.KERNEL_BVH8 | _abi.KERNEL_STACK_REGISTER):
        with pytest.raises(atlas_rt.AtlasRTError) as e:
            blas.trace_closest_noalloc(d_rays, d_hits, kernel=k)
        assert e.value.code == _abi.E_UNSUPPORTED
    blas.close()
