"""GPU parity on the coherent heightfield workload: primary closest-hit rays
from a 1920 x 1080 pinhole camera pitched 30 degrees down, then one any-hit
shadow ray per hit toward a directional light.

The heightfield's shared grid edges put many rays exactly on two triangles at
once, so this is the workload where equal-t ties happen.  The kernels follow
the engine's traversal order (HitClosest, data/shader/raytracer/bvh.hsh:191-273;
HitAny, bvh.hsh:359-441), so even a tie resolves to the engine's slot: the bar
is bit equality of slot, caller index, t, u, v and the occlusion flag, with
the tie count printed.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import workloads as W

pytestmark = pytest.mark.gpu

# A low sun (about 12 degrees above the horizon) shadows much of the terrain;
# the workload's own light direction is high and shadows little.
LOW_SUN = np.array([0.7, 0.21, 0.68]) / np.linalg.norm([0.7, 0.21, 0.68])


def _to_dev(torch, rays):
    return torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()


def _closest(torch, blas, d_rays, n, kernel):
    d_hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    d_slots = torch.empty(n, dtype=torch.int32, device="cuda")
    blas.trace_closest(d_rays, d_hits, d_slots, kernel=kernel)
    torch.cuda.synchronize()
    return d_hits.cpu().numpy().view(atlas_rt.HIT_DTYPE).reshape(-1), d_slots.cpu().numpy()


def _any(torch, blas, rays, kernel):
    d = _to_dev(torch, rays)
    out = torch.empty(len(rays), dtype=torch.int32, device="cuda")
    blas.trace_any(d, out, kernel=kernel)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.uint32)


def _check_heightfield(torch, grid):
    verts, prim = W.c1(grid=grid)
    blas = atlas_rt.Blas.build(verts)
    nodes, tris, orig = blas.export()
    href, sref, ties = oracle_rt.trace_closest(nodes, tris, orig, prim, ties=True)
    n_hits = int((sref >= 0).sum())
    print(f"c1 grid {grid}: {len(prim)} primary rays, {n_hits} hits, {int(ties.sum())} equal-t ties", flush=True)
    assert n_hits > len(prim) // 4
    d_prim = _to_dev(torch, prim)
    # grid kernel (the engine's dispatch shape), resident leaf-batched waves,
    # and AUTO (what an unadorned call picks); twice each, so the persistent
    # kernel's self-resetting counter is exercised too.
    for kernel in (1, 2 | 0x40, 0, 2 | 0x50):
        for _ in range(2):
            h, s = _closest(torch, blas, d_prim, len(prim), kernel)
            bad = np.nonzero(np.any(rt_cases.hit_bits(h) != rt_cases.hit_bits(href), axis=1) | (s != sref))[0]
            assert len(bad) == 0, (f"grid {grid} kernel {kernel}: {len(bad)} rays differ, first {bad[:5]}, "
                                   f"of which {int(ties[bad].sum())} are ties")
    for light in (W.C1_LIGHT_DIR, LOW_SUN):
        shadow = W.shadow_rays(prim, href, verts, light_dir=light)
        ref = oracle_rt.trace_any(nodes, tris, shadow)
        for kernel in (1, 2, 0):
            got = _any(torch, blas, shadow, kernel)
            assert np.array_equal(got, ref), (grid, kernel, int(np.sum(got != ref)))
        print(f"c1 grid {grid}: shadow rays toward {np.round(light, 3)}: {int(ref.sum())} of {n_hits} occluded",
              flush=True)
    return ties


def test_c1_heightfield_256(torch_cuda):
    """A 256 x 256 heightfield (130,050 triangles) under the full 1920 x 1080 ray set."""
    print("case c1 grid 256", flush=True)
    _check_heightfield(torch_cuda, 256)


def test_c1_heightfield_full(torch_cuda):
    """The full workload: 1024 x 1024 grid (2,093,058 triangles), 2,073,600 primary rays."""
    print("case c1 grid 1024", flush=True)
    _check_heightfield(torch_cuda, 1024)
