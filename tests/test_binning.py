"""Ray coherence sorting (direction bins).

CPU: the oracle's bin function pinned by the axis directions in closed form
and by the inverse octahedral map (every point strictly inside a cell decodes
to a direction the oracle puts back in that cell, at any length); its sort
pinned by numpy's stable argsort of the bins.
GPU: atlas_rt_sort_rays_by_direction equals the oracle's sort (rays, permutation,
bin offsets) bit for bit; the path tracer renders the same image, bit for bit,
with and without binning.

Reference: DetermineRayBin / AddRayToBin (data/shader/raytracer/tracing.hsh:18-29),
UnitVectorToOctahedron (data/shader/common/octahedron.hsh:23-37), binning.csh,
binningOffset.csh, RayTracingHelper.cpp:304-344.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import pt_cases
import rt_cases
from atlas_rt import workloads as W

ENV = (0.25, 0.3, 0.4)


def test_oracle_bins_of_the_axes():
    """+y maps to the centre (4, 4); -y to the folded corner (1, 1), which the
    engine truncates to cell 8 and the build clamps to 7; +-x and +-z to the
    edge midpoints.  NaN and zero directions fall in bin 0."""
    d = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1],
                  [np.nan, 0, 0], [0, 0, 0]], np.float32)
    assert list(oracle_rt.ray_bins(d)) == [4 * 8 + 4, 7 * 8 + 7, 4 * 8 + 7, 4 * 8 + 0, 7 * 8 + 4, 0 * 8 + 4, 0, 0]


def _octahedron_to_unit_vector(u, v):
    """Inverse octahedral map (octahedron.hsh:39-53): the square [0, 1]^2 onto
    the sphere, lower hemisphere folded over the diagonals."""
    cx, cz = 2.0 * u - 1.0, 2.0 * v - 1.0
    y = 1.0 - np.abs(cx) - np.abs(cz)
    lower = y < 0.0
    fx = np.where(cx >= 0.0, 1.0, -1.0) * (1.0 - np.abs(cz))
    fz = np.where(cz >= 0.0, 1.0, -1.0) * (1.0 - np.abs(cx))
    x = np.where(lower, fx, cx)
    z = np.where(lower, fz, cz)
    d = np.stack([x, y, z], axis=-1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_oracle_bin_inverts_the_octahedral_map():
    g = np.random.default_rng(3)
    n = 64 * 200
    cell = np.repeat(np.arange(64), 200)
    cx, cy = cell % 8, cell // 8
    # strictly inside the cell, away from float rounding at its edges
    u = (cx + g.uniform(0.01, 0.99, n)) / 8.0
    v = (cy + g.uniform(0.01, 0.99, n)) / 8.0
    d = _octahedron_to_unit_vector(u, v) * g.uniform(0.01, 100.0, (n, 1))
    assert np.array_equal(oracle_rt.ray_bins(d.astype(np.float32)), cell.astype(np.uint32))


def _rays(n, seed, specials=True):
    _, rays = rt_cases.soup(4, n, seed)
    rays = rays.copy()
    if specials and n >= 16:
        rays["direction"][:8] = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1],
                                 [np.nan, 0, 0], [0, 0, 0]]
        # coherent runs: many rays of one bin in a row
        rays["direction"][8:16] = rays["direction"][8]
    return rays


@pytest.mark.parametrize("n", [0, 1, 17, 5000, 70001])
def test_oracle_sort_is_the_stable_sort_of_the_bins(n):
    rays = _rays(n, 40 + n)
    out, perm, offsets = oracle_rt.sort_rays_by_direction(rays)
    bins = oracle_rt.ray_bins(rays["direction"])
    ref = np.argsort(bins, kind="stable")
    assert np.array_equal(perm, ref)
    assert out.tobytes() == rays[ref].tobytes()
    assert np.array_equal(offsets, np.searchsorted(bins[ref], np.arange(65), side="left"))
    assert offsets[64] == n


def _gpu_sort(torch, rays, with_perm=True):
    n = len(rays)
    d_in = torch.from_numpy(rays.view(np.float32).reshape(-1, 8).copy()).cuda()
    d_out = torch.full((max(n, 1), 8), -1.0, dtype=torch.float32, device="cuda")
    d_perm = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda") if with_perm else None
    d_off = torch.full((65,), -1, dtype=torch.int32, device="cuda")
    atlas_rt.sort_rays_by_direction(d_in, d_out, d_perm, d_off, count=n)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()[:n].copy().view(oracle_rt.RAY_DTYPE).reshape(-1)
    perm = d_perm.cpu().numpy()[:n].astype(np.uint32) if with_perm else None
    return out, perm, d_off.cpu().numpy().astype(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 17, 4096, 4097, 70001, 1 << 20])
def test_gpu_sort_equals_oracle(torch_cuda, n):
    print(f"case ray sort n={n}", flush=True)
    rays = _rays(n, 50 + n)
    out, perm, offsets = _gpu_sort(torch_cuda, rays)
    r_out, r_perm, r_off = oracle_rt.sort_rays_by_direction(rays)
    assert out.tobytes() == r_out.tobytes()
    assert np.array_equal(perm, r_perm)
    assert np.array_equal(offsets, r_off)
    if n:
        out2, _, _ = _gpu_sort(torch_cuda, rays, with_perm=False)
        assert out2.tobytes() == r_out.tobytes()


@pytest.mark.gpu
def test_gpu_pt_binning_leaves_the_image_unchanged(torch_cuda):
    """Binning reorders each bounce's paths; every path keeps its pixel, RNG
    stream and accumulation order, so the image is the same bit for bit."""
    print("case pt binning", flush=True)
    meshes, normals, mats, xf, which = pt_cases.small_c2()
    S = pt_cases.oracle_scene(meshes, normals, mats, xf, which)
    w_, blases = pt_cases.product_scene(meshes, normals, xf, which, S["boxes"])
    W_, H_, bounces, frames = 96, 64, 4, 3
    cam = atlas_rt.camera((0.0, 15.0, -70.0), 0.0, 12.0, 50.0, W_ / H_)
    imgs = []
    for binning in (False, True):
        pt = atlas_rt.PathTracer(w_, W_, H_, bounces, mats, W.sun_light(), env=ENV, binning=binning)
        pt.render(cam, frames)
        imgs.append(pt.read().cpu().numpy())
        pt.close()
    assert imgs[0].tobytes() == imgs[1].tobytes()
    assert imgs[0][..., :3].max() > 0.2
