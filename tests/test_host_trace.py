"""The engine's CPU traversal in the product (atlas_rt_trace_closest_host /
_any_host: BVH::GetIntersection / GetIntersectionAny, BVH.cpp:103-217, over
Ray::Intersects, Ray.cpp:28-76) against the oracle's R-CPU restatement, on the
host with host-only BLASes.

Bar: the hit triangle (caller index) and the FP32 hit distance bit for bit,
and any-hit flags equal.  The oracle's R-CPU is pinned separately
(test_oracle.py): with tmin = 0 it agrees with the R-GPU walk bit for bit,
and with tmin > 0 it differs only by accepting hits in [0, tmin].  Here the
product's u, v (the engine's FP32 solution) must lie within a first-order
FP32 rounding bound of the oracle's FP64 refinement of the same triangle (a swapped u / v or a wrong
divisor fails that) and reconstruct the hit point o + t d.
"""
import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import workloads as W


def _check(verts, rays):
    blas = atlas_rt.Blas.build(verts, host_only=True)
    nodes, tris, orig = blas.export()
    got = blas.trace_closest_host(rays)
    ref, slots, raw_t = oracle_rt.trace_closest(nodes, tris, orig, rays, mode=oracle_rt.MODE_RCPU, raw_t=True)
    assert np.array_equal(got["prim_id"], ref["prim_id"])
    hit = ref["prim_id"] >= 0
    assert np.array_equal(got["t"][hit].view(np.uint32), raw_t[hit].view(np.uint32))
    assert np.array_equal(got["t"][~hit].view(np.uint32), rays["tmax"][~hit].view(np.uint32))
    assert np.all(got["u"][~hit] == 0) and np.all(got["v"][~hit] == 0)
    # The FP32 solution against the FP64 refinement of the same triangle,
    # within a first-order bound of FP32 rounding in (q.s, p.d) / q.e0.
    tri = np.asarray(verts, np.float64).reshape(-1, 3, 3)[got["prim_id"][hit]]
    o = rays["origin"][hit].astype(np.float64)
    d = rays["direction"][hit].astype(np.float64)
    e0, e1, sv = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], o - tri[:, 0]
    pv, qv = np.cross(sv, e0), np.cross(d, e1)
    den = np.abs(np.sum(qv * e0, axis=1))
    nrm = lambda a: np.linalg.norm(a, axis=1)
    eps = 2.0 ** -24
    for name, num in (("u", nrm(qv) * nrm(sv)), ("v", nrm(pv) * nrm(d))):
        bound = 64 * eps * (num + np.abs(ref[name][hit]) * nrm(qv) * nrm(e0)) / den + 1e-6
        err = np.abs(got[name][hit].astype(np.float64) - ref[name][hit])
        assert np.all(err <= bound), (name, err.max(), np.argmax(err / bound))
    # and they name the hit point: o + t d = v0 + u e0 + v e1
    p_bar = tri[:, 0] + got["u"][hit, None] * e0 + got["v"][hit, None] * e1
    assert np.median(np.linalg.norm(o + got["t"][hit, None] * d - p_bar, axis=1)) < 1e-4
    any_got = blas.trace_any_host(rays)
    any_ref = oracle_rt.trace_any(nodes, tris, rays, mode=oracle_rt.MODE_RCPU)
    assert np.array_equal(any_got, any_ref)
    # an any hit exists exactly where a closest hit does (both accept [0, tMax))
    assert np.array_equal(any_got.astype(bool), hit)
    blas.close()
    return hit.mean()


@pytest.mark.parametrize("n_tris,n_rays,salt", [(1, 2048, 1), (2, 2048, 2), (7, 4096, 3), (4096, 65536, 4),
                                                (65536, 100000, 5)])
def test_host_rcpu_equals_oracle_soup(n_tris, n_rays, salt):
    verts, rays = rt_cases.soup(n_tris, n_rays, salt)
    if n_tris <= 7:  # aim some rays at the few triangles
        c = verts.reshape(-1, 3, 3).mean(axis=1)
        g = np.random.default_rng(salt)
        tgt = c[g.integers(0, len(c), n_rays // 2)]
        d = tgt - rays["origin"][: n_rays // 2]
        rays["direction"][: n_rays // 2] = d / np.linalg.norm(d, axis=1, keepdims=True)
    frac = _check(verts, rays)
    assert frac > 0.0


def test_host_rcpu_equals_oracle_heightfield_ties_and_tmin():
    """Grid edges give exact equal-t ties; tmin = 0.5 exercises the engine's
    tmin-free triangle acceptance (a hit closer than tmin is still taken when
    its boxes pass the slab test)."""
    verts = W.heightfield(grid=64, amplitude=6.0, seed=3)
    rays = W.pinhole_rays(160, 90, (32.0, 30.0, -20.0), 0.0, 40.0, 60.0)
    assert _check(verts, rays) > 0.2
    rays["tmin"] = 0.5
    _check(verts, rays)


def test_host_rcpu_edge_cases():
    verts, _ = rt_cases.soup(256, 8, 9)
    c = verts.reshape(-1, 3, 3).mean(axis=1)
    o = np.array([[0.0, 0.0, -30.0]] * 4, np.float32)
    d = np.array([[np.nan, 0, 1], [0, 0, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    rays = rt_cases.make_rays(o, d, tmax=1e12)
    tgt = rt_cases.make_rays(np.zeros((len(c), 3)), c / np.linalg.norm(c, axis=1, keepdims=True), tmax=1e12)
    short = tgt.copy()
    short["tmax"] = 1e-3  # nothing within reach
    _check(verts, np.concatenate([rays, tgt, short]))
    blas = atlas_rt.Blas.build(verts, host_only=True)
    h = blas.trace_closest_host(short)
    assert np.all(h["prim_id"] == -1)
    assert len(blas.trace_closest_host(short[:0])) == 0
    blas.close()
