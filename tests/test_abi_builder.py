"""CPU checks of the product library: the C ABI loads and exports every
declared symbol, error paths return status codes, and the host builder
reproduces the oracle's restatement of the engine builder bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import atlas_rt
import oracle_rt
import rt_cases
from atlas_rt import _abi
from atlas_rt import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "atlas_rt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(atlas_rt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(atlas_rt.lib_path())
    syms = declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(syms) <= set(_abi.PROTOTYPES), set(syms) - set(_abi.PROTOTYPES)


def test_version_and_errors_without_device():
    L = atlas_rt.lib()
    assert L.atlas_rt_version().startswith(b"atlas_rt")
    out = C.c_void_p()
    assert L.atlas_rt_blas_build(None, None, 10, 0, C.byref(out)) == _abi.E_INVALID_ARG
    assert b"NULL" in L.atlas_rt_last_error()
    tris = atlas_rt.as_bvh_triangles(np.zeros((1, 9), np.float32))
    assert L.atlas_rt_blas_build(None, tris.ctypes.data, 0, 0, C.byref(out)) == _abi.E_EMPTY
    assert L.atlas_rt_blas_build(None, tris.ctypes.data, 1, 0x80, C.byref(out)) == _abi.E_INVALID_ARG
    assert L.atlas_rt_blas_build(None, tris.ctypes.data, 1, 0, None) == _abi.E_INVALID_ARG


def test_device_trace_on_host_only_blas_is_an_error():
    """A host-only BLAS has no device arrays: the device entry points refuse
    it (before touching any buffer), while the engine's CPU traversal
    (atlas_rt_trace_closest_host) runs on its host copy."""
    verts, _ = rt_cases.soup(64, 1, 1)
    blas = atlas_rt.Blas.build(verts, host_only=True)
    rays = np.zeros(4, dtype=atlas_rt.RAY_DTYPE)
    hits = np.zeros(4, dtype=atlas_rt.HIT_DTYPE)
    L = atlas_rt.lib()
    st = L.atlas_rt_trace_closest_rays(blas.handle, rays.ctypes.data, hits.ctypes.data, None, 4, 0, None)
    assert st == _abi.E_INVALID_ARG and b"host-only" in L.atlas_rt_last_error()
    h = blas.trace_closest_host(rays)
    assert np.all(h["prim_id"] == -1)


def test_create_rejects_malformed_trees():
    nodes, tris, _ = rt_cases.chain_tree(4)
    bad = nodes.copy()
    bad[2]["left_ptr"] = 1  # points backwards: would loop
    with pytest.raises(atlas_rt.AtlasRTError) as e:
        atlas_rt.Blas.create(bad, tris)
    assert e.value.code == _abi.E_INVALID_ARG
    bad = nodes.copy()
    bad[0]["right_ptr"] = ~100  # leaf beyond the triangle array
    with pytest.raises(atlas_rt.AtlasRTError):
        atlas_rt.Blas.create(bad, tris)
    t2 = tris.copy()
    t2["v0"][:, 3] = -1.0  # no end-of-leaf marker anywhere
    with pytest.raises(atlas_rt.AtlasRTError):
        atlas_rt.Blas.create(nodes, t2)


def _same_as_oracle(verts, parallel):
    blas = atlas_rt.Blas.build(verts, parallel=parallel, host_only=True)
    nodes, tris, orig = blas.export()
    ob = oracle_rt.build_triangles(verts)
    assert nodes.tobytes() == ob.nodes.tobytes()
    assert tris.tobytes() == ob.tris.tobytes()
    assert np.array_equal(orig, ob.ref_idx)
    assert blas.info()[2] == (ob.depth if len(ob.nodes) else 1)


@pytest.mark.parametrize("n,salt", [(1, 1), (2, 2), (3, 3), (5, 4), (64, 5), (1000, 6), (4096, 7), (65536, 8)])
def test_builder_bit_identical_to_oracle(n, salt):
    verts, _ = rt_cases.soup(n, 1, salt)
    _same_as_oracle(verts, parallel=True)


def test_builder_serial_equals_parallel():
    verts, _ = rt_cases.soup(20000, 1, 9)
    a = atlas_rt.Blas.build(verts, parallel=False, host_only=True).export()
    b = atlas_rt.Blas.build(verts, parallel=True, host_only=True).export()
    assert all(x.tobytes() == y.tobytes() for x, y in zip(a, b))


def test_builder_heightfield_ties_identical():
    # a regular grid has many equal centroids/extents: exercises the median
    # fallback's std::sort and equal-cost bins
    _same_as_oracle(W.heightfield(grid=128), parallel=True)


def test_builder_degenerate_inputs_identical():
    g = np.random.default_rng(4)
    # all triangles at the same place (object split impossible -> median/sort)
    v = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (33, 1))
    _same_as_oracle(v, parallel=True)
    # zero-area triangles and a flat (extent < 1e-3) axis
    v = g.uniform(-5, 5, (500, 9)).astype(np.float32)
    v[:, 2::3] = 0.0
    v[:50, 3:6] = v[:50, 0:3]
    _same_as_oracle(v, parallel=True)


def test_builder_explicit_aabbs_match_oracle():
    verts, _ = rt_cases.soup(3000, 1, 10)
    tri = verts.reshape(-1, 3, 3)
    boxes = np.concatenate([tri.min(axis=1) - 0.01, tri.max(axis=1) + 0.01], axis=1).astype(np.float32)
    blas = atlas_rt.Blas.build(verts, aabbs=boxes, host_only=True)
    nodes, _, orig = blas.export()
    ob = oracle_rt.build_triangles(verts, aabbs=boxes)
    assert nodes.tobytes() == ob.nodes.tobytes() and np.array_equal(orig, ob.ref_idx)
