"""TEST INFRASTRUCTURE - the oracle (ctypes bindings of liboracle_rt.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  It restates the reference
algorithms (src/engine/volume/BVH.cpp, Ray.cpp, AABB.cpp,
data/shader/raytracer/bvh.hsh, intersections.hsh) on the CPU; see the C++
sources in this directory for file:line citations.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NODE_DTYPE = np.dtype([
    ("left_min", "<f4", 3), ("left_max", "<f4", 3),
    ("right_min", "<f4", 3), ("right_max", "<f4", 3),
    ("left_ptr", "<i4"), ("right_ptr", "<i4"), ("padding", "<i4", 2),
])
GPU_TRI_DTYPE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4)])
INSTANCE_DTYPE = np.dtype([
    ("inverse_matrix", "<f4", (3, 4)), ("mesh_offset", "<i4"), ("material_offset", "<i4"),
    ("next_instance", "<i4"), ("mask", "<u4"),
])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim_id", "<i4")])

MODE_RGPU, MODE_RCPU, MODE_BRUTE = 0, 1, 2


def lib_path():
    return os.path.join(_HERE, "liboracle_rt.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(lib_path()):
            raise RuntimeError(f"{lib_path()} is not built (make -C oracle)")
        L = C.CDLL(lib_path())
        vp, u32, i32, f32 = C.c_void_p, C.c_uint32, C.c_int32, C.c_float
        L.oracle_bvh_build_tris.restype = vp
        L.oracle_bvh_build_tris.argtypes = [vp, vp, u32]
        L.oracle_bvh_build_aabbs.restype = vp
        L.oracle_bvh_build_aabbs.argtypes = [vp, u32]
        L.oracle_bvh_node_count.restype = u32
        L.oracle_bvh_node_count.argtypes = [vp]
        L.oracle_bvh_ref_count.restype = u32
        L.oracle_bvh_ref_count.argtypes = [vp]
        L.oracle_bvh_export.restype = None
        L.oracle_bvh_export.argtypes = [vp, vp, vp, vp, vp, vp, i32, f32]
        L.oracle_bvh_free.restype = None
        L.oracle_bvh_free.argtypes = [vp]
        L.oracle_bvh_depth.restype = u32
        L.oracle_bvh_depth.argtypes = [vp, u32]
        L.oracle_trace_closest.restype = None
        L.oracle_trace_closest.argtypes = [vp, u32, vp, vp, u32, vp, u32, vp, vp, vp, vp, vp, vp, i32, i32]
        L.oracle_trace_any.restype = None
        L.oracle_trace_any.argtypes = [vp, u32, vp, u32, vp, u32, vp, vp, vp, i32, i32]
        L.oracle_trace_world.restype = None
        L.oracle_trace_world.argtypes = [vp, u32, vp, u32, vp, vp, vp, vp, vp, u32, vp, u32, u32, i32,
                                         vp, vp, vp, vp, vp, vp, vp, vp, i32]
        L.oracle_triangle_test.restype = i32
        L.oracle_triangle_test.argtypes = [vp, vp, vp, vp, vp, vp]
        L.oracle_slab_gpu.restype = i32
        L.oracle_slab_gpu.argtypes = [vp, vp, vp, vp, f32, f32, vp]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


def threads_default():
    return os.cpu_count() or 1


class OracleBVH:
    """Result of the restated reference builder, packed like MeshData::BuildBVH."""

    def __init__(self, nodes, tris, ref_idx, end_of_node):
        self.nodes = nodes
        self.tris = tris
        self.ref_idx = ref_idx
        self.end_of_node = end_of_node

    @property
    def depth(self):
        return int(lib().oracle_bvh_depth(_p(self.nodes), len(self.nodes)))


def build_triangles(verts, aabbs=None, material_idx=0, opacity=1.0):
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
    if aabbs is not None:
        aabbs = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    h = lib().oracle_bvh_build_tris(_p(aabbs), _p(verts), len(verts))
    if not h:
        raise ValueError("oracle build failed")
    try:
        n = lib().oracle_bvh_node_count(h)
        r = lib().oracle_bvh_ref_count(h)
        nodes = np.zeros(n, dtype=NODE_DTYPE)
        ref = np.zeros(r, dtype=np.uint32)
        eon = np.zeros(r, dtype=np.uint8)
        tris = np.zeros(r, dtype=GPU_TRI_DTYPE)
        lib().oracle_bvh_export(h, _p(nodes) if n else None, _p(ref), _p(eon), _p(verts), _p(tris),
                                material_idx, opacity)
    finally:
        lib().oracle_bvh_free(h)
    return OracleBVH(nodes, tris, ref, eon)


def build_aabbs(aabbs):
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    h = lib().oracle_bvh_build_aabbs(_p(aabbs), len(aabbs))
    if not h:
        raise ValueError("oracle build failed")
    try:
        n = lib().oracle_bvh_node_count(h)
        r = lib().oracle_bvh_ref_count(h)
        nodes = np.zeros(n, dtype=NODE_DTYPE)
        ref = np.zeros(r, dtype=np.uint32)
        eon = np.zeros(r, dtype=np.uint8)
        lib().oracle_bvh_export(h, _p(nodes) if n else None, _p(ref), _p(eon), None, None, 0, 0.0)
    finally:
        lib().oracle_bvh_free(h)
    return OracleBVH(nodes, None, ref, eon)


def trace_closest(bvh_nodes, bvh_tris, orig_idx, rays, mode=MODE_RGPU, threads=None, stats=False, raw_t=False,
                  ties=False):
    """Closest hit on the packed arrays.

    Returns (hits, slots[, raw_t][, node_visits, tri_tests][, ties]).  hits
    carry the refined (t, u, v) of the selected triangle; raw_t is the
    engine's float32 hit distance that drove the traversal; ties (R-GPU mode)
    flags rays whose hit had a rival triangle at exactly the same float32 t,
    so the traversal order chose the winner.
    """
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    bvh_nodes = np.ascontiguousarray(bvh_nodes, dtype=NODE_DTYPE)
    bvh_tris = np.ascontiguousarray(bvh_tris, dtype=GPU_TRI_DTYPE)
    n = len(rays)
    hits = np.zeros(n, dtype=HIT_DTYPE)
    slots = np.zeros(n, dtype=np.int32)
    rt = np.zeros(n, dtype=np.float32) if raw_t else None
    nv = np.zeros(n, dtype=np.uint32) if stats else None
    tt = np.zeros(n, dtype=np.uint32) if stats else None
    tie = np.zeros(n, dtype=np.uint8) if ties else None
    orig = None if orig_idx is None else np.ascontiguousarray(orig_idx, dtype=np.uint32)
    lib().oracle_trace_closest(_p(bvh_nodes) if len(bvh_nodes) else None, len(bvh_nodes), _p(bvh_tris), _p(orig),
                               len(bvh_tris), _p(rays), n, _p(hits), _p(slots), _p(rt), _p(nv), _p(tt), _p(tie),
                               mode, threads or threads_default())
    out = (hits, slots)
    if raw_t:
        out += (rt,)
    if stats:
        out += (nv, tt)
    if ties:
        out += (tie.astype(bool),)
    return out


def trace_any(bvh_nodes, bvh_tris, rays, mode=MODE_RGPU, threads=None, stats=False):
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = len(rays)
    hit = np.zeros(n, dtype=np.uint32)
    nv = np.zeros(n, dtype=np.uint32) if stats else None
    tt = np.zeros(n, dtype=np.uint32) if stats else None
    lib().oracle_trace_any(_p(bvh_nodes) if len(bvh_nodes) else None, len(bvh_nodes), _p(bvh_tris), len(bvh_tris),
                           _p(rays), n, _p(hit), _p(nv), _p(tt), mode, threads or threads_default())
    if stats:
        return hit, nv, tt
    return hit


WORLD_CLOSEST, WORLD_ANY, WORLD_BRUTE, WORLD_ANY_RESTORE = 0, 1, 2, 3


def trace_world(tlas_nodes, instances, blases, rays, mask=1 << 7, any_hit=False, threads=None, stats=False,
                mode=None, raw_t=False, ties=False):
    """Two-level trace.  blases: list of (nodes, tris, orig_idx) packed arrays.

    mode: WORLD_CLOSEST (R-GPU HitClosest), WORLD_ANY (HitAny; hits.prim_id =
    1/0) or WORLD_BRUTE (every instance x triangle).  Returns (hits, slots,
    instance slots[, raw_t][, node_visits, tri_tests, instance_visits][, ties]).
    """
    if mode is None:
        mode = WORLD_ANY if any_hit else WORLD_CLOSEST
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = len(rays)
    k = len(blases)
    keep = []
    node_ptrs = (C.c_void_p * k)()
    tri_ptrs = (C.c_void_p * k)()
    orig_ptrs = (C.c_void_p * k)()
    node_counts = np.zeros(k, dtype=np.uint32)
    tri_counts = np.zeros(k, dtype=np.uint32)
    for i, (bn, bt, bo) in enumerate(blases):
        bn = np.ascontiguousarray(bn, dtype=NODE_DTYPE)
        bt = np.ascontiguousarray(bt, dtype=GPU_TRI_DTYPE)
        bo = np.ascontiguousarray(bo, dtype=np.uint32)
        keep += [bn, bt, bo]
        node_ptrs[i] = bn.ctypes.data if len(bn) else None
        tri_ptrs[i] = bt.ctypes.data
        orig_ptrs[i] = bo.ctypes.data
        node_counts[i] = len(bn)
        tri_counts[i] = len(bt)
    tlas_nodes = np.ascontiguousarray(tlas_nodes, dtype=NODE_DTYPE)
    instances = np.ascontiguousarray(instances, dtype=INSTANCE_DTYPE)
    hits = np.zeros(n, dtype=HIT_DTYPE)
    slots = np.zeros(n, dtype=np.int32)
    inst = np.zeros(n, dtype=np.int32)
    rt = np.zeros(n, dtype=np.float32) if raw_t else None
    nv = np.zeros(n, dtype=np.uint32) if stats else None
    tt = np.zeros(n, dtype=np.uint32) if stats else None
    iv = np.zeros(n, dtype=np.uint32) if stats else None
    tie = np.zeros(n, dtype=np.uint8) if ties else None
    lib().oracle_trace_world(_p(tlas_nodes), len(tlas_nodes), _p(instances), len(instances),
                             C.cast(node_ptrs, C.c_void_p), _p(node_counts), C.cast(tri_ptrs, C.c_void_p),
                             C.cast(orig_ptrs, C.c_void_p), _p(tri_counts), k, _p(rays), n, mask,
                             mode, _p(hits), _p(slots), _p(inst), _p(rt), _p(nv), _p(tt), _p(iv), _p(tie),
                             threads or threads_default())
    out = (hits, slots, inst)
    if raw_t:
        out += (rt,)
    if stats:
        out += (nv, tt, iv)
    if ties:
        out += (tie.astype(bool),)
    return out


def _f32(*xs):
    return [np.ascontiguousarray(x, dtype=np.float32) for x in xs]


def _scene_args(tlas_nodes, instances, xform, blases, shading, materials, lights, env):
    """ctypes argument list of the scene part shared by the path tracers, and
    the arrays it points into (keep them alive across the call)."""
    k = len(blases)
    keep = []
    node_ptrs, tri_ptrs = (C.c_void_p * k)(), (C.c_void_p * k)()
    nrm_ptrs, mat_ptrs = (C.c_void_p * k)(), (C.c_void_p * k)()
    node_counts = np.zeros(k, dtype=np.uint32)
    tri_counts = np.zeros(k, dtype=np.uint32)
    for i, ((bn, bt), (sn, sm)) in enumerate(zip(blases, shading)):
        bn = np.ascontiguousarray(bn, dtype=NODE_DTYPE)
        bt = np.ascontiguousarray(bt, dtype=GPU_TRI_DTYPE)
        sn = np.ascontiguousarray(sn, dtype=np.float32)
        sm = np.ascontiguousarray(sm, dtype=np.int32)
        keep += [bn, bt, sn, sm]
        node_ptrs[i] = bn.ctypes.data if len(bn) else None
        tri_ptrs[i] = bt.ctypes.data
        nrm_ptrs[i] = sn.ctypes.data
        mat_ptrs[i] = sm.ctypes.data
        node_counts[i], tri_counts[i] = len(bn), len(bt)
    tlas_nodes = np.ascontiguousarray(tlas_nodes, dtype=NODE_DTYPE)
    instances = np.ascontiguousarray(instances, dtype=INSTANCE_DTYPE)
    xform = np.ascontiguousarray(xform, dtype=np.float32)
    materials = np.ascontiguousarray(materials)
    lights = np.ascontiguousarray(lights)
    env = np.ascontiguousarray(env, dtype=np.float32)
    keep += [node_ptrs, tri_ptrs, nrm_ptrs, mat_ptrs, node_counts, tri_counts, tlas_nodes, instances, xform,
             materials, lights, env]
    args = [_p(tlas_nodes), len(tlas_nodes), _p(instances), len(instances), C.cast(node_ptrs, C.c_void_p),
            _p(node_counts), C.cast(tri_ptrs, C.c_void_p), _p(tri_counts), C.cast(nrm_ptrs, C.c_void_p),
            C.cast(mat_ptrs, C.c_void_p), k, _p(xform), _p(materials), len(materials),
            _p(lights) if len(lights) else None, len(lights), _p(env)]
    return args, keep


class _Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.c_void_p), ("channels", C.c_uint32),
                ("reserved", C.c_uint32)]


class _Surface(C.Structure):
    _fields_ = [("opacity", C.c_void_p), ("uvs", C.c_void_p), ("textures", C.c_void_p),
                ("texture_count", C.c_uint32), ("emitters", C.c_void_p), ("emitter_count", C.c_uint32),
                ("lights_out", C.c_void_p), ("lights_cap", C.c_uint32), ("lights_count", C.c_void_p),
                ("tangents", C.c_void_p), ("colors", C.c_void_p)]


GPU_LIGHT_DTYPE = np.dtype([("P", "<f4", 4), ("N", "<f4", 4), ("color", "<f4", 4), ("data", "<f4", 4)])


def _surface_arg(surface, n_blas, keep):
    """OracleSurface from a dict with optional keys: opacity (per BLAS, slot
    order, or None), uvs (per BLAS, (n, 6) slot order, or None), tangents
    (per BLAS, (n, 6) slot order, or None), colors (per BLAS, (n, 12) vertex
    colours RGBA in slot order, or None), textures (list of (h, w) or
    (h, w, channels) uint8), emitters ((m, 4) float32), lights_cap (int: the
    light list is returned in surface["lights"])."""
    if surface is None:
        return None
    S = _Surface()
    def per_blas(key, width):
        vals = surface.get(key)
        if vals is None:
            return None
        ptrs = (C.c_void_p * n_blas)()
        for i, a in enumerate(vals):
            if a is None:
                continue
            a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, width)
            keep.append(a)
            ptrs[i] = a.ctypes.data
        keep.append(ptrs)
        return C.cast(ptrs, C.c_void_p)
    S.opacity = per_blas("opacity", 1)
    S.uvs = per_blas("uvs", 6)
    S.tangents = per_blas("tangents", 6)
    S.colors = per_blas("colors", 12)
    tex = surface.get("textures") or []
    if tex:
        arr = (_Texture * len(tex))()
        for i, t in enumerate(tex):
            t = np.ascontiguousarray(t, dtype=np.uint8)
            keep.append(t)
            arr[i].height, arr[i].width = t.shape[:2]
            arr[i].channels = t.shape[2] if t.ndim == 3 else 1
            arr[i].texels = t.ctypes.data
        keep.append(arr)
        S.textures = C.cast(arr, C.c_void_p)
        S.texture_count = len(tex)
    em = surface.get("emitters")
    if em is not None:
        em = np.ascontiguousarray(em, dtype=np.float32).reshape(-1, 4)
        keep.append(em)
        S.emitters = em.ctypes.data
        S.emitter_count = len(em)
    cap = int(surface.get("lights_cap", 0))
    if cap:
        out = np.zeros(cap, GPU_LIGHT_DTYPE)
        cnt = np.zeros(1, np.uint32)
        keep += [out, cnt]
        S.lights_out, S.lights_cap, S.lights_count = out.ctypes.data, cap, cnt.ctypes.data
        surface["_lights"] = (out, cnt)
    keep.append(S)
    return C.addressof(S)


def _collect_lights(surface):
    if surface is not None and "_lights" in surface:
        out, cnt = surface.pop("_lights")
        surface["lights"] = out[:min(int(cnt[0]), len(out))].copy()


_SCENE_TYPES = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                C.c_void_p]


def pathtrace(tlas_nodes, instances, xform, blases, shading, materials, lights, env, cam, width, height, bounces,
              frames, first_sample=0, pcg_draws=0, accum=None, threads=None, surface=None):
    """CPU restatement of the progressive path tracer.

    blases: list of (nodes, tris) packed arrays; shading: list of (normals
    (n, 9) float32, material (n,) int32) in BVH slot order; xform: (m, 12)
    object-to-world rows per packed instance; cam: a ctypes camera struct
    (location, origin, right, bottom); surface: optional opacity, texture and
    triangle-light inputs (see _surface_arg).  Returns the accumulated
    (h, w, 4) float32 image (rgb sum, sample count).
    """
    args, keep = _scene_args(tlas_nodes, instances, xform, blases, shading, materials, lights, env)
    ext = _surface_arg(surface, len(blases), keep)
    if accum is None:
        accum = np.zeros((height, width, 4), dtype=np.float32)
    L = lib()
    if not getattr(L, "_pt_bound", False):
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
        L.oracle_pathtrace.restype = None
        L.oracle_pathtrace.argtypes = _SCENE_TYPES + [vp, u32, u32, i32, u32, u32, C.c_uint64, vp, i32, vp]
        L._pt_bound = True
    L.oracle_pathtrace(*args, C.addressof(cam), width, height, bounces, frames, first_sample, pcg_draws, _p(accum),
                       threads or threads_default(), ext)
    _collect_lights(surface)
    return accum


def pathtrace_realtime(tlas_nodes, instances, xform, blases, shading, materials, lights, env, cam, width, height,
                       bounces, spf, matrices, last_xform, max_radiance=65535.0, pcg_draws=0, threads=None,
                       surface=None):
    """One realtime frame: (frame (h, w, 4) radiance as the temporal pass
    fetches it, depth (h, w), velocity (h, w) uint32 binary16 pairs, normal
    (h, w) uint32 binary16 pairs, material (h, w) uint16).  matrices: a
    ctypes atlas_rt_frame_matrices; last_xform: (m, 12) last frame's rows."""
    args, keep = _scene_args(tlas_nodes, instances, xform, blases, shading, materials, lights, env)
    ext = _surface_arg(surface, len(blases), keep)
    L = lib()
    if not getattr(L, "_ptr_bound", False):
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
        L.oracle_pathtrace_realtime.restype = None
        L.oracle_pathtrace_realtime.argtypes = _SCENE_TYPES + [vp, u32, u32, i32, u32, C.c_uint64, vp, vp, vp, vp, vp,
                                                               C.c_float, vp, vp, vp, vp, vp, i32, vp]
        L._ptr_bound = True
    frame = np.zeros((height, width, 4), np.float32)
    depth = np.zeros((height, width), np.float32)
    vel = np.zeros((height, width), np.uint32)
    nrm = np.zeros((height, width), np.uint32)
    mat = np.zeros((height, width), np.uint16)
    last = np.ascontiguousarray(last_xform, dtype=np.float32)
    m = matrices
    L.oracle_pathtrace_realtime(*args, C.addressof(cam), width, height, bounces, spf, pcg_draws,
                                C.addressof(m.view), C.addressof(m.projection), C.addressof(m.last_view_projection),
                                C.addressof(m.jitter), _p(last), max_radiance, _p(frame), _p(depth), _p(vel), _p(nrm),
                                _p(mat), threads or threads_default(), ext)
    _collect_lights(surface)
    return frame, depth, vel, nrm, mat


def temporal(frame, depth, velocity, normal, material, hdepth, hnormal, hmaterial, hradiance,
             history_clip_max=0.2, current_clip_factor=2.0, max_history_length=32.0):
    """The temporal resolve restated: (h, w, 4) (resolved rgb, history length)."""
    h, w = depth.shape
    L = lib()
    if not getattr(L, "_tmp_bound", False):
        L.oracle_temporal.restype = None
        L.oracle_temporal.argtypes = [C.c_uint32, C.c_uint32] + [C.c_void_p] * 11
        L._tmp_bound = True
    arrs = [np.ascontiguousarray(a, dtype=t) for a, t in
            ((frame, np.float32), (depth, np.float32), (velocity, np.uint32), (normal, np.uint32),
             (material, np.uint16), (hdepth, np.float32), (hnormal, np.uint32), (hmaterial, np.uint16),
             (hradiance, np.float32))]
    params = np.array([history_clip_max, current_clip_factor, max_history_length], np.float32)
    out = np.zeros((h, w, 4), np.float32)
    L.oracle_temporal(w, h, *[_p(a) for a in arrs], _p(params), _p(out))
    return out


def f2h(values):
    """binary16 bits (round to nearest even) of float32 values, by the oracle."""
    L = lib()
    if not getattr(L, "_f2h_bound", False):
        L.oracle_f2h.restype = C.c_uint16
        L.oracle_f2h.argtypes = [C.c_float]
        L.oracle_h2f.restype = C.c_float
        L.oracle_h2f.argtypes = [C.c_uint16]
        L._f2h_bound = True
    return np.array([L.oracle_f2h(float(v)) for v in np.asarray(values, np.float32).ravel()], np.uint16)


GPU_TRIANGLE_DTYPE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4), ("d0", "<f4", 4),
                               ("d1", "<f4", 4), ("d2", "<f4", 4)])


def unpack_triangles(records):
    """UnpackTriangle (raytracer/common.hsh:17-44) of GPUTriangle records:
    dict of normals (n, 9), material (n,), opacity (n,), uvs (n, 6),
    tangents (n, 6), colors (n, 12)."""
    r = np.ascontiguousarray(records, dtype=GPU_TRIANGLE_DTYPE)
    n = len(r)
    out = {"normals": np.zeros((n, 9), np.float32), "material": np.zeros(n, np.int32),
           "opacity": np.zeros(n, np.float32), "uvs": np.zeros((n, 6), np.float32),
           "tangents": np.zeros((n, 6), np.float32), "colors": np.zeros((n, 12), np.float32)}
    L = lib()
    L.oracle_unpack_triangles.restype = None
    L.oracle_unpack_triangles.argtypes = [C.c_void_p, C.c_uint32] + [C.c_void_p] * 6
    L.oracle_unpack_triangles(_p(r), n, _p(out["normals"]), _p(out["material"]), _p(out["opacity"]), _p(out["uvs"]),
                              _p(out["tangents"]), _p(out["colors"]))
    return out


def pathtrace_counts():
    """Traversal work of the last pathtrace() call: rays traced (closest hits
    and shadow rays), inner nodes fetched (TLAS and BLAS), triangles tested,
    instances entered."""
    out = (C.c_uint64 * 4)()
    L = lib()
    L.oracle_pathtrace_counts.restype = None
    L.oracle_pathtrace_counts.argtypes = [C.c_void_p]
    L.oracle_pathtrace_counts(C.cast(out, C.c_void_p))
    return {"rays": out[0], "node_visits": out[1], "tri_tests": out[2], "instance_visits": out[3]}


def sincos(x):
    """The restatement's (sin, cos, x^5) of float32 values (the path tracer's
    sample-angle sine / cosine and pow(x, 5))."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    sn, cs, p5 = (np.zeros_like(x) for _ in range(3))
    L = lib()
    if not getattr(L, "_sincos_bound", False):
        L.oracle_sincos.restype = None
        L.oracle_sincos.argtypes = [C.c_void_p] * 4 + [C.c_uint32]
        L._sincos_bound = True
    L.oracle_sincos(_p(x), _p(sn), _p(cs), _p(p5), len(x))
    return sn, cs, p5


def h2f(bits):
    f2h([0.0])
    return np.array([lib().oracle_h2f(int(b)) for b in np.asarray(bits, np.uint16).ravel()], np.float32)


def pathtrace_cache(tlas_nodes, instances, xform, blases, shading, materials, lights, env, cam, width, height,
                    bounces, train_stride, frame_index=0, sample_index=0, pcg_draws=0, threads=None, surface=None):
    """One frame of the radiance-cache mode.  Returns (radiance (h, w, 3)
    added before any cache term, query (h, w, 16): cache query (12), the
    throughput reaching the vertex (3), flag; train (h, w) bool; rec
    (h, w, bounces + 1, 20): per bounce of a training path its query (12),
    throughput reaching it (3), surface flag, radiance added (3), 0)."""
    args, keep = _scene_args(tlas_nodes, instances, xform, blases, shading, materials, lights, env)
    ext = _surface_arg(surface, len(blases), keep)
    L = lib()
    if not getattr(L, "_ptc_bound", False):
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
        L.oracle_pathtrace_cache.restype = None
        L.oracle_pathtrace_cache.argtypes = _SCENE_TYPES + [vp, u32, u32, i32, u32, u32, C.c_uint64, u32, vp, vp, vp,
                                                            vp, i32, vp]
        L._ptc_bound = True
    rad = np.zeros((height, width, 4), np.float32)
    query = np.zeros((height, width, 16), np.float32)
    train = np.zeros((height, width), np.int32)
    rec = np.zeros((height, width, bounces + 1, 20), np.float32)
    L.oracle_pathtrace_cache(*args, C.addressof(cam), width, height, bounces, frame_index, sample_index, pcg_draws,
                             train_stride, _p(rad), _p(query), _p(train), _p(rec), threads or threads_default(), ext)
    _collect_lights(surface)
    return rad[..., :3], query, train.astype(bool), rec


def build_lbvh(verts):
    """CPU restatement of the GPU linear-BVH builder: (nodes, tris, orig,
    max_depth) in the engine's packed layout."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
    n = len(verts)
    nodes = np.zeros(max(n - 1, 0), dtype=NODE_DTYPE)
    tris = np.zeros(n, dtype=GPU_TRI_DTYPE)
    orig = np.zeros(n, dtype=np.uint32)
    depth = np.zeros(1, dtype=np.uint32)
    L = lib()
    if not getattr(L, "_lbvh_bound", False):
        L.oracle_lbvh_build.restype = None
        L.oracle_lbvh_build.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_radix_tree.restype = None
        L.oracle_radix_tree.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L._lbvh_bound = True
    L.oracle_lbvh_build(_p(verts), n, _p(nodes) if len(nodes) else None, _p(tris), _p(orig), _p(depth))
    return nodes, tris, orig, int(depth[0])


def radix_tree(keys):
    """Children (n - 1, 2) of the binary radix tree of sorted distinct uint64
    keys (Karras' per-node search): >= 0 inner node, < 0 ~leaf."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros((max(len(keys) - 1, 0), 2), dtype=np.int32)
    build_lbvh(np.zeros((1, 9), np.float32))  # binds the entry points
    lib().oracle_radix_tree(_p(keys), len(keys), _p(out) if len(out) else None)
    return out


def world_refit(tlas, inst, xform, order, descs, blases):
    """The device refit restated: copies of (tlas, inst, xform) after moving
    the instances to descs (caller order), and the instances' world boxes
    ((n, 6), slot order); blases: (nodes, tris) per BLAS of the world's table."""
    tlas = np.ascontiguousarray(tlas).copy()
    inst = np.ascontiguousarray(inst).copy()
    xform = np.ascontiguousarray(xform, dtype=np.float32).copy()
    order = np.ascontiguousarray(order, dtype=np.uint32)
    descs = np.ascontiguousarray(descs)
    L = lib()
    if not getattr(L, "_refit_bound", False):
        L.oracle_world_refit.restype = None
        L.oracle_world_refit.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        L._refit_bound = True
    k = len(blases)
    keep = [(np.ascontiguousarray(n), np.ascontiguousarray(t)) for n, t in blases]
    np_ptrs = (C.c_void_p * k)(*[n.ctypes.data if len(n) else None for n, _ in keep])
    tp_ptrs = (C.c_void_p * k)(*[t.ctypes.data for _, t in keep])
    ncnt = np.array([len(n) for n, _ in keep], np.uint32)
    tcnt = np.array([len(t) for _, t in keep], np.uint32)
    boxes = np.zeros((len(inst), 6), np.float32)
    L.oracle_world_refit(_p(tlas) if len(tlas) else None, len(tlas), _p(inst), _p(xform), _p(order), len(inst),
                         _p(descs), C.cast(np_ptrs, C.c_void_p), _p(ncnt), C.cast(tp_ptrs, C.c_void_p), _p(tcnt),
                         _p(boxes))
    return tlas, inst, xform, boxes


def _bind_binning():
    L = lib()
    if not getattr(L, "_bin_bound", False):
        L.oracle_ray_bin.restype = C.c_uint32
        L.oracle_ray_bin.argtypes = [C.c_float, C.c_float, C.c_float]
        L.oracle_ray_bins.restype = None
        L.oracle_ray_bins.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_sort_rays_by_direction.restype = None
        L.oracle_sort_rays_by_direction.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L._bin_bound = True
    return L


def ray_bins(dirs):
    """Octahedral direction bin (0..63) of each (n, 3) float32 direction."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(dirs), dtype=np.uint32)
    if len(dirs):
        _bind_binning().oracle_ray_bins(_p(dirs), len(dirs), _p(out))
    return out


def sort_rays_by_direction(rays):
    """(sorted rays, perm, 65 bin offsets): stable order by direction bin."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    out = np.zeros_like(rays)
    perm = np.zeros(len(rays), dtype=np.uint32)
    offsets = np.zeros(65, dtype=np.uint32)
    _bind_binning().oracle_sort_rays_by_direction(_p(rays) if len(rays) else None, len(rays),
                                                  _p(out) if len(rays) else None, _p(perm) if len(rays) else None,
                                                  _p(offsets))
    return out, perm, offsets


def sample_opacity(texels, x, y):
    """Bilinear, repeat-addressed sample of an (h, w) uint8 opacity texture."""
    texels = np.ascontiguousarray(texels, dtype=np.uint8)
    L = lib()
    if not getattr(L, "_so_bound", False):
        L.oracle_sample_opacity.restype = C.c_float
        L.oracle_sample_opacity.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_float, C.c_float]
        L._so_bound = True
    return float(L.oracle_sample_opacity(texels.shape[1], texels.shape[0], _p(texels), x, y))


def trace_world_alpha(tlas_nodes, instances, blases, opacity, rays, mask=0xFFFFFFFF):
    """Per ray: HitAnyTransparency's transparency and HitClosestTransparency's
    (slot, instance slot, t), with baked per-slot opacity per BLAS.
    blases: list of (nodes, tris); opacity: list of (n,) float32 in slot order."""
    k = len(blases)
    keep = []
    node_ptrs, tri_ptrs, op_ptrs = (C.c_void_p * k)(), (C.c_void_p * k)(), (C.c_void_p * k)()
    node_counts, tri_counts = np.zeros(k, np.uint32), np.zeros(k, np.uint32)
    for i, ((bn, bt), op) in enumerate(zip(blases, opacity)):
        bn = np.ascontiguousarray(bn, dtype=NODE_DTYPE)
        bt = np.ascontiguousarray(bt, dtype=GPU_TRI_DTYPE)
        op = np.ascontiguousarray(op, dtype=np.float32)
        keep += [bn, bt, op]
        node_ptrs[i] = bn.ctypes.data if len(bn) else None
        tri_ptrs[i], op_ptrs[i] = bt.ctypes.data, op.ctypes.data
        node_counts[i], tri_counts[i] = len(bn), len(bt)
    tlas_nodes = np.ascontiguousarray(tlas_nodes, dtype=NODE_DTYPE)
    instances = np.ascontiguousarray(instances, dtype=INSTANCE_DTYPE)
    rays = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
    n = len(rays)
    tr, t = np.zeros(n, np.float32), np.zeros(n, np.float32)
    slot, inst = np.zeros(n, np.int32), np.zeros(n, np.int32)
    L = lib()
    if not getattr(L, "_twa_bound", False):
        vp, u32 = C.c_void_p, C.c_uint32
        L.oracle_trace_world_alpha.restype = None
        L.oracle_trace_world_alpha.argtypes = [vp, u32, vp, u32, vp, vp, vp, vp, u32, vp, vp, u32, u32, vp, vp, vp, vp]
        L._twa_bound = True
    L.oracle_trace_world_alpha(_p(tlas_nodes), len(tlas_nodes), _p(instances), len(instances),
                               C.cast(node_ptrs, C.c_void_p), _p(node_counts), C.cast(tri_ptrs, C.c_void_p),
                               _p(tri_counts), k, C.cast(op_ptrs, C.c_void_p), _p(rays), n, mask, _p(tr), _p(slot),
                               _p(inst), _p(t))
    return tr, slot, inst, t


def triangle_test(o, d, v0, v1, v2):
    args = _f32(o, d, v0, v1, v2)  # keep the arrays alive across the call
    sol = np.zeros(3, dtype=np.float32)
    h = lib().oracle_triangle_test(*[_p(a) for a in args], _p(sol))
    return bool(h), sol


def slab_test(o, d, bmin, bmax, tmin, tmax):
    args = _f32(o, d, bmin, bmax)
    dist = np.zeros(1, dtype=np.float32)
    h = lib().oracle_slab_gpu(*[_p(a) for a in args], tmin, tmax, _p(dist))
    return bool(h), float(dist[0])
