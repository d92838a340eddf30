"""Python view of the atlas_rt C ABI (include/atlas_rt.h).

This module is plumbing for tests and benchmarks: it loads the in-tree
libatlas_rt.so, mirrors the engine-facing objects (a BLAS is the engine's
Volume::BVH + packed MeshData arrays, src/engine/volume/BVH.h:114-138,
src/engine/mesh/MeshData.cpp:172-269) and passes torch device buffers to the
HIP kernels.  All compute happens in the library; there is no Python or CPU
fallback, and every call raises AtlasRTError when the library reports a
non-OK status or is missing.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (AABB_DTYPE, BVH_TRI_DTYPE, GPU_TRI_DTYPE, HIT_DTYPE, INSTANCE_DTYPE, NODE_DTYPE,
                   PACKED_RAY_DTYPE, RAY_DTYPE)

__all__ = [
    "AtlasRTError", "lib", "lib_path", "Blas", "NODE_DTYPE", "GPU_TRI_DTYPE", "INSTANCE_DTYPE",
    "PACKED_RAY_DTYPE", "BVH_TRI_DTYPE", "RAY_DTYPE", "HIT_DTYPE", "AABB_DTYPE",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path():
    return os.path.join(os.path.dirname(_HERE), "lib", "libatlas_rt.so")


class AtlasRTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"atlas_rt status {code}: {msg}")
        self.code = code


def lib():
    """Load libatlas_rt.so and bind every prototype; raises if it is not built."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise AtlasRTError(-1, f"{path} is not built (run __graft_entry__.build())")
        handle = C.CDLL(path)
        for name, (res, args) in _abi.PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(st):
    if st != _abi.OK:
        raise AtlasRTError(st, lib().atlas_rt_last_error().decode())


def _ptr(a):
    """Device pointer of a torch tensor, host pointer of a numpy array, or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    if isinstance(a, int):
        return a
    assert a.is_contiguous()
    return a.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def as_bvh_triangles(verts):
    """(n, 9) float32 vertices -> BVH_TRI_DTYPE records with idx = row."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
    out = np.zeros(len(verts), dtype=BVH_TRI_DTYPE)
    out["v0"] = verts[:, 0:3]
    out["v1"] = verts[:, 3:6]
    out["v2"] = verts[:, 6:9]
    out["idx"] = np.arange(len(verts), dtype=np.uint32)
    return out


class Blas:
    """A bottom-level BVH living on the current HIP device."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def build(cls, verts, aabbs=None, parallel=True, host_only=False):
        tris = verts if (isinstance(verts, np.ndarray) and verts.dtype == BVH_TRI_DTYPE) else as_bvh_triangles(verts)
        tris = np.ascontiguousarray(tris)
        if aabbs is not None:
            aabbs = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        flags = (_abi.BUILD_PARALLEL if parallel else 0) | (_abi.BUILD_HOST_ONLY if host_only else 0)
        out = C.c_void_p()
        check(lib().atlas_rt_blas_build(_ptr(aabbs), _ptr(tris), len(tris), flags, C.byref(out)))
        return cls(out.value)

    @classmethod
    def create(cls, nodes, tris, orig_idx=None):
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        tris = np.ascontiguousarray(tris, dtype=GPU_TRI_DTYPE)
        if orig_idx is not None:
            orig_idx = np.ascontiguousarray(orig_idx, dtype=np.uint32)
        out = C.c_void_p()
        check(lib().atlas_rt_blas_create(_ptr(nodes) if len(nodes) else None, len(nodes), _ptr(tris),
                                         _ptr(orig_idx), len(tris), C.byref(out)))
        return cls(out.value)

    def close(self):
        if self._h is not None and self._h.value:
            check(lib().atlas_rt_blas_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        n, t, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().atlas_rt_blas_info(self._h, C.byref(n), C.byref(t), C.byref(d)))
        return n.value, t.value, d.value

    def export(self):
        n, t, _ = self.info()
        nodes = np.zeros(n, dtype=NODE_DTYPE)
        tris = np.zeros(t, dtype=GPU_TRI_DTYPE)
        orig = np.zeros(t, dtype=np.uint32)
        check(lib().atlas_rt_blas_export(self._h, _ptr(nodes) if n else None, _ptr(tris), _ptr(orig)))
        return nodes, tris, orig

    def trace_closest(self, d_rays, d_hits, d_slots=None, count=None, kernel=_abi.KERNEL_AUTO, stream=None):
        """d_rays: device buffer of RAY_DTYPE records (e.g. float32 (n, 8)); d_hits: (n, 4) 32-bit."""
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_trace_closest_rays(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_slots), n, kernel,
                                                _stream_ptr(stream)))

    def trace_closest_noalloc(self, d_rays, d_hits, d_slots=None, count=None, kernel=_abi.KERNEL_AUTO, stream=None):
        """trace_closest through the entry point that allocates nothing
        (binary persistent kernels, fixed device scratch)."""
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_trace_closest_rays_noalloc(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_slots), n, kernel,
                                                        _stream_ptr(stream)))

    def trace_closest_stats(self, d_rays, d_hits, d_visits, d_slots=None, count=None, stream=None):
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_trace_closest_stats(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_slots),
                                                 _ptr(d_visits), n, _stream_ptr(stream)))

    def trace_any(self, d_rays, d_hit, count=None, kernel=_abi.KERNEL_AUTO, stream=None):
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_trace_any_rays(self._h, _ptr(d_rays), _ptr(d_hit), n, kernel, _stream_ptr(stream)))

    def trace_closest_host(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        check(lib().atlas_rt_trace_closest_host(self._h, _ptr(rays), _ptr(hits), len(rays)))
        return hits

    def trace_any_host(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hit = np.zeros(len(rays), dtype=np.uint32)
        check(lib().atlas_rt_trace_any_host(self._h, _ptr(rays), _ptr(hit), len(rays)))
        return hit

    def last_kernel_ms(self):
        ms = C.c_float()
        check(lib().atlas_rt_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def set_stack_limit(self, entries):
        """Cap the traversal stack at `entries` (0: sized from the tree)."""
        check(lib().atlas_rt_blas_set_stack_limit(self._h, int(entries)))

    def trace_status(self, stream=None):
        """The overflow status of the traces since the last call (waits for `stream`)."""
        return lib().atlas_rt_blas_trace_status(self._h, _stream_ptr(stream))


def aabb_tree_build(boxes, parallel=True):
    """The engine's AABB-only tree (Volume::BVH(aabbs)) on the host: (nodes,
    ref_idx, end_of_node) for (n, 6) float32 boxes (min xyz, max xyz)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    cap_nodes, cap_refs = max(n, 1), max(n, 2)
    nodes = np.zeros(cap_nodes, dtype=NODE_DTYPE)
    ref_idx = np.zeros(cap_refs, dtype=np.uint32)
    end = np.zeros(cap_refs, dtype=np.uint8)
    nn, nr = C.c_uint32(cap_nodes), C.c_uint32(cap_refs)
    check(lib().atlas_rt_aabb_tree_build(_ptr(boxes), n, 1 if parallel else 0, _ptr(nodes), C.byref(nn),
                                         _ptr(ref_idx), _ptr(end), C.byref(nr)))
    return nodes[:nn.value], ref_idx[:nr.value], end[:nr.value]


def instance_descs(object_to_world, blas_index, mask=_abi.MASK_ALL | _abi.MASK_SHADOW, material_offset=0):
    """Instance descriptors from (n, 3, 4) or (n, 4, 4) object-to-world matrices.

    The inverse is computed in float64 and rounded to float32, standing in for
    TransformComponent::inverseGlobalMatrix.
    """
    m = np.asarray(object_to_world, dtype=np.float64)
    n = len(m)
    full = np.zeros((n, 4, 4))
    full[:, :3, :] = m[:, :3, :]
    full[:, 3, 3] = 1.0
    inv = np.linalg.inv(full)
    d = np.zeros(n, dtype=_abi.INSTANCE_DESC_DTYPE)
    d["object_to_world"] = full[:, :3, :].astype(np.float32)
    d["world_to_object"] = inv[:, :3, :].astype(np.float32)
    d["blas"] = np.asarray(blas_index, dtype=np.uint32)
    d["material_offset"] = material_offset
    d["mask"] = mask
    return d


class World:
    """TLAS over instances of BLASes (RayTracingWorld, software path)."""

    def __init__(self):
        out = C.c_void_p()
        check(lib().atlas_rt_world_create(C.byref(out)))
        self._h = out
        self._blases = []

    def update(self, descs, blases, world_aabbs=None):
        descs = np.ascontiguousarray(descs, dtype=_abi.INSTANCE_DESC_DTYPE)
        self._blases = list(blases)
        table = (C.c_void_p * len(self._blases))(*[b.handle.value for b in self._blases])
        if world_aabbs is not None:
            world_aabbs = np.ascontiguousarray(world_aabbs, dtype=np.float32).reshape(-1, 6)
        check(lib().atlas_rt_world_update(self._h, _ptr(descs), len(descs), _ptr(world_aabbs),
                                          C.cast(table, C.c_void_p), len(self._blases)))

    def refit(self, descs, stream=None):
        """Device refit with new matrices (atlas_rt_world_refit): descs is a
        numpy array (copied to the device) or a uint8 CUDA tensor holding
        INSTANCE_DESC_DTYPE records, in the order of the last update."""
        import torch
        if isinstance(descs, np.ndarray):
            descs = np.ascontiguousarray(descs, dtype=_abi.INSTANCE_DESC_DTYPE)
            descs = torch.from_numpy(descs.view(np.uint8).copy()).cuda()
        check(lib().atlas_rt_world_refit(self._h, _ptr(descs), _stream_ptr(stream)))
        return descs

    def close(self):
        if self._h is not None and self._h.value:
            check(lib().atlas_rt_world_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, i, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().atlas_rt_world_info(self._h, C.byref(n), C.byref(i), C.byref(d)))
        return n.value, i.value, d.value

    def export(self):
        n, i, _ = self.info()
        nodes = np.zeros(n, dtype=NODE_DTYPE)
        inst = np.zeros(i, dtype=INSTANCE_DTYPE)
        order = np.zeros(i, dtype=np.uint32)
        check(lib().atlas_rt_world_export(self._h, _ptr(nodes) if n else None, _ptr(inst), _ptr(order)))
        return nodes, inst, order

    def set_stack_limit(self, entries):
        """Cap the traversal stack of the world's traces (0: sized from the depths)."""
        check(lib().atlas_rt_world_set_stack_limit(self._h, int(entries)))

    def trace_status(self, stream=None):
        """The overflow status of the world's traces since the last call."""
        return lib().atlas_rt_world_trace_status(self._h, _stream_ptr(stream))

    def set_history(self, last_object_to_world):
        """Last frame's object-to-world rows (n, 3, 4) of the last update's
        instances in its order (includeObjectHistory); None clears it."""
        if last_object_to_world is None:
            check(lib().atlas_rt_world_set_history(self._h, None, 0))
            return
        m = np.ascontiguousarray(last_object_to_world, dtype=np.float32).reshape(-1, 12)
        check(lib().atlas_rt_world_set_history(self._h, _ptr(m), len(m)))

    def trace_closest_rays(self, d_rays, d_hits, d_inst=None, count=None, mask=_abi.MASK_ALL, stream=None):
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_world_trace_closest_rays(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_inst), n, mask,
                                                      _stream_ptr(stream)))

    def trace_any_rays(self, d_rays, d_hit, count=None, mask=_abi.MASK_SHADOW, flags=_abi.ANY_ENGINE_HITANY,
                       stream=None):
        n = count if count is not None else d_rays.numel() // 8
        check(lib().atlas_rt_world_trace_any_rays(self._h, _ptr(d_rays), _ptr(d_hit), n, mask, flags,
                                                  _stream_ptr(stream)))

    def trace_closest_packed(self, d_in, d_out, d_count, max_count, tmin=0.0, tmax=1e12, mask=_abi.MASK_ALL,
                             stream=None):
        check(lib().atlas_rt_trace_closest(self._h, _ptr(d_in), _ptr(d_out), _ptr(d_count), max_count, tmin, tmax,
                                           mask, _stream_ptr(stream)))

    def trace_any_packed(self, d_in, d_occluded, d_count, max_count, tmin=0.0, tmax=1e12, mask=_abi.MASK_SHADOW,
                         flags=_abi.ANY_ENGINE_HITANY, stream=None):
        check(lib().atlas_rt_trace_any(self._h, _ptr(d_in), _ptr(d_occluded), _ptr(d_count), max_count, tmin, tmax,
                                       mask, flags, _stream_ptr(stream)))

    def trace_closest_transparency_packed(self, d_in, d_out, d_count, max_count, tmin=0.0, tmax=1e12,
                                          mask=_abi.MASK_ALL, stream=None):
        """HitClosestTransparency over PackedRay buffers (baked opacity)."""
        check(lib().atlas_rt_trace_closest_transparency(self._h, _ptr(d_in), _ptr(d_out), _ptr(d_count), max_count,
                                                        tmin, tmax, mask, _stream_ptr(stream)))

    def trace_transparency_packed(self, d_in, d_trans, d_count, max_count, tmin=0.0, tmax=1e12,
                                  mask=_abi.MASK_SHADOW, stream=None):
        """HitAnyTransparency over PackedRay buffers: float32 transparency per ray."""
        check(lib().atlas_rt_trace_transparency(self._h, _ptr(d_in), _ptr(d_trans), _ptr(d_count), max_count, tmin,
                                                tmax, mask, _stream_ptr(stream)))


def build_many(meshes, host_only=False):
    """Build one BLAS per (n, 9) vertex array, one host thread per mesh."""
    recs = [as_bvh_triangles(v) for v in meshes]
    k = len(recs)
    ptrs = (C.c_void_p * k)(*[r.ctypes.data for r in recs])
    counts = np.array([len(r) for r in recs], dtype=np.uint32)
    out = (C.c_void_p * k)()
    flags = _abi.BUILD_HOST_ONLY if host_only else 0
    check(lib().atlas_rt_blas_build_many(C.cast(ptrs, C.c_void_p), _ptr(counts), k, flags, C.cast(out, C.c_void_p)))
    return [Blas(out[i]) for i in range(k)]


def sort_rays_by_direction(d_rays, d_out, d_perm=None, d_bin_offsets=None, count=None, stream=None):
    """Stable sort of (n, 8) float32 device rays into d_out by the octahedral
    bin of their direction (atlas_rt_sort_rays_by_direction); optional int32
    permutation (d_out[i] = d_rays[d_perm[i]]) and 65 bin offsets."""
    import torch
    n = d_rays.shape[0] if count is None else count
    need = C.c_size_t()
    check(lib().atlas_rt_sort_rays_workspace(n, C.byref(need)))
    ws = torch.empty(max(need.value, 4), dtype=torch.uint8, device=d_rays.device)
    check(lib().atlas_rt_sort_rays_by_direction(_ptr(d_rays), n, _ptr(d_out), _ptr(d_perm), _ptr(d_bin_offsets),
                                                _ptr(ws), need.value, _stream_ptr(stream)))
    return ws


def build_many_device(meshes, stream=None):
    """GPU linear-BVH build (atlas_rt_blas_build_device) of one BLAS per (n, 9)
    vertex array; returns (blases, GPU milliseconds of the build kernels)."""
    arrs = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 9) for v in meshes]
    k = len(arrs)
    ptrs = (C.c_void_p * k)(*[a.ctypes.data for a in arrs])
    counts = np.array([len(a) for a in arrs], dtype=np.uint32)
    out = (C.c_void_p * k)()
    ms = C.c_float()
    check(lib().atlas_rt_blas_build_device(C.cast(ptrs, C.c_void_p), _ptr(counts), k, 0, _stream_ptr(stream),
                                           C.cast(out, C.c_void_p), C.byref(ms)))
    return [Blas(out[i]) for i in range(k)], ms.value


class BatchSet:
    """Many (BLAS, ray range) pairs traced by one persistent launch."""

    def __init__(self, blases, batches):
        self._blases = list(blases)
        self.batches = np.ascontiguousarray(batches, dtype=_abi.BATCH_DTYPE)
        table = (C.c_void_p * len(self._blases))(*[b.handle.value for b in self._blases])
        out = C.c_void_p()
        check(lib().atlas_rt_batch_create(C.cast(table, C.c_void_p), len(self._blases), _ptr(self.batches),
                                          len(self.batches), C.byref(out)))
        self._h = out

    def close(self):
        if self._h is not None and self._h.value:
            check(lib().atlas_rt_batch_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _cfg(kernel, block, refill):
        if kernel is None:
            return None
        return C.byref(_abi.LaunchConfig(kernel, block or 0, refill or 0, 0))

    def trace_closest(self, d_rays, d_hits, d_count=None, kernel=None, block=0, refill=0, stream=None):
        check(lib().atlas_rt_batch_trace_closest(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_count),
                                                 self._cfg(kernel, block, refill), _stream_ptr(stream)))

    def tuning(self):
        best = _abi.LaunchConfig()
        n = C.c_uint32(64)
        trials = (_abi.LaunchConfig * 64)()
        ms = np.zeros(64, dtype=np.float32)
        check(lib().atlas_rt_batch_tuning(self._h, C.byref(best), C.byref(n), C.cast(trials, C.c_void_p), _ptr(ms)))
        as_dict = lambda c: {"kernel": c.kernel, "block": c.block, "refill": c.refill}
        return as_dict(best), [(as_dict(trials[i]), float(ms[i])) for i in range(min(n.value, 64))]

    def last_kernel_ms(self):
        ms = C.c_float()
        check(lib().atlas_rt_batch_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def graph_capture(self, d_rays, d_hits, d_count=None, kernel=None, block=0, refill=0):
        check(lib().atlas_rt_batch_graph_capture(self._h, _ptr(d_rays), _ptr(d_hits), _ptr(d_count),
                                                 self._cfg(kernel, block, refill)))

    def graph_launch(self, stream=None):
        check(lib().atlas_rt_batch_graph_launch(self._h, _stream_ptr(stream)))


def _blas_set_shading(self, normals, material=None):
    """Per-triangle vertex normals ((n, 9), caller order) and material indices."""
    normals = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
    if material is not None:
        material = np.ascontiguousarray(material, dtype=np.int32)
    check(lib().atlas_rt_blas_set_shading(self._h, _ptr(normals), _ptr(material)))


Blas.set_shading = _blas_set_shading


def _blas_set_opacity(self, opacity=None, uvs=None):
    """Baked per-triangle opacity ((n,), caller order; -1 = opacity map) and
    texture coordinates ((n, 6): uv0, uv1, uv2).  Call after set_shading and
    before creating a PathTracer over the world."""
    if opacity is not None:
        opacity = np.ascontiguousarray(opacity, dtype=np.float32).reshape(-1)
    if uvs is not None:
        uvs = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 6)
    check(lib().atlas_rt_blas_set_opacity(self._h, _ptr(opacity), _ptr(uvs)))


Blas.set_opacity = _blas_set_opacity


def _blas_set_tangents(self, tangents):
    """Per-triangle tangent and bitangent ((n, 6), caller order) for normal maps."""
    tangents = np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 6)
    self._tangents = tangents
    check(lib().atlas_rt_blas_set_tangents(self._h, _ptr(tangents)))


Blas.set_tangents = _blas_set_tangents


def _blas_set_triangles(self, records):
    """The engine's GPUTriangle records (GPU_TRIANGLE_DTYPE, BVH slot order)
    as the path tracer's shading data (atlas_rt_blas_set_triangles)."""
    r = np.ascontiguousarray(records, dtype=_abi.GPU_TRIANGLE_DTYPE)
    check(lib().atlas_rt_blas_set_triangles(self._h, _ptr(r), len(r)))


Blas.set_triangles = _blas_set_triangles


def camera(eye, yaw_deg, pitch_deg, fov_y_deg, aspect, near=0.1):
    """atlas_rt_camera from an eye point and yaw/pitch (y up), built like
    CameraComponent::GetFrustumCorners' near-plane corners 4, 5, 6."""
    import math
    pitch, yaw = math.radians(pitch_deg), math.radians(yaw_deg)
    fwd = np.array([math.cos(pitch) * math.sin(yaw), -math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    h = near * math.tan(math.radians(fov_y_deg) / 2.0)
    w = aspect * h
    eye = np.asarray(eye, dtype=np.float64)
    near_pt = eye + fwd * near
    c4 = near_pt + h * up - w * right
    c5 = near_pt + h * up + w * right
    c6 = near_pt - h * up - w * right
    cam = _abi.Camera()
    cam.location[:] = eye.astype(np.float32)
    cam.origin[:] = c4.astype(np.float32)
    cam.right[:] = (c5 - c4).astype(np.float32)
    cam.bottom[:] = (c6 - c4).astype(np.float32)
    return cam


def view_matrix(eye, yaw_deg, pitch_deg):
    """4x4 world-to-view matrix (numpy, acting on column vectors) of the same
    eye and orientation as camera(): right-handed, looking down -z."""
    import math
    pitch, yaw = math.radians(pitch_deg), math.radians(yaw_deg)
    fwd = np.array([math.cos(pitch) * math.sin(yaw), -math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    eye = np.asarray(eye, dtype=np.float64)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = right, up, -fwd
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def projection_matrix(fov_y_deg, aspect, near=0.1, far=1000.0):
    """4x4 perspective projection (numpy), depth zero to one, y down (Vulkan)."""
    import math
    f = 1.0 / math.tan(math.radians(fov_y_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = -f
    m[2, 2] = far / (near - far)
    m[2, 3] = near * far / (near - far)
    m[3, 2] = -1.0
    return m


def frame_matrices(view, proj, last_view_proj, jitter=(0.0, 0.0)):
    """atlas_rt_frame_matrices from numpy 4x4 matrices acting on column
    vectors (stored column-major, as glm does)."""
    fm = _abi.FrameMatrices()
    for name, m in (("view", view), ("projection", proj), ("last_view_projection", last_view_proj)):
        getattr(fm, name)[:] = np.asarray(m, np.float64).T.astype(np.float32).ravel()
    fm.jitter[:] = np.asarray(jitter, np.float32)
    return fm


class PathTracer:
    """Progressive path tracer over a World (PathTracingRenderer).

    frames_in_flight: progressive frames overlapped on the tracer's own
    streams (1-4, None for the library's default); wavefront: bounce-by-bounce
    launches instead of the resident frame kernel.  The image depends on
    neither."""

    def __init__(self, world, width, height, bounces, materials, lights, env=(0.0, 0.0, 0.0), exposure=1.0,
                 binning=False, frames_in_flight=None, wavefront=False):
        flags = _abi.PT_RAY_BINNING if binning else 0
        if wavefront:
            flags |= _abi.PT_WAVEFRONT
        if frames_in_flight is not None:
            flags |= int(frames_in_flight) << _abi.PT_FRAMES_IN_FLIGHT_SHIFT
        p = _abi.PtParams(width, height, bounces, exposure, (C.c_float * 3)(*env), flags)
        self._mats = np.ascontiguousarray(materials, dtype=_abi.MATERIAL_DTYPE)
        self._lights = np.ascontiguousarray(lights, dtype=_abi.LIGHT_DTYPE)
        self._world = world
        self.width, self.height = width, height
        out = C.c_void_p()
        check(lib().atlas_rt_pathtracer_create(world._h, C.byref(p), _ptr(self._mats), len(self._mats),
                                               _ptr(self._lights) if len(self._lights) else None,
                                               len(self._lights), C.byref(out)))
        self._h = out

    def close(self):
        if self._h is not None and self._h.value:
            check(lib().atlas_rt_pathtracer_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, cam, frames=1, stream=None):
        check(lib().atlas_rt_pathtracer_render(self._h, C.byref(cam), frames, _stream_ptr(stream)))

    def capture(self, cam, frames=1):
        """Capture `frames` progressive frames of `cam` as a HIP graph."""
        check(lib().atlas_rt_pathtracer_capture(self._h, C.byref(cam), frames))

    def replay(self, stream=None):
        """Render the captured frames again (render(cam, frames) as of now)."""
        check(lib().atlas_rt_pathtracer_replay(self._h, _stream_ptr(stream)))

    def set_realtime(self, samples_per_frame=4, history_length_max=32, history_clip_max=0.2,
                     current_clip_factor=2.0, max_radiance=65535.0, enable=True):
        """Realtime mode (defaults of PathTracingRenderer.h:38-43), or off."""
        p = _abi.RealtimeParams(samples_per_frame, history_length_max, history_clip_max, current_clip_factor,
                                max_radiance)
        check(lib().atlas_rt_pathtracer_set_realtime(self._h, C.byref(p) if enable else None))

    def render_realtime(self, cam, matrices, stream=None):
        check(lib().atlas_rt_pathtracer_render_realtime(self._h, C.byref(cam), C.byref(matrices),
                                                        _stream_ptr(stream)))

    def read_realtime(self, buffer, stream=None):
        """One of this frame's realtime buffers (_abi.RT_*) as a CUDA tensor."""
        import torch
        shape, dtype = {
            _abi.RT_RESOLVED: ((self.height, self.width, 4), torch.float32),
            _abi.RT_FRAME: ((self.height, self.width, 4), torch.float32),
            _abi.RT_DEPTH: ((self.height, self.width), torch.float32),
            _abi.RT_VELOCITY: ((self.height, self.width), torch.int32),
            _abi.RT_NORMAL: ((self.height, self.width), torch.int32),
            _abi.RT_MATERIAL: ((self.height, self.width), torch.int16),
        }[buffer]
        out = torch.empty(shape, dtype=dtype, device="cuda")
        check(lib().atlas_rt_pathtracer_read_realtime(self._h, buffer, _ptr(out), _stream_ptr(stream)))
        return out

    def reset(self):
        check(lib().atlas_rt_pathtracer_reset(self._h))

    def read(self, out=None, stream=None):
        """(height, width, 4) float32 torch tensor: rgb radiance sum, sample count."""
        import torch
        if out is None:
            out = torch.empty((self.height, self.width, 4), dtype=torch.float32, device="cuda")
        check(lib().atlas_rt_pathtracer_read(self._h, _ptr(out), _stream_ptr(stream)))
        return out

    def status(self, stream=None):
        """The overflow status of the frames since the last call (waits for them)."""
        return lib().atlas_rt_pathtracer_status(self._h, _stream_ptr(stream))

    def sample_count(self):
        n = C.c_int32()
        check(lib().atlas_rt_pathtracer_sample_count(self._h, C.byref(n)))
        return n.value

    WORK_FIELDS = ("closest_rays", "shadow_rays", "hits", "nodes", "triangles", "instances")

    def count_work(self, enable=True):
        """Start (from zero) or stop counting the frames' traversal work."""
        check(lib().atlas_rt_pathtracer_count_work(self._h, 1 if enable else 0))

    def work_counts(self):
        """The traversal work since count_work(): a dict of WORK_FIELDS."""
        c = np.zeros(len(self.WORK_FIELDS), np.uint64)
        check(lib().atlas_rt_pathtracer_work_counts(self._h, _ptr(c)))
        return {k: int(v) for k, v in zip(self.WORK_FIELDS, c)}

    def set_textures(self, textures):
        """Material textures: a list of (height, width) or (height, width,
        channels) uint8 arrays (UNORM8, 1-4 channels)."""
        keep = [np.ascontiguousarray(t, dtype=np.uint8) for t in textures]
        arr = (_abi.Texture * max(len(keep), 1))()
        for i, t in enumerate(keep):
            arr[i].height, arr[i].width = t.shape[:2]
            arr[i].channels = t.shape[2] if t.ndim == 3 else 1
            arr[i].texels = t.ctypes.data
        check(lib().atlas_rt_pathtracer_set_textures(self._h, arr if keep else None, len(keep)))

    def set_emissive_lights(self, emitters):
        """Triangle lights from emissive materials (PathTracingRenderer's
        sampleEmissives): emitters (m, 4) float32 per material table entry,
        (sRGB emissive colour, emissive intensity); None turns them off."""
        if emitters is None:
            check(lib().atlas_rt_pathtracer_set_emissive_lights(self._h, None, 0))
            return
        em = np.ascontiguousarray(emitters, dtype=np.float32).reshape(-1, 4)
        check(lib().atlas_rt_pathtracer_set_emissive_lights(self._h, _ptr(em), len(em)))

    def lights(self):
        """The light list in use: atlas_rt_gpu_light records (GPU_LIGHT_DTYPE)."""
        n = C.c_uint32()
        check(lib().atlas_rt_pathtracer_lights(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, _abi.GPU_LIGHT_DTYPE)
        check(lib().atlas_rt_pathtracer_lights(self._h, _ptr(out) if n.value else None, n.value, C.byref(n)))
        return out

    def set_cache(self, cache, train_stride=32, train_steps=4, query_lag=0):
        """Attach a RadianceCache (None: plain path tracing).  query_lag 1:
        frame i of a render call queries the network as frame i - 2's training
        left it, so training overlaps the next frame (0: frame i - 1's)."""
        p = _abi.PtCacheParams(train_stride, train_steps, query_lag, 0)
        check(lib().atlas_rt_pathtracer_set_cache(self._h, cache._h if cache is not None else None, C.byref(p)))
        self._cache = cache

    def cache_samples(self, max_count):
        """The last frame's training samples: ((n, 12), (n, 4)) float32 device tensors."""
        import torch
        q = torch.empty((max_count, 12), dtype=torch.float32, device="cuda")
        t = torch.empty((max_count, 4), dtype=torch.float32, device="cuda")
        n = C.c_uint32()
        check(lib().atlas_rt_pathtracer_cache_samples(self._h, _ptr(q), _ptr(t), max_count, C.byref(n)))
        return q[:n.value], t[:n.value]

    def cache_stats(self):
        """(training samples, cache queries, last training loss) of the last frame."""
        n, q, l = C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().atlas_rt_pathtracer_cache_stats(self._h, C.byref(n), C.byref(q), C.byref(l)))
        return n.value, q.value, l.value


class RadianceCache:
    """Online-trained hash-grid + MLP radiance cache (atlas_rt_nrc_*).

    Queries are (N, 12) float32 device tensors: (position xyz, roughness),
    (outgoing direction xyz, unused), (normal xyz, unused); targets and
    outputs are (N, 4) float32 (rgb, unused)."""

    def __init__(self, bounds_min, bounds_max, levels=12, log2_table=19, base_resolution=16, max_resolution=2048,
                 learning_rate=1e-2, seed=0):
        p = _abi.NrcParams((C.c_float * 3)(*bounds_min), (C.c_float * 3)(*bounds_max), levels, log2_table,
                           base_resolution, max_resolution, learning_rate, seed)
        self.params = p
        out = C.c_void_p()
        check(lib().atlas_rt_nrc_create(C.byref(p), C.byref(out)))
        self._h = out
        m, t = C.c_uint32(), C.c_uint32()
        check(lib().atlas_rt_nrc_sizes(self._h, C.byref(m), C.byref(t)))
        self.mlp_size, self.table_size = m.value, t.value

    def close(self):
        if self._h is not None and self._h.value:
            check(lib().atlas_rt_nrc_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def infer(self, d_queries, out=None, stream=None):
        import torch
        n = d_queries.shape[0]
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float32, device=d_queries.device)
        check(lib().atlas_rt_nrc_infer(self._h, _ptr(d_queries), _ptr(out), n, _stream_ptr(stream)))
        return out

    def train(self, d_queries, d_targets, d_loss=None, stream=None):
        check(lib().atlas_rt_nrc_train(self._h, _ptr(d_queries), _ptr(d_targets), d_queries.shape[0],
                                       _ptr(d_loss) if d_loss is not None else None, _stream_ptr(stream)))

    def gradients(self, d_queries, d_targets, stream=None):
        """(weight gradients (mlp_size,), table gradients (table_size,)), FP32 device tensors."""
        import torch
        gm = torch.empty(self.mlp_size, dtype=torch.float32, device=d_queries.device)
        gt = torch.empty(self.table_size, dtype=torch.float32, device=d_queries.device)
        check(lib().atlas_rt_nrc_gradients(self._h, _ptr(d_queries), _ptr(d_targets), d_queries.shape[0], _ptr(gm),
                                           _ptr(gt), _stream_ptr(stream)))
        return gm, gt

    def get_params(self):
        """(FP32 master weights, FP16 table as int16 bits) device tensors."""
        import torch
        m = torch.empty(self.mlp_size, dtype=torch.float32, device="cuda")
        t = torch.empty(self.table_size, dtype=torch.int16, device="cuda")
        check(lib().atlas_rt_nrc_get_params(self._h, _ptr(m), _ptr(t)))
        return m, t

    def set_params(self, mlp=None, table=None):
        check(lib().atlas_rt_nrc_set_params(self._h, _ptr(mlp) if mlp is not None else None,
                                            _ptr(table) if table is not None else None))
