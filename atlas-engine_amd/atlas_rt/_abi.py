"""ctypes view of include/atlas_rt.h (struct layouts, status codes, prototypes).

numpy structured dtypes mirror the C structs byte for byte so host arrays can
be passed straight through the C ABI.
"""
import ctypes as C

import numpy as np

OK = 0
E_INVALID_ARG = 1
E_SIZE_MISMATCH = 2
E_EMPTY = 3
E_HIP = 4
E_STACK_OVERFLOW = 5
E_OUT_OF_MEMORY = 6
E_UNSUPPORTED = 7
E_NO_DEVICE = 8

BUILD_SERIAL = 0
BUILD_PARALLEL = 1
BUILD_HOST_ONLY = 2

KERNEL_AUTO = 0
KERNEL_BVH2_EXACT = 1
KERNEL_BVH2_PERSISTENT = 2
KERNEL_BVH4 = 3
KERNEL_BVH8 = 4
KERNEL_VARIANT_MASK = 0xF
KERNEL_STACK_REGISTER = 0x10
KERNEL_SLAB_FLOAT = 0x20
KERNEL_LEAF_BATCH = 0x40

PT_RAY_BINNING = 0x1
PT_WAVEFRONT = 0x2
PT_FRAMES_IN_FLIGHT_SHIFT = 8  # flags bits 8-11: progressive frames in flight (0 = default)
RAY_BINS = 64

ANY_ENGINE_HITANY = 0
ANY_RESTORE_ALWAYS = 1

MASK_ALL = 1 << 7
MASK_SHADOW = 1 << 6

# GPUBVHNode (RTStructures.h:95-103), 64 B.
NODE_DTYPE = np.dtype([
    ("left_min", "<f4", 3), ("left_max", "<f4", 3),
    ("right_min", "<f4", 3), ("right_max", "<f4", 3),
    ("left_ptr", "<i4"), ("right_ptr", "<i4"), ("padding", "<i4", 2),
])
# GPUBVHTriangle (RTStructures.h:23-27), 48 B.
GPU_TRI_DTYPE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4)])
# GPUBVHInstance (RTStructures.h:85-93), 64 B.
INSTANCE_DTYPE = np.dtype([
    ("inverse_matrix", "<f4", (3, 4)), ("mesh_offset", "<i4"), ("material_offset", "<i4"),
    ("next_instance", "<i4"), ("mask", "<u4"),
])
# PackedRay (structures.hsh:9-13), 48 B.
PACKED_RAY_DTYPE = np.dtype([("origin", "<f4", 4), ("direction", "<f4", 4), ("hit", "<f4", 4)])
# Volume::BVHTriangle input, 40 B.
BVH_TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3), ("idx", "<u4")])
# Compact ray (32 B) and hit (16 B).
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim_id", "<i4")])
AABB_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3)])
# atlas_rt_instance_desc, 112 B.
INSTANCE_DESC_DTYPE = np.dtype([
    ("object_to_world", "<f4", (3, 4)), ("world_to_object", "<f4", (3, 4)), ("blas", "<u4"),
    ("material_offset", "<i4"), ("mask", "<u4"), ("padding", "<u4"),
])
assert INSTANCE_DESC_DTYPE.itemsize == 112
# atlas_rt_batch, 16 B.
BATCH_DTYPE = np.dtype([("blas", "<u4"), ("ray_offset", "<u4"), ("ray_count", "<u4"), ("reserved", "<u4")])


class LaunchConfig(C.Structure):
    """atlas_rt_launch_config."""
    _fields_ = [("kernel", C.c_int32), ("block", C.c_uint32), ("refill", C.c_uint32), ("reserved", C.c_uint32)]


# atlas_rt_material (64 B), atlas_rt_light (32 B), atlas_rt_gpu_light (64 B).
MATERIAL_DTYPE = np.dtype([
    ("base_color", "<f4", 3), ("opacity", "<f4"), ("emissive", "<f4", 3), ("roughness", "<f4"),
    ("metalness", "<f4"), ("ao", "<f4"), ("reflectance", "<f4"), ("two_sided", "<i4"),
    ("opacity_texture", "<i4"), ("invert_uvs", "<i4"), ("id", "<i4"), ("normal_scale", "<f4"),
    ("base_color_texture", "<i4"), ("roughness_texture", "<i4"), ("metalness_texture", "<i4"),
    ("ao_texture", "<i4"), ("normal_texture", "<i4"), ("use_vertex_colors", "<i4"), ("reserved", "<i4", 2),
])
# GPUTriangle (96 B): v0..v2 = (vertex, 10-10-10-2 normal bits), d0 = (half2 uv
# bits x3, material index bits), d1 = (tangent bits, bitangent bits, end of
# leaf +-1, 0), d2 = (RGBA8 colour bits x3, opacity)
GPU_TRIANGLE_DTYPE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4), ("d0", "<f4", 4),
                               ("d1", "<f4", 4), ("d2", "<f4", 4)])
assert GPU_TRIANGLE_DTYPE.itemsize == 96
TEXTURE_FIELDS = ("base_color_texture", "opacity_texture", "roughness_texture", "metalness_texture", "ao_texture",
                  "normal_texture")
LIGHT_DTYPE = np.dtype([("direction", "<f4", 3), ("type", "<u4"), ("radiance", "<f4", 3), ("reserved", "<f4")])
GPU_LIGHT_DTYPE = np.dtype([("P", "<f4", 4), ("N", "<f4", 4), ("color", "<f4", 4), ("data", "<f4", 4)])
assert MATERIAL_DTYPE.itemsize == 96 and LIGHT_DTYPE.itemsize == 32 and GPU_LIGHT_DTYPE.itemsize == 64
LIGHT_DIRECTIONAL = 0
LIGHT_TRIANGLE = 1


class Texture(C.Structure):
    """atlas_rt_texture (UNORM8, 1-4 channels, row-major, host memory)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.c_void_p), ("channels", C.c_uint32),
                ("reserved", C.c_uint32)]


class Camera(C.Structure):
    """atlas_rt_camera."""
    _fields_ = [("location", C.c_float * 3), ("origin", C.c_float * 3), ("right", C.c_float * 3),
                ("bottom", C.c_float * 3)]


class RealtimeParams(C.Structure):
    """atlas_rt_realtime_params (defaults: PathTracingRenderer.h:38-43)."""
    _fields_ = [("samples_per_frame", C.c_int32), ("history_length_max", C.c_int32),
                ("history_clip_max", C.c_float), ("current_clip_factor", C.c_float), ("max_radiance", C.c_float)]


class FrameMatrices(C.Structure):
    """atlas_rt_frame_matrices: glm column-major mat4 (m[column][row])."""
    _fields_ = [("view", C.c_float * 16), ("projection", C.c_float * 16), ("last_view_projection", C.c_float * 16),
                ("jitter", C.c_float * 2), ("reserved", C.c_float * 2)]


RT_RESOLVED, RT_FRAME, RT_DEPTH, RT_VELOCITY, RT_NORMAL, RT_MATERIAL = range(6)


class PtParams(C.Structure):
    """atlas_rt_pt_params."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("bounces", C.c_int32), ("exposure", C.c_float),
                ("env_radiance", C.c_float * 3), ("flags", C.c_uint32)]

class NrcParams(C.Structure):
    """atlas_rt_nrc_params."""
    _fields_ = [("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("levels", C.c_uint32),
                ("log2_table", C.c_uint32), ("base_resolution", C.c_uint32), ("max_resolution", C.c_uint32),
                ("learning_rate", C.c_float), ("seed", C.c_uint32)]


class PtCacheParams(C.Structure):
    """atlas_rt_pt_cache_params."""
    _fields_ = [("train_stride", C.c_uint32), ("train_steps", C.c_uint32), ("query_lag", C.c_uint32),
                ("reserved", C.c_uint32)]


assert NODE_DTYPE.itemsize == 64 and GPU_TRI_DTYPE.itemsize == 48
assert INSTANCE_DTYPE.itemsize == 64 and PACKED_RAY_DTYPE.itemsize == 48
assert BVH_TRI_DTYPE.itemsize == 40 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16

vp = C.c_void_p
u32 = C.c_uint32
i32 = C.c_int32
f32 = C.c_float
status = C.c_int32

# name -> (restype, argtypes).  Every symbol include/atlas_rt.h declares.
PROTOTYPES = {
    "atlas_rt_last_error": (C.c_char_p, []),
    "atlas_rt_version": (C.c_char_p, []),
    "atlas_rt_device_count": (i32, []),
    "atlas_rt_blas_build": (status, [vp, vp, u32, u32, C.POINTER(vp)]),
    "atlas_rt_blas_create": (status, [vp, u32, vp, vp, u32, C.POINTER(vp)]),
    "atlas_rt_blas_destroy": (status, [vp]),
    "atlas_rt_blas_info": (status, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
    "atlas_rt_blas_export": (status, [vp, vp, vp, vp]),
    "atlas_rt_trace_closest_rays": (status, [vp, vp, vp, vp, u32, i32, vp]),
    "atlas_rt_trace_closest_rays_noalloc": (status, [vp, vp, vp, vp, u32, i32, vp]),
    "atlas_rt_trace_closest_stats": (status, [vp, vp, vp, vp, vp, u32, vp]),
    "atlas_rt_sort_rays_workspace": (status, [u32, C.POINTER(C.c_size_t)]),
    "atlas_rt_sort_rays_by_direction": (status, [vp, u32, vp, vp, vp, vp, C.c_size_t, vp]),
    "atlas_rt_trace_any_rays": (status, [vp, vp, vp, u32, i32, vp]),
    "atlas_rt_trace_closest_host": (status, [vp, vp, vp, u32]),
    "atlas_rt_trace_any_host": (status, [vp, vp, vp, u32]),
    "atlas_rt_last_kernel_ms": (status, [vp, C.POINTER(f32)]),
    "atlas_rt_blas_set_stack_limit": (status, [vp, u32]),
    "atlas_rt_blas_trace_status": (status, [vp, vp]),
    "atlas_rt_world_set_stack_limit": (status, [vp, u32]),
    "atlas_rt_world_trace_status": (status, [vp, vp]),
    "atlas_rt_world_set_history": (status, [vp, vp, u32]),
    "atlas_rt_pathtracer_status": (status, [vp, vp]),
    "atlas_rt_world_create": (status, [C.POINTER(vp)]),
    "atlas_rt_world_destroy": (status, [vp]),
    "atlas_rt_world_update": (status, [vp, vp, u32, vp, vp, u32]),
    "atlas_rt_world_refit": (status, [vp, vp, vp]),
    "atlas_rt_world_info": (status, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
    "atlas_rt_world_export": (status, [vp, vp, vp, vp]),
    "atlas_rt_trace_closest": (status, [vp, vp, vp, vp, u32, f32, f32, u32, vp]),
    "atlas_rt_trace_any": (status, [vp, vp, vp, vp, u32, f32, f32, u32, u32, vp]),
    "atlas_rt_world_trace_closest_rays": (status, [vp, vp, vp, vp, u32, u32, vp]),
    "atlas_rt_world_trace_any_rays": (status, [vp, vp, vp, u32, u32, u32, vp]),
    "atlas_rt_blas_build_many": (status, [vp, vp, u32, u32, vp]),
    "atlas_rt_batch_create": (status, [vp, u32, vp, u32, C.POINTER(vp)]),
    "atlas_rt_batch_destroy": (status, [vp]),
    "atlas_rt_batch_trace_closest": (status, [vp, vp, vp, vp, vp, vp]),
    "atlas_rt_batch_tuning": (status, [vp, vp, vp, vp, vp]),
    "atlas_rt_batch_last_kernel_ms": (status, [vp, C.POINTER(f32)]),
    "atlas_rt_batch_graph_capture": (status, [vp, vp, vp, vp, vp]),
    "atlas_rt_batch_graph_launch": (status, [vp, vp]),
    "atlas_rt_aabb_tree_build": (status, [vp, u32, u32, vp, C.POINTER(u32), vp, vp, C.POINTER(u32)]),
    "atlas_rt_blas_set_triangles": (status, [vp, vp, u32]),
    "atlas_rt_pathtracer_count_work": (status, [vp, i32]),
    "atlas_rt_pathtracer_capture": (status, [vp, vp, u32]),
    "atlas_rt_pathtracer_replay": (status, [vp, vp]),
    "atlas_rt_pathtracer_work_counts": (status, [vp, vp]),
    "atlas_rt_blas_set_shading": (status, [vp, vp, vp]),
    "atlas_rt_blas_set_opacity": (status, [vp, vp, vp]),
    "atlas_rt_blas_set_tangents": (status, [vp, vp]),
    "atlas_rt_blas_build_device": (status, [vp, vp, u32, u32, vp, vp, C.POINTER(f32)]),
    "atlas_rt_trace_closest_transparency": (status, [vp, vp, vp, vp, u32, f32, f32, u32, vp]),
    "atlas_rt_trace_transparency": (status, [vp, vp, vp, vp, u32, f32, f32, u32, vp]),
    "atlas_rt_pathtracer_set_textures": (status, [vp, vp, u32]),
    "atlas_rt_pathtracer_set_realtime": (status, [vp, vp]),
    "atlas_rt_pathtracer_render_realtime": (status, [vp, vp, vp, vp]),
    "atlas_rt_pathtracer_read_realtime": (status, [vp, u32, vp, vp]),
    "atlas_rt_pathtracer_set_emissive_lights": (status, [vp, vp, u32]),
    "atlas_rt_pathtracer_lights": (status, [vp, vp, u32, C.POINTER(u32)]),
    "atlas_rt_pathtracer_create": (status, [vp, vp, vp, u32, vp, u32, C.POINTER(vp)]),
    "atlas_rt_pathtracer_destroy": (status, [vp]),
    "atlas_rt_pathtracer_render": (status, [vp, vp, u32, vp]),
    "atlas_rt_pathtracer_reset": (status, [vp]),
    "atlas_rt_pathtracer_read": (status, [vp, vp, vp]),
    "atlas_rt_pathtracer_sample_count": (status, [vp, C.POINTER(i32)]),
    "atlas_rt_nrc_create": (status, [vp, C.POINTER(vp)]),
    "atlas_rt_nrc_destroy": (status, [vp]),
    "atlas_rt_nrc_infer": (status, [vp, vp, vp, u32, vp]),
    "atlas_rt_nrc_train": (status, [vp, vp, vp, u32, vp, vp]),
    "atlas_rt_nrc_gradients": (status, [vp, vp, vp, u32, vp, vp, vp]),
    "atlas_rt_nrc_get_params": (status, [vp, vp, vp]),
    "atlas_rt_nrc_set_params": (status, [vp, vp, vp]),
    "atlas_rt_nrc_sizes": (status, [vp, C.POINTER(u32), C.POINTER(u32)]),
    "atlas_rt_pathtracer_set_cache": (status, [vp, vp, vp]),
    "atlas_rt_pathtracer_cache_stats": (status, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(f32)]),
    "atlas_rt_pathtracer_cache_samples": (status, [vp, vp, vp, u32, C.POINTER(u32)]),
}
