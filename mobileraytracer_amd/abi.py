"""ctypes mirror of include/mobilert_gpu.h (structs, enums, prototypes).

Loading fails loudly when libmrt_gpu.so is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmrt_gpu.so"

# ---- status / enums ---------------------------------------------------------
MRT_OK = 0
MRT_E_INVALID = -1
MRT_E_HIP = -2
MRT_E_NOMEM = -3
MRT_E_UNSUPPORTED = -4
MRT_E_BUDGET = -5
MRT_E_STOPPED = -6
MRT_E_NODEVICE = -7
MRT_E_TIMEOUT = -8

NAIVE, REGULAR_GRID, BVH = 0, 1, 2
SHADER_NOSHADOWS, SHADER_WHITTED, SHADER_PATHTRACER, SHADER_DEPTHMAP, SHADER_DIFFUSE = 0, 1, 2, 3, 4
CAT_PLANE, CAT_SPHERE, CAT_TRIANGLE, CAT_LIGHT = 0, 1, 2, 3
ID_NONE = 0xFFFFFFFF
SAMPLER_CONSTANT, SAMPLER_RANDOM, SAMPLER_HALTON, SAMPLER_STRATIFIED = 0, 1, 2, 3
SAMPLER_HALTON_COMPAT, SAMPLER_STRATIFIED_COMPAT = 4, 5
LIGHT_AREA, LIGHT_POINT = 0, 1
CAMERA_PERSPECTIVE, CAMERA_ORTHOGRAPHIC = 0, 1
KERNEL_AUTO, KERNEL_MEGAKERNEL, KERNEL_WAVEFRONT = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1
BUILDER_AUTO, BUILDER_HOST, BUILDER_GPU = 0, 1, 2
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3
CORNELL_TRIANGLES, CORNELL_MATERIALS, CORNELL_LIGHTS = 34, 7, 2


def make_id(cat: int, idx) -> np.ndarray:
    return (np.uint32(cat) << np.uint32(30)) | np.asarray(idx, dtype=np.uint32)


F3 = C.c_float * 3


class Triangle(C.Structure):
    _fields_ = [("ac", F3), ("ab", F3), ("a", F3), ("na", F3), ("nb", F3), ("nc", F3),
                ("ca", F3), ("cb", F3), ("cc", F3), ("material_id", C.c_int32)]


class Sphere(C.Structure):
    _fields_ = [("center", F3), ("sq_radius", C.c_float), ("material_id", C.c_int32)]


class Plane(C.Structure):
    _fields_ = [("normal", F3), ("point", F3), ("material_id", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("le", F3), ("kd", F3), ("ks", F3), ("kt", F3), ("ior", C.c_float)]


class Sampler(C.Structure):
    _fields_ = [("kind", C.c_int32), ("value", C.c_float), ("seed", C.c_uint32), ("domain_size", C.c_uint32)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_int32), ("radiance", Material), ("triangle", Triangle), ("position", F3),
                ("sampler", Sampler)]


class SceneDesc(C.Structure):
    _fields_ = [("triangles", C.c_void_p), ("n_triangles", C.c_uint32),
                ("spheres", C.c_void_p), ("n_spheres", C.c_uint32),
                ("planes", C.c_void_p), ("n_planes", C.c_uint32),
                ("materials", C.c_void_p), ("n_materials", C.c_uint32),
                ("lights", C.c_void_p), ("n_lights", C.c_uint32)]


class Ray(C.Structure):
    _fields_ = [("org", F3), ("dir", F3), ("tmax", C.c_float), ("exclude", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("id", C.c_uint32), ("u", C.c_float), ("v", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("kind", C.c_int32), ("position", F3), ("direction", F3), ("right", F3), ("up", F3),
                ("fov_h", C.c_float), ("fov_v", C.c_float)]


class RenderDesc(C.Structure):
    _fields_ = [("shader", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples_pixel", C.c_uint32), ("samples_light", C.c_uint32),
                ("ray_depth_min", C.c_int32), ("ray_depth_max", C.c_int32),
                ("camera", Camera),
                ("sampler_pixel", Sampler), ("sampler_rr", Sampler), ("sampler_light", Sampler),
                ("sampler_hemi", Sampler),
                ("tile_values", C.c_float * 256),
                ("kernel", C.c_int32), ("first_sample", C.c_uint32), ("instrument", C.c_uint32),
                ("ray_sort", C.c_uint32), ("persistent", C.c_uint32), ("depth_max_point", F3),
                ("tile_skip", C.c_uint32 * 8)]


class RenderStats(C.Structure):
    _fields_ = [("rays_camera", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("rays_specular", C.c_uint64), ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64),
                ("light_node_visits", C.c_uint64), ("light_prim_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_prim_tests", C.c_uint64),
                ("samples_done", C.c_uint32), ("covered_pixels", C.c_uint32),
                ("ms_render", C.c_float), ("ms_kernel", C.c_float), ("ms_trace", C.c_float),
                ("trace_launches", C.c_uint32), ("trace_lane_iters", C.c_uint64),
                ("trace_lane_slots", C.c_uint64), ("ms_trace_busy", C.c_float), ("trace_groups", C.c_uint32),
                ("slot_lanes", C.c_uint32), ("persistent_megakernel", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("refill_lanes", C.c_uint32), ("fetch_chunk", C.c_uint32), ("launch_share_pct", C.c_uint32),
                ("max_slot_lanes", C.c_uint32), ("max_groups", C.c_uint32), ("overlap_loads", C.c_uint32),
                ("graph_loop", C.c_uint32), ("service_lanes", C.c_uint32)]


class ServerDesc(C.Structure):
    _fields_ = [("max_slots", C.c_uint32), ("max_jobs", C.c_uint32), ("max_scenes", C.c_uint32),
                ("pipelines", C.c_uint32), ("job_memory", C.c_uint64), ("scene_memory", C.c_uint64),
                ("sample_memory", C.c_uint64), ("idle_direct", C.c_uint32), ("reserved", C.c_uint32)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("t_submit_us", C.c_double),
                ("t_admit_us", C.c_double), ("t_gpu_done_us", C.c_double), ("t_done_us", C.c_double),
                ("lanes", C.c_uint32), ("direct", C.c_uint32)]


class ServingStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64), ("jobs_failed", C.c_uint64),
                ("rays_camera", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("jobs_queued", C.c_uint32), ("jobs_running", C.c_uint32), ("slots_total", C.c_uint32),
                ("slots_used", C.c_uint32), ("scene_bytes", C.c_uint64), ("job_bytes_used", C.c_uint64),
                ("batches", C.c_uint64)]


class DenoiserDesc(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("channels", C.c_uint32 * 5), ("depth_scale", C.c_float),
                ("reserved", C.c_uint32 * 5)]


class DenoiserTrainDesc(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("crop", C.c_uint32), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("seed", C.c_uint64), ("reserved", C.c_uint32 * 8)]


class DenoiserTrainStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("last_loss", C.c_double), ("last_ms", C.c_float),
                ("ms_per_step", C.c_float), ("flops_per_step", C.c_double), ("reserved", C.c_uint32 * 8)]


class AccelParams(C.Structure):
    _fields_ = [("camera_min", F3), ("camera_max", F3), ("leaf_size", C.c_uint32), ("builder", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


def np_dtype(struct) -> np.dtype:
    """numpy dtype with the byte layout of a ctypes struct (for bulk arrays)."""
    return np.dtype((np.void, C.sizeof(struct)))


TRIANGLE_DTYPE = np.dtype({
    "names": ["ac", "ab", "a", "na", "nb", "nc", "ca", "cb", "cc", "material_id"],
    "formats": [("<f4", 3)] * 9 + ["<i4"],
    "offsets": [0, 12, 24, 36, 48, 60, 72, 84, 96, 108], "itemsize": 112})
RAY_DTYPE = np.dtype({"names": ["org", "dir", "tmax", "exclude"],
                      "formats": [("<f4", 3), ("<f4", 3), "<f4", "<u4"],
                      "offsets": [0, 12, 24, 28], "itemsize": 32})
HIT_DTYPE = np.dtype({"names": ["t", "id", "u", "v"], "formats": ["<f4", "<u4", "<f4", "<f4"],
                      "offsets": [0, 4, 8, 12], "itemsize": 16})
MATERIAL_DTYPE = np.dtype({"names": ["le", "kd", "ks", "kt", "ior"],
                           "formats": [("<f4", 3)] * 4 + ["<f4"], "offsets": [0, 12, 24, 36, 48],
                           "itemsize": 52})
SPHERE_DTYPE = np.dtype({"names": ["center", "sq_radius", "material_id"],
                         "formats": [("<f4", 3), "<f4", "<i4"], "offsets": [0, 12, 16], "itemsize": 20})
PLANE_DTYPE = np.dtype({"names": ["normal", "point", "material_id"],
                        "formats": [("<f4", 3), ("<f4", 3), "<i4"], "offsets": [0, 12, 24], "itemsize": 28})
LIGHT_DTYPE = np_dtype(Light)

_lib = None


def _proto(lib):
    P = C.c_void_p
    u32, i32, f32, u64 = C.c_uint32, C.c_int32, C.c_float, C.c_uint64
    sig = {
        "mrt_last_error": (C.c_char_p, []),
        "mrt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "mrt_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "mrt_ctx_destroy": (C.c_int, [P]),
        "mrt_ctx_set_budget": (C.c_int, [P, u64]),
        "mrt_ctx_synchronize": (C.c_int, [P]),
        "mrt_scene_create": (C.c_int, [P, C.POINTER(SceneDesc), C.POINTER(P)]),
        "mrt_scene_destroy": (C.c_int, [P]),
        "mrt_scene_build": (C.c_int, [P, C.c_int, C.POINTER(AccelParams)]),
        "mrt_scene_footprint": (C.c_int, [P, C.POINTER(u64)]),
        "mrt_trace_closest": (C.c_int, [P, P, u32, P, C.c_int]),
        "mrt_trace_closest_noalloc": (C.c_int, [P, P, u32, P]),
        "mrt_trace_any": (C.c_int, [P, P, u32, P, C.c_int]),
        "mrt_shade_rays": (C.c_int, [P, C.POINTER(RenderDesc), P, P, u32, u32, P, P, C.c_int]),
        "mrt_camera_rays": (C.c_int, [P, C.POINTER(Camera), P, u32, P, C.c_int]),
        "mrt_intersect_pairs": (C.c_int, [P, C.c_int, P, P, u32, P, P, C.c_int]),
        "mrt_box_rays": (C.c_int, [P, P, P, u32, P, P, C.c_int]),
        "mrt_render": (C.c_int, [P, C.POINTER(RenderDesc), P, u32, P, C.c_int, C.POINTER(RenderStats)]),
        "mrt_render_stop": (C.c_int, [P]),
        "mrt_render_progress": (C.c_int, [P, C.POINTER(u32)]),
        "mrt_render_aovs": (C.c_int, [P, C.POINTER(RenderDesc), P, C.c_int]),
        "mrt_render_with_aovs": (C.c_int, [P, C.POINTER(RenderDesc), P, u32, P, P, C.c_int, C.POINTER(RenderStats)]),
        "mrt_server_create": (C.c_int, [P, C.POINTER(ServerDesc), C.POINTER(P)]),
        "mrt_server_destroy": (C.c_int, [P]),
        "mrt_server_add_scene": (C.c_int, [P, P, C.POINTER(u32)]),
        "mrt_server_submit": (C.c_int, [P, u32, C.POINTER(RenderDesc), P, u32, P, C.c_int, C.POINTER(u64)]),
        "mrt_server_wait": (C.c_int, [P, u64, C.c_int64, C.POINTER(JobInfo)]),
        "mrt_server_poll": (C.c_int, [P, u64, C.POINTER(JobInfo)]),
        "mrt_server_cancel": (C.c_int, [P, u64]),
        "mrt_server_stats": (C.c_int, [P, C.POINTER(ServingStats)]),
        "mrt_denoiser_weight_count": (C.c_int, [C.POINTER(DenoiserDesc), C.POINTER(u64)]),
        "mrt_denoiser_create": (C.c_int, [P, C.POINTER(DenoiserDesc), P, u64, C.POINTER(P)]),
        "mrt_denoiser_run": (C.c_int, [P, P, f32, P, u32, u32, P, C.c_int]),
        "mrt_denoiser_last_ms": (C.c_int, [P, C.POINTER(f32)]),
        "mrt_denoiser_destroy": (C.c_int, [P]),
        "mrt_conv3x3_bf16": (C.c_int, [P, P, u32, u32, u32, P, P, u32, C.c_int, P, C.c_int]),
        "mrt_denoiser_trainer_create": (C.c_int, [P, C.POINTER(DenoiserDesc), P, u64, C.POINTER(DenoiserTrainDesc),
                                                  C.POINTER(P)]),
        "mrt_denoiser_trainer_set_pairs": (C.c_int, [P, P, P, u32, u32, u32, C.c_int]),
        "mrt_denoiser_trainer_crops": (C.c_int, [P, u64, P]),
        "mrt_denoiser_trainer_step": (C.c_int, [P, u32, C.c_double, C.POINTER(C.c_double)]),
        "mrt_denoiser_trainer_gradient": (C.c_int, [P, P, P, u32, C.c_double, C.POINTER(C.c_double), P, u64, C.c_int]),
        "mrt_denoiser_trainer_weights": (C.c_int, [P, P, u64]),
        "mrt_denoiser_trainer_stats": (C.c_int, [P, C.POINTER(DenoiserTrainStats)]),
        "mrt_denoiser_trainer_destroy": (C.c_int, [P]),
        "mrt_triangle_make": (C.c_int, [P, P, P, i32, P, P, P]),
        "mrt_camera_make": (C.c_int, [C.c_int, P, P, P, f32, f32, C.POINTER(Camera)]),
        "mrt_camera_aabb": (C.c_int, [C.POINTER(Camera), P]),
        "mrt_ctx_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "mrt_ctx_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "mrt_tile_values": (C.c_int, [u32, P]),
        "mrt_triangle_intersect_box": (C.c_int, [P, P, P]),
        "mrt_primitive_intersect_box": (C.c_int, [C.c_int, P, P, P]),
        "mrt_primitive_aabbs": (C.c_int, [C.c_int, P, u32, P]),
        "mrt_bvh_build": (C.c_int, [P, u32, u32, P, u32, C.POINTER(u32)]),
        "mrt_wide_build": (C.c_int, [P, P, u32, P, u32, C.POINTER(u32), C.POINTER(u32)]),
        "mrt_bvh_build_gpu": (C.c_int, [P, P, u32, u32, P, u32, C.POINTER(u32), C.c_int]),
        "mrt_grid_build_gpu": (C.c_int, [P, C.c_int, P, u32, P, i32, P, P, u32, C.POINTER(u32), C.c_int]),
        "mrt_scene_cornell": (C.c_int, [u32, P, P, P]),
        "mrt_scene_soup": (C.c_int, [u32, u32, u32, f32, f32, P, P, C.c_int]),
        "mrt_scene_icospheres": (C.c_int, [u32, u32, u32, f32, u32, f32, P, P, C.c_int]),
        "mrt_obj_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
        "mrt_obj_counts": (C.c_int, [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
        "mrt_obj_fill": (C.c_int, [P, P, P, P, C.POINTER(Sampler)]),
        "mrt_obj_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return sig


EXPORTS = None


def lib():
    """The loaded libmrt_gpu.so (raises if it was not built)."""
    global _lib, EXPORTS
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        l = C.CDLL(str(LIB_PATH))
        EXPORTS = list(_proto(l))
        _lib = l
    return _lib


class MrtError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().mrt_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != MRT_OK:
        raise MrtError(code, where)


def ptr(a) -> int | None:
    """Address of a numpy array, torch tensor or ctypes object (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return C.addressof(a)
