"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/_build/liboracle.so, the CPU restatement of
MobileRT's ray-scene path (mrt_oracle.hpp / mrt_oracle.cpp).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker or the timed CPU baseline; the product never
loads it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    """Compile the oracle (make -C oracle)."""
    res = subprocess.run(["make", "-C", str(HERE), "-B"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: build it with `make -C oracle`")
        l = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        u32, i32, f32 = C.c_uint32, C.c_int32, C.c_float
        sig = {
            "oracle_scene_create": (P, [P]),
            "oracle_scene_destroy": (None, [P]),
            "oracle_scene_build": (C.c_int, [P, C.c_int, P]),
            "oracle_bvh_nodes": (u32, [P, C.c_int, P, u32]),
            "oracle_grid_info": (C.c_int, [P, C.c_int, P, P, P]),
            "oracle_grid_cells": (C.c_int, [P, C.c_int, P, P]),
            "oracle_trace_closest": (C.c_int, [P, P, u32, P, u32, P, P]),
            "oracle_trace_any": (C.c_int, [P, P, u32, P, u32]),
            "oracle_shade_rays": (C.c_int, [P, P, P, P, u32, u32, P, P, C.c_int]),
            "oracle_render": (C.c_int, [P, P, P, u32, P, u32, P, C.c_int, P, P]),
            "oracle_camera_rays": (C.c_int, [P, P, u32, P]),
            "oracle_triangle_make": (C.c_int, [P, P, P, i32, P, P, P]),
            "oracle_camera_make": (C.c_int, [C.c_int, P, P, P, f32, f32, P]),
            "oracle_triangle_intersect": (C.c_int, [P, P, P, P]),
            "oracle_triangle_intersect_any": (C.c_int, [P, P]),
            "oracle_triangle_aabb": (C.c_int, [P, P]),
            "oracle_triangle_intersect_box": (C.c_int, [P, P, P]),
            "oracle_plane_intersect_box": (C.c_int, [P, P, P]),
            "oracle_sphere_intersect_box": (C.c_int, [P, P, P]),
            "oracle_grid_walk": (C.c_int, [P, P, P, u32, P, P]),
            "oracle_ordered_traffic": (C.c_int, [P, P, u32, u32, P]),
            "oracle_plane_distance": (f32, [P, P]),
            "oracle_plane_make": (C.c_int, [P, P, i32, P]),
            "oracle_slab": (C.c_int, [P, P, P]),
            "oracle_halton": (f32, [u32, u32]),
            "oracle_fast_arctan": (f32, [f32]),
            "oracle_incremental_avg": (u32, [P, u32, u32]),
            "oracle_draw": (f32, [P, u32, u32, u32, u32]),
            "oracle_grid_size_for": (i32, [u32]),
            "oracle_struct_sizes": (C.c_int, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return C.addressof(a)


BVH_NODE_DTYPE = np.dtype({"names": ["mn", "mx", "offset", "count"],
                           "formats": [("<f4", 3), ("<f4", 3), "<u4", "<u4"],
                           "offsets": [0, 12, 24, 28], "itemsize": 32})


class OracleScene:
    """Float and double CPU scenes built from the same stored forms."""

    def __init__(self, scene_desc):
        self.desc = scene_desc
        self.h = lib().oracle_scene_create(C.addressof(scene_desc))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def build(self, accel: int, params=None):
        return lib().oracle_scene_build(self.h, accel, _p(params))

    def bvh_nodes(self, category: int) -> np.ndarray:
        n = lib().oracle_bvh_nodes(self.h, category, None, 0)
        out = np.zeros(n, BVH_NODE_DTYPE)
        lib().oracle_bvh_nodes(self.h, category, out.ctypes.data, n)
        return out

    def grid(self, category: int):
        gs = C.c_int32()
        bounds = np.zeros(6, np.float32)
        nref = C.c_uint32()
        lib().oracle_grid_info(self.h, category, C.byref(gs), bounds.ctypes.data, C.byref(nref))
        off = np.zeros(gs.value ** 3 + 1, np.uint32)
        refs = np.zeros(max(1, nref.value), np.uint32)
        lib().oracle_grid_cells(self.h, category, off.ctypes.data, refs.ctypes.data)
        return gs.value, bounds, off, refs[: nref.value]

    def grid_walk(self, ray: np.ndarray, hit_dtype, cap: int = 4096):
        """Triangle-grid walk of one ray: (visited cell indices in walk
        order, hit record)."""
        cells = np.zeros(cap, np.uint32)
        n = C.c_uint32()
        hit = np.zeros(1, hit_dtype)
        rc = lib().oracle_grid_walk(self.h, ray.ctypes.data, cells.ctypes.data, cap, C.byref(n), hit.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"oracle_grid_walk failed: {rc}")
        if n.value > cap:
            raise RuntimeError(f"grid walk visited {n.value} cells, more than {cap}")
        return cells[: n.value].copy(), hit[0]

    def ordered_traffic(self, desc, pixel_mod: int = 256, threads: int = 8) -> dict:
        """Node-pair visits and triangle tests per ray of an ordered
        (near-first, culled) traversal of the same trees, over the rays of
        the pixels whose id is a multiple of pixel_mod (BVH scenes)."""
        out = np.zeros(8, np.float64)
        rc = lib().oracle_ordered_traffic(self.h, C.addressof(desc), pixel_mod, threads, out.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"oracle_ordered_traffic failed: {rc}")
        keys = ["rays_closest", "pairs_closest", "tests_closest", "rays_shadow", "pairs_shadow", "tests_shadow",
                "pairs_lights", "tests_lights"]
        return dict(zip(keys, out.tolist()))

    def trace_closest(self, rays: np.ndarray, threads: int = 8, hit_dtype=None):
        out = np.zeros(len(rays), hit_dtype)
        pairs, prims = C.c_uint64(), C.c_uint64()
        lib().oracle_trace_closest(self.h, rays.ctypes.data, len(rays), out.ctypes.data, threads, C.byref(pairs),
                                   C.byref(prims))
        return out, pairs.value, prims.value

    def trace_any(self, rays: np.ndarray, threads: int = 8) -> np.ndarray:
        out = np.zeros(len(rays), np.uint8)
        lib().oracle_trace_any(self.h, rays.ctypes.data, len(rays), out.ctypes.data, threads)
        return out

    def shade_rays(self, desc, rays: np.ndarray, depths=None, sample: int = 0, use_double: bool = False):
        n = len(rays)
        rgb = np.zeros((n, 3), np.float32)
        flag = np.zeros(n, np.uint8)
        dp = None if depths is None else np.ascontiguousarray(depths, np.int32)
        lib().oracle_shade_rays(self.h, C.addressof(desc), rays.ctypes.data, _p(dp), n, sample, rgb.ctypes.data,
                                flag.ctypes.data, 1 if use_double else 0)
        return rgb, flag

    def render(self, desc, bitmap=None, accum=None, threads: int = 8, tile_mask=None, use_double: bool = False,
               stride_bytes=None):
        if bitmap is None:
            bitmap = np.zeros((desc.height, desc.width), np.uint32)
        if stride_bytes is None:
            stride_bytes = desc.width * 4
        counters = np.zeros(4, np.uint64)
        secs = C.c_double()
        tm = None if tile_mask is None else np.ascontiguousarray(tile_mask, np.uint8)
        rc = lib().oracle_render(self.h, C.addressof(desc), bitmap.ctypes.data, stride_bytes, _p(accum), threads,
                                 _p(tm), 1 if use_double else 0, counters.ctypes.data, C.byref(secs))
        if rc != 0:
            raise RuntimeError(f"oracle_render failed: {rc}")
        return bitmap, counters, secs.value


def camera_rays(cam, uvdd: np.ndarray, ray_dtype) -> np.ndarray:
    uvdd = np.ascontiguousarray(uvdd, np.float32)
    out = np.zeros(len(uvdd), ray_dtype)
    lib().oracle_camera_rays(C.addressof(cam), uvdd.ctypes.data, len(uvdd), out.ctypes.data)
    return out


def struct_sizes() -> list[int]:
    out = np.zeros(16, np.uint32)
    n = lib().oracle_struct_sizes(out.ctypes.data)
    return out[:n].tolist()
