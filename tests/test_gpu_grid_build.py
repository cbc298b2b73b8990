"""GPU build of the RegularGrid cell lists (mrt_grid_build_gpu) against the
oracle's RegularGrid (the reference's cell registration restated on the
CPU): for the oracle's own grid size and bounds, identical offsets and
identical cell lists (ascending primitive order) for triangles, spheres and
planes; scenes built with the GPU grid render the same frames as scenes
built on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import DeviceArray, mirror_scene, oracle_scene
from mobileraytracer_amd import abi

pytestmark = pytest.mark.gpu

CATS = {abi.CAT_TRIANGLE: "triangles", abi.CAT_SPHERE: "spheres", abi.CAT_PLANE: "planes"}


def gpu_grid(ctx, category, prims, gs, bounds):
    n = len(prims)
    off = np.zeros(gs ** 3 + 1, np.uint32)
    nref = C.c_uint32()
    abi.check(abi.lib().mrt_grid_build_gpu(ctx.h, category, prims.ctypes.data if n else None, n,
                                           np.ascontiguousarray(bounds, np.float32).ctypes.data, gs,
                                           off.ctypes.data, None, 0, C.byref(nref), abi.MEM_HOST),
              "mrt_grid_build_gpu")
    refs = np.zeros(max(1, nref.value), np.uint32)
    abi.check(abi.lib().mrt_grid_build_gpu(ctx.h, category, prims.ctypes.data if n else None, n,
                                           np.ascontiguousarray(bounds, np.float32).ctypes.data, gs,
                                           off.ctypes.data, refs.ctypes.data, len(refs), C.byref(nref),
                                           abi.MEM_HOST), "mrt_grid_build_gpu")
    return off, refs[:nref.value]


def check_scene(ctx, data, cam_min=(0, 0, 0), cam_max=(0, 0, 0)):
    sc = oracle_scene(data, abi.REGULAR_GRID, cam_min, cam_max)
    for cat, field in CATS.items():
        prims = getattr(data, field)
        if len(prims) == 0:
            continue
        gs, bounds, off, refs = sc.grid(cat)
        goff, grefs = gpu_grid(ctx, cat, prims, gs, bounds)
        assert np.array_equal(goff, off), f"{field}: offsets differ"
        assert np.array_equal(grefs, refs), f"{field}: cell lists differ"


def test_soup(gpu_ctx):
    check_scene(gpu_ctx, mrt.soup_scene(8, 100_000, 1000))


def test_icospheres(gpu_ctx):
    check_scene(gpu_ctx, mrt.icosphere_scene(3, 10, 4, 100))


def test_cornell(gpu_ctx):
    check_scene(gpu_ctx, mrt.cornell_scene(), (0, 0, -3.4), (0, 0, -3.4))


def test_planes_spheres_triangles(gpu_ctx):
    check_scene(gpu_ctx, mirror_scene(), (-0.2, 0.1, -3.0), (0.2, 0.3, -2.9))


def test_many_spheres(gpu_ctx):
    rng = np.random.default_rng(4)
    sph = np.zeros(3000, abi.SPHERE_DTYPE)
    sph["center"] = rng.uniform(0, 1, (3000, 3))
    sph["sq_radius"] = (rng.uniform(0.002, 0.03, 3000) ** 2).astype(np.float32)
    sph["material_id"] = 0
    mats = np.zeros(1, abi.MATERIAL_DTYPE)
    mats[0]["kd"] = (0.5, 0.5, 0.5)
    check_scene(gpu_ctx, mrt.SceneData(materials=mats, spheres=sph))


def test_device_memory(gpu_ctx):
    data = mrt.icosphere_scene(6, 4, 3, 20)
    sc = oracle_scene(data, abi.REGULAR_GRID)
    gs, bounds, off, refs = sc.grid(abi.CAT_TRIANGLE)
    tris = data.triangles
    bufs = [DeviceArray(tris.nbytes), DeviceArray(off.nbytes), DeviceArray(refs.nbytes)]
    try:
        bufs[0].upload(tris)
        nref = C.c_uint32()
        abi.check(abi.lib().mrt_grid_build_gpu(gpu_ctx.h, abi.CAT_TRIANGLE, bufs[0].data_ptr(), len(tris),
                                               bounds.ctypes.data, gs, bufs[1].data_ptr(), bufs[2].data_ptr(),
                                               len(refs), C.byref(nref), abi.MEM_DEVICE), "mrt_grid_build_gpu")
        assert nref.value == len(refs)
        goff = bufs[1].download(np.zeros_like(off))
        grefs = bufs[2].download(np.zeros_like(refs))
    finally:
        for b in bufs:
            b.free()
    assert np.array_equal(goff, off) and np.array_equal(grefs, refs)


def test_non_finite_box_is_refused(gpu_ctx):
    data = mrt.soup_scene(5, 500, 5)
    tris = data.triangles.copy()
    tris[10]["a"] = (np.inf, 0.0, 0.0)
    off = np.zeros(9, np.uint32)
    nref = C.c_uint32()
    bounds = np.array([0, 0, 0, 1, 1, 1], np.float32)
    rc = abi.lib().mrt_grid_build_gpu(gpu_ctx.h, abi.CAT_TRIANGLE, tris.ctypes.data, len(tris), bounds.ctypes.data, 2,
                                      off.ctypes.data, None, 0, C.byref(nref), abi.MEM_HOST)
    assert rc == abi.MRT_E_UNSUPPORTED


def test_scene_grid_builders_render_the_same(gpu_ctx):
    """mrt_scene_build with the GPU grid against the host grid: identical
    PathTracer frames (megakernel, RegularGrid traversal)."""
    data = mrt.soup_scene(9, 120_000, 1200)
    cam = mrt.make_camera(abi.CAMERA_PERSPECTIVE, (0.5, 0.5, -1.0), (0.5, 0.5, 0.5), (0, 1, 0),
                          float(np.float32(45.0) * np.float32(16.0 / 9.0)), 45.0)
    cmin, cmax = mrt.camera_aabb(cam)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 256, 144, 2, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, 3),
                           rr=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 4), light=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 5),
                           hemi=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 6), depth_min=2, depth_max=4,
                           kernel=abi.KERNEL_MEGAKERNEL)
    frames = []
    for builder in (abi.BUILDER_HOST, abi.BUILDER_GPU):
        sc = mrt.Scene(gpu_ctx, data)
        try:
            sc.build(abi.REGULAR_GRID, cmin, cmax, builder=builder)
            acc = np.zeros((144, 256, 3), np.float32)
            bm, _ = sc.render(desc, accum=acc)
            frames.append((bm, acc))
        finally:
            sc.close()
    assert np.array_equal(frames[0][0], frames[1][0])
    assert np.array_equal(frames[0][1], frames[1][1])
