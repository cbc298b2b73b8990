"""Parity of the headline workload at its full size: the 1M-triangle soup
(990,000 diffuse triangles, 10,000 area lights), 1920x1080, PathTracer at 16
samples per pixel, depth 6 with Russian roulette from depth 3, per-pixel
Halton jitter, rendered exactly as the benchmark renders it (AUTO kernel:
the persistent wavefront, three slot groups, 8 path slots per pixel, walk
trees written on the device by the GPU BVH build and compressed there, the
light BVH over the 10k lights), against the oracle's Renderer::renderFrame
restatement (Renderer.cpp:67-116, PathTracer.cpp:25-164, BVH.hpp:158-271) on
a fixed subset of tiles.  The persistent megakernel runs the same 16-spp
frame; the per-pass megakernel and the 128^3 RegularGrid
(RegularGrid.hpp:157-465) run the frame at 2 samples per pixel.

The oracle renders only the subset's tiles (its tile mask); the GPU renders
the whole frame as benchmarked, and the subset's pixels are compared:
  * FP32 radiance sums within 1e-3 relative (1e-5 absolute) per channel (the
    GPU sums a path front to back, the reference back to front);
  * the RGBA8 running mean within 1 code value.
A second GPU render of the same desc with every other tile skipped
(tile_skip) gives the subset's own ray counts, which must equal the oracle's
camera / extension / shadow counts, and its subset pixels must be
bit-identical to the whole frame's.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import oracle_scene, soup_camera
from mobileraytracer_amd import abi

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
BSX, BSY = W // 16, H // 16
# Tile ids (row * 16 + column): the centre of the soup, two off-centre
# tiles and one at the left edge of the cube's image (hits and misses).
TILES_16SPP = (7 * 16 + 7, 2 * 16 + 5, 13 * 16 + 10, 8 * 16 + 3)
TILES_2SPP = (7 * 16 + 8, 11 * 16 + 6)
ORACLE_THREADS = 16


def bench_desc(cam, spp, kernel=abi.KERNEL_AUTO, tiles=None, persistent=0):
    """The benchmark's descriptor (bench.py make_desc)."""
    return mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, kernel=kernel,
                           pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, 5),
                           rr=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 6),
                           light=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 7),
                           hemi=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 8),
                           depth_min=2, depth_max=6, tiles=tiles, persistent=persistent)


def tile_pixels(tiles):
    m = np.zeros((H, W), bool)
    for t in tiles:
        ty, tx = divmod(t, 16)
        m[ty * BSY:(ty + 1) * BSY, tx * BSX:(tx + 1) * BSX] = True
    return m


def claim_mask(desc, tiles):
    """Oracle tile mask: indexed by claim rank (the tile table order)."""
    vals = np.frombuffer(bytes(desc.tile_values), np.float32)
    ids = np.rint(vals * 256).astype(int)
    keep = set(tiles)
    return np.array([1 if t in keep else 0 for t in ids], np.uint8)


@pytest.fixture(scope="module")
def soup():
    t0 = time.time()
    data = mrt.soup_scene(2, 1_000_000, 10_000, 5.0, 0.005, abi.SAMPLER_RANDOM)
    print(f"soup generated in {time.time() - t0:.1f} s", flush=True)
    return data


# Oracle renders of a tile subset, shared by the kernels that render the
# same frame: (accel, samples per pixel, tiles) -> (bitmap, accum, counts).
_ORACLE = {}


def oracle_render(data, accel, desc_full, tiles, cmin, cmax):
    key = (accel, desc_full.samples_pixel, tuple(tiles))
    if key not in _ORACLE:
        t0 = time.time()
        orc = oracle_scene(data, accel, cmin, cmax)
        t1 = time.time()
        acc_o = np.zeros((H, W, 3), np.float64)
        bm_o, cnt, secs = orc.render(desc_full, accum=acc_o, threads=ORACLE_THREADS,
                                     tile_mask=claim_mask(desc_full, tiles))
        print(f"oracle: build {t1 - t0:.1f} s, render {secs:.1f} s: {int(cnt[0])} camera {int(cnt[1])} extension "
              f"{int(cnt[2])} shadow rays", flush=True)
        _ORACLE[key] = (bm_o, acc_o, cnt)
    return _ORACLE[key]


def check_against_oracle(ctx, data, accel, desc_full, desc_sub, tiles, shape, lanes, cam_min=None, cam_max=None):
    """shape: the kernel the frame must run on: "wavefront" (three slot
    groups, `lanes` path slots per pixel), "persistent" (the persistent
    megakernel, `lanes` sample lanes per pixel) or "pass" (the megakernel
    per sample pass)."""
    cmin = cam_min if cam_min is not None else (0, 0, 0)
    cmax = cam_max if cam_max is not None else (0, 0, 0)
    gpu = mrt.Scene(ctx, data)
    try:
        if accel == abi.REGULAR_GRID:
            gpu.build(accel, cmin, cmax)
        else:
            gpu.build(accel)
        acc_full = np.zeros((H, W, 3), np.float32)
        bm_full, st_full = gpu.render(desc_full, accum=acc_full)
        print(f"GPU whole frame: {st_full.rays_camera} camera {st_full.rays_extension} extension "
              f"{st_full.rays_shadow} shadow rays, {st_full.ms_render:.1f} ms, {st_full.trace_groups} slot groups, "
              f"{st_full.slot_lanes} slots per pixel, persistent megakernel {st_full.persistent_megakernel}",
              flush=True)
        acc_sub = np.zeros((H, W, 3), np.float32)
        bm_sub, st_sub = gpu.render(desc_sub, accum=acc_sub)
    finally:
        gpu.close()
    if shape == "wavefront":
        assert st_full.trace_launches > 0 and st_full.trace_groups == 3 and st_full.slot_lanes == lanes
        assert st_full.persistent_megakernel == 0
    elif shape == "persistent":
        assert st_full.trace_launches == 0 and st_full.persistent_megakernel == 1 and st_full.slot_lanes == lanes
    else:
        assert st_full.trace_launches == 0 and st_full.persistent_megakernel == 0
    sel = tile_pixels(tiles)
    assert st_full.covered_pixels == 256 * BSX * BSY
    assert st_sub.covered_pixels == len(tiles) * BSX * BSY
    assert st_full.samples_done == st_sub.samples_done == desc_full.samples_pixel
    # Skipping the other tiles changes none of the subset's pixels and writes
    # no other pixel.
    assert np.array_equal(bm_sub[sel], bm_full[sel])
    assert np.array_equal(acc_sub[sel].view(np.uint32), acc_full[sel].view(np.uint32))
    assert not bm_sub[~sel].any() and not acc_sub[~sel].any()

    bm_o, acc_o, cnt = oracle_render(data, accel, desc_full, tiles, cmin, cmax)
    assert (st_sub.rays_camera, st_sub.rays_extension, st_sub.rays_shadow, st_sub.rays_specular) == \
        tuple(int(c) for c in cnt), "ray counts of the tile subset"
    assert int(cnt[0]) == len(tiles) * BSX * BSY * desc_full.samples_pixel
    ag, ao = acc_full[sel], acc_o[sel]
    close = np.isclose(ag, ao, rtol=1e-3, atol=1e-5).all(axis=1)
    assert close.all(), f"{int((~close).sum())} of {close.size} pixels differ in radiance"
    assert (ao > 0).any(axis=1).mean() > 0.05, "the subset should hold lit pixels"
    d8 = np.abs(bm_full[sel].view(np.uint8).astype(int) - bm_o[sel].view(np.uint8).astype(int))
    assert d8.max() <= 1, f"bitmap differs by {d8.max()}"


def test_headline_frame_bvh(gpu_ctx, soup):
    """The benchmarked frame: 16 spp, AUTO kernel (the persistent wavefront,
    3 slot groups, 8 slots per pixel), GPU-built trees, 10k-light tree.  The
    4-tile subset render is below the megakernel threshold, so it runs the
    megakernel: its pixels equal the wavefront frame's bit for bit."""
    cam = soup_camera(W, H)
    check_against_oracle(gpu_ctx, soup, abi.BVH, bench_desc(cam, 16), bench_desc(cam, 16, tiles=TILES_16SPP),
                         TILES_16SPP, "wavefront", lanes=8)


def test_headline_frame_bvh_megakernel(gpu_ctx, soup):
    """The same frame on the persistent megakernel (8 sample lanes per
    pixel)."""
    cam = soup_camera(W, H)
    k = abi.KERNEL_MEGAKERNEL
    check_against_oracle(gpu_ctx, soup, abi.BVH, bench_desc(cam, 16, k), bench_desc(cam, 16, k, tiles=TILES_16SPP),
                         TILES_16SPP, "persistent", lanes=8)


def test_headline_frame_bvh_megakernel_per_pass(gpu_ctx, soup):
    cam = soup_camera(W, H)
    k = abi.KERNEL_MEGAKERNEL
    check_against_oracle(gpu_ctx, soup, abi.BVH, bench_desc(cam, 2, k, persistent=2),
                         bench_desc(cam, 2, k, tiles=TILES_2SPP, persistent=2), TILES_2SPP, "pass", lanes=0)


@pytest.mark.parametrize("kernel", [abi.KERNEL_AUTO, abi.KERNEL_MEGAKERNEL])
def test_headline_frame_regular_grid(gpu_ctx, soup, kernel):
    """RegularGrid 128^3 over the 990k triangles (bounds joined with the
    camera's, Shader.cpp:52-62): AUTO runs the persistent wavefront's grid
    walk, MEGAKERNEL the megakernel per sample pass."""
    cam = soup_camera(W, H)
    cmin, cmax = mrt.camera_aabb(cam)
    auto = kernel == abi.KERNEL_AUTO
    check_against_oracle(gpu_ctx, soup, abi.REGULAR_GRID, bench_desc(cam, 2, kernel),
                         bench_desc(cam, 2, kernel, tiles=TILES_2SPP), TILES_2SPP,
                         "wavefront" if auto else "pass", lanes=2 if auto else 0, cam_min=cmin, cam_max=cmax)
