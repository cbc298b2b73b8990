"""GPU renderer parity (mrt_render / mrt_shade_rays through the C ABI) against
the oracle's Renderer::renderFrame restatement.

Tolerances:
  * Whitted without reflection/transmission: the RGBA8 bitmap and the FP32
    radiance are bit-exact (same operations in the same order);
  * Whitted / PathTracer with reflection or transmission, or with paths of
    three or more vertices: FP32 radiance within 1e-3 per channel relative
    (the GPU sums a path front to back, the reference back to front), the
    bitmap within 1 code value;
  * PathTracer statistics: per-pixel-block means of the GPU within 3 sigma of
    the FP64 oracle's (independent seeds).
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import cornell_camera, mirror_scene, oracle_scene, soup_camera
from mobileraytracer_amd import abi

pytestmark = pytest.mark.gpu

RANDOM = abi.SAMPLER_RANDOM


def rnd(seed):
    return mrt.sampler(RANDOM, 0.5, seed)


def render_both(ctx, data, desc, accel=abi.BVH, threads=8, cam_min=None, cam_max=None):
    if cam_min is None:
        cam_min, cam_max = mrt.camera_aabb(desc.camera)
    gpu = mrt.Scene(ctx, data)
    gpu.build(accel, cam_min, cam_max)
    acc_g = np.zeros((desc.height, desc.width, 3), np.float32)
    bm_g, stats = gpu.render(desc, accum=acc_g)
    orc = oracle_scene(data, accel, cam_min, cam_max)
    acc_o = np.zeros((desc.height, desc.width, 3), np.float64)
    bm_o, cnt, _ = orc.render(desc, accum=acc_o, threads=threads)
    return (bm_g, acc_g, stats), (bm_o, acc_o, cnt)


@pytest.mark.parametrize("accel", [abi.BVH, abi.REGULAR_GRID, abi.NAIVE])
def test_whitted_cornell_exact(gpu_ctx, accel):
    """Config 1: Cornell box, Whitted, 256x256, 1 spp, constant samplers."""
    cam = cornell_camera(256, 256)
    desc = mrt.render_desc(abi.SHADER_WHITTED, 256, 256, 1, cam)
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, mrt.cornell_scene(1), desc, accel)
    print(f"whitted cornell accel={accel}: rays camera={st.rays_camera} shadow={st.rays_shadow}", flush=True)
    assert np.array_equal(bg, bo), f"bitmap mismatches: {int(np.sum(bg != bo))}"
    assert np.array_equal(ag.view(np.uint32), ao.astype(np.float32).view(np.uint32))
    assert (st.rays_camera, st.rays_extension, st.rays_shadow, st.rays_specular) == tuple(int(c) for c in cnt)
    assert st.rays_camera == 256 * 256


def test_whitted_cornell_multisample_bitmap(gpu_ctx):
    """4 spp with random jitter: the 8-bit running mean (incrementalAvg) over
    four passes must match exactly."""
    cam = cornell_camera(128, 128)
    desc = mrt.render_desc(abi.SHADER_WHITTED, 128, 128, 4, cam, pixel=rnd(3))
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, mrt.cornell_scene(1), desc)
    assert np.array_equal(bg, bo)
    np.testing.assert_allclose(ag, ao, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("kind", [abi.SAMPLER_HALTON_COMPAT, abi.SAMPLER_STRATIFIED_COMPAT])
def test_compat_pixel_sampler(gpu_ctx, kind):
    """The reference's frame-wide HaltonSeq/Stratified counter in single-thread
    claim order (tile table shuffled with seed 9)."""
    cam = cornell_camera(256, 256)
    desc = mrt.render_desc(abi.SHADER_WHITTED, 256, 256, 2, cam, tile_seed=9,
                           pixel=mrt.sampler(kind, 0.0, 0, 256))
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, mrt.cornell_scene(1), desc, threads=1)
    assert np.array_equal(bg, bo)


@pytest.mark.parametrize("shader", [abi.SHADER_WHITTED, abi.SHADER_PATHTRACER])
def test_mirror_scene_branches(gpu_ctx, shader):
    """Planes, spheres, mirror and glass: reflection/transmission branches."""
    cam = cornell_camera(96, 96)
    desc = mrt.render_desc(shader, 96, 96, 2, cam, pixel=rnd(1), rr=rnd(2), light=rnd(3), hemi=rnd(4),
                           depth_min=2, depth_max=6)
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, mirror_scene(), desc)
    assert (st.rays_camera, st.rays_extension, st.rays_shadow, st.rays_specular) == tuple(int(c) for c in cnt)
    assert st.rays_specular > 0
    np.testing.assert_allclose(ag, ao, rtol=1e-3, atol=1e-5)
    assert np.max(np.abs(bg.view(np.uint8).astype(int) - bo.view(np.uint8).astype(int))) <= 1


@pytest.mark.parametrize("accel,kernel", [(abi.BVH, abi.KERNEL_MEGAKERNEL), (abi.REGULAR_GRID, abi.KERNEL_MEGAKERNEL),
                                          (abi.BVH, abi.KERNEL_WAVEFRONT), (abi.REGULAR_GRID, abi.KERNEL_WAVEFRONT)])
def test_pathtracer_soup(gpu_ctx, accel, kernel):
    """Config-2 shape at reduced size: PathTracer, depth 6, RR from depth 3."""
    data = mrt.soup_scene(4, 60000, 600)
    cam = soup_camera(128, 72)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 128, 80, 4, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                           rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6, kernel=kernel)
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, data, desc, accel, threads=16)
    print(f"pt soup accel={accel}: {st.rays_camera} cam {st.rays_extension} ext {st.rays_shadow} shadow", flush=True)
    assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(c) for c in cnt[:3])
    close = np.isclose(ag, ao, rtol=1e-3, atol=1e-5).all(axis=2)
    assert close.mean() == 1.0, f"{int((~close).sum())} pixels differ"
    assert np.max(np.abs(bg.view(np.uint8).astype(int) - bo.view(np.uint8).astype(int))) <= 1


def test_shade_rays_matches_oracle(gpu_ctx):
    data = mirror_scene()
    cam = cornell_camera(64, 64)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 64, 64, 1, cam, rr=rnd(2), light=rnd(3), hemi=rnd(4),
                           depth_min=2, depth_max=6)
    gpu = mrt.Scene(gpu_ctx, data)
    gpu.build(abi.BVH)
    orc = oracle_scene(data, abi.BVH)
    from helpers import pixel_uvdd
    rays = gpu_ctx.camera_rays(cam, pixel_uvdd(64, 64))
    rg, fg = gpu.shade_rays(desc, rays)
    ro, fo = orc.shade_rays(desc, rays)
    np.testing.assert_allclose(rg, ro, rtol=1e-3, atol=1e-5)
    assert np.array_equal(fg, fo)


def test_pathtracer_fp64_statistics(gpu_ctx):
    """GPU PathTracer means against the FP64 oracle with independent seeds,
    8x8-pixel blocks, 3 sigma (binomial slack on the outlier count)."""
    data = mrt.soup_scene(12, 20000, 400)
    cam = soup_camera(64, 64)
    spp = 64
    mk = lambda s: mrt.render_desc(abi.SHADER_PATHTRACER, 64, 64, spp, cam, pixel=rnd(s), rr=rnd(s + 1),
                                   light=rnd(s + 2), hemi=rnd(s + 3), depth_min=2, depth_max=6)
    gpu = mrt.Scene(gpu_ctx, data)
    gpu.build(abi.BVH)
    acc_g = np.zeros((64, 64, 3), np.float32)
    gpu.render(mk(100), accum=acc_g)
    orc = oracle_scene(data, abi.BVH)
    acc_o = np.zeros((64, 64, 3), np.float64)
    orc.render(mk(200), accum=acc_o, threads=16, use_double=True)
    # Per-block sample statistics from per-pixel means (pixels in a block are
    # treated as independent samples of the block mean).
    g = (acc_g / spp).reshape(8, 8, 8, 8, 3).transpose(0, 2, 1, 3, 4).reshape(64, 64, 3)
    o = (acc_o / spp).reshape(8, 8, 8, 8, 3).transpose(0, 2, 1, 3, 4).reshape(64, 64, 3)
    mg, mo = g.mean(axis=1), o.mean(axis=1)
    se = np.sqrt(g.var(axis=1, ddof=1) / 64 + o.var(axis=1, ddof=1) / 64) + 1e-6
    z = np.abs(mg - mo) / se
    frac = float(np.mean(z > 3.0))
    print(f"fp64 statistics: max z {z.max():.2f}, fraction beyond 3 sigma {frac:.4f}", flush=True)
    assert frac <= 0.03
    assert abs(float(mg.mean() - mo.mean())) < 4 * float(np.sqrt((se ** 2).mean() / se.size)) + 1e-4


def test_render_stop_and_progress(gpu_ctx):
    data = mrt.soup_scene(4, 60000, 600)
    cam = soup_camera(256, 256)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 256, 256, 64, cam, pixel=rnd(1), rr=rnd(2), light=rnd(3),
                           hemi=rnd(4), depth_min=2, depth_max=6)
    gpu = mrt.Scene(gpu_ctx, data)
    gpu.build(abi.BVH)
    result = {}

    def run():
        try:
            gpu.render(desc)
            result["rc"] = 0
        except abi.MrtError as e:
            result["rc"] = e.code

    t = threading.Thread(target=run)
    t.start()
    while gpu.progress() < 2 and t.is_alive():
        pass
    gpu.stop()
    t.join(timeout=60)
    assert not t.is_alive()
    assert result["rc"] in (abi.MRT_E_STOPPED, 0)
    if result["rc"] == abi.MRT_E_STOPPED:
        assert 2 <= gpu.progress() < 64


def test_graph_loop_stop_then_frames(gpu_ctx):
    """A stop request during a graph-loop wavefront frame (the groups' graphs
    re-captured with the stop flag) ends it with its completed samples; the
    scene's next frames, replayed from graphs again, equal the launched
    loop's."""
    data = mrt.soup_scene(4, 60000, 600)
    cam = soup_camera(1024, 576)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 1024, 576, 256, cam, pixel=rnd(1), rr=rnd(2), light=rnd(3),
                           hemi=rnd(4), depth_min=2, depth_max=6, kernel=abi.KERNEL_WAVEFRONT)
    ctx = mrt.Context(0)
    ctx.set_tuning(graph_loop=1)
    gpu = mrt.Scene(ctx, data)
    try:
        gpu.build(abi.BVH)
        result = {}

        def run():
            try:
                gpu.render(desc)
                result["rc"] = 0
            except abi.MrtError as e:
                result["rc"] = e.code

        t = threading.Thread(target=run)
        t.start()
        while gpu.progress() < 2 and t.is_alive():
            pass
        gpu.stop()
        t.join(timeout=60)
        assert not t.is_alive()
        print(f"graph-loop frame: rc {result['rc']}, {gpu.progress()} of 256 samples", flush=True)
        assert result["rc"] in (abi.MRT_E_STOPPED, 0)
        if result["rc"] == abi.MRT_E_STOPPED:
            assert 2 <= gpu.progress() < 256
        short = mrt.render_desc(abi.SHADER_PATHTRACER, 1024, 576, 3, cam, pixel=rnd(1), rr=rnd(2), light=rnd(3),
                                hemi=rnd(4), depth_min=2, depth_max=6, kernel=abi.KERNEL_WAVEFRONT)
        frames = [gpu.render(short, accum=np.zeros((576, 1024, 3), np.float32)) for _ in range(2)]
        ctx.set_tuning(graph_loop=2)
        acc = np.zeros((576, 1024, 3), np.float32)
        launched, _ = gpu.render(short, accum=acc)
        for bm, _ in frames:
            assert np.array_equal(bm, launched)
    finally:
        gpu.close()
        ctx.close()


def test_render_deterministic(gpu_ctx):
    data = mrt.soup_scene(4, 60000, 600)
    cam = soup_camera(128, 128)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, 128, 128, 4, cam, pixel=rnd(1), rr=rnd(2), light=rnd(3),
                           hemi=rnd(4), depth_min=2, depth_max=6)
    gpu = mrt.Scene(gpu_ctx, data)
    gpu.build(abi.BVH)
    a1 = np.zeros((128, 128, 3), np.float32)
    a2 = np.zeros((128, 128, 3), np.float32)
    b1, _ = gpu.render(desc, accum=a1)
    b2, _ = gpu.render(desc, accum=a2)
    assert np.array_equal(b1, b2) and np.array_equal(a1, a2)


@pytest.mark.parametrize("generic,size", [(False, (256, 144)), (True, (256, 144)), (False, (512, 288))])
def test_wavefront_equals_megakernel(gpu_ctx, generic, size):
    """Both kernel shapes evaluate every path with the same operations in the
    same order: bitmaps, FP32 radiance and ray counts are identical (the
    megakernel per sample pass and, on scenes of triangles and lights, as one
    persistent launch; the wavefront with and without reordering and
    refilling lanes).  At 512x288 the wavefront frame runs as three slot
    groups on three streams."""
    W, H = size
    data = mrt.soup_scene(9, 80000, 800)
    if generic:
        # Planes and spheres without reflection/transmission.
        data.planes = np.zeros(1, abi.PLANE_DTYPE)
        data.planes[0]["point"] = (0, -0.01, 0)
        data.planes[0]["normal"] = (0, 1, 0)
        data.planes[0]["material_id"] = 0
        data.spheres = np.zeros(2, abi.SPHERE_DTYPE)
        data.spheres["center"] = [(0.3, 0.4, 0.5), (0.7, 0.6, 0.4)]
        data.spheres["sq_radius"] = np.float32(0.1) * np.float32(0.1)
        data.spheres["material_id"] = 0
        data.materials = np.zeros(1, abi.MATERIAL_DTYPE)
        data.materials[0]["kd"] = (0.5, 0.6, 0.7)
        data.materials[0]["ior"] = 1.0
    cam = soup_camera(W, H)
    res = []
    for kernel, sort, pers in ((abi.KERNEL_MEGAKERNEL, 0, 2), (abi.KERNEL_MEGAKERNEL, 0, 0), (abi.KERNEL_WAVEFRONT, 2, 2),
                               (abi.KERNEL_WAVEFRONT, 1, 2), (abi.KERNEL_WAVEFRONT, 2, 1),
                               (abi.KERNEL_WAVEFRONT, 1, 1)):
        desc = mrt.render_desc(abi.SHADER_PATHTRACER, W, H, 3, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                               rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6, kernel=kernel,
                               ray_sort=sort, persistent=pers)
        gpu = mrt.Scene(gpu_ctx, data)
        gpu.build(abi.BVH)
        acc = np.zeros((H, W, 3), np.float32)
        bm, st = gpu.render(desc, accum=acc)
        res.append((bm, acc, (st.rays_camera, st.rays_extension, st.rays_shadow, st.rays_specular)))
    b0, a0, c0 = res[0]
    for b1, a1, c1 in res[1:]:
        print(f"megakernel vs wavefront rays {c0} {c1}", flush=True)
        assert c0 == c1
        assert np.array_equal(b0, b1)
        assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))


@pytest.mark.parametrize("shader", [abi.SHADER_NOSHADOWS, abi.SHADER_DEPTHMAP, abi.SHADER_DIFFUSE])
@pytest.mark.parametrize("accel", [abi.BVH, abi.REGULAR_GRID, abi.NAIVE])
def test_one_hit_shaders_exact(gpu_ctx, shader, accel):
    """NoShadows, DepthMap and DiffuseMaterial (one closest hit, no
    recursion): bitmap and FP32 radiance bit-exact against the oracle, on the
    Cornell box (two area lights, light samples drawn per pixel) and on the
    mirror scene (planes, spheres, reflective and transmissive materials,
    which DiffuseMaterial reports as Ks / Kt)."""
    for name, data, cam in (("cornell", mrt.cornell_scene(1), cornell_camera(128, 128)),
                            ("mirror", mirror_scene(), cornell_camera(128, 128))):
        desc = mrt.render_desc(shader, 128, 128, 2, cam, samples_light=2, pixel=rnd(3), light=rnd(4),
                               depth_max_point=(1, 1, 1))
        (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, data, desc, accel)
        print(f"{name} shader {shader} accel {accel}: {st.rays_camera} camera rays", flush=True)
        assert (st.rays_camera, st.rays_shadow) == (int(cnt[0]), int(cnt[2]))
        assert np.array_equal(ag.view(np.uint32), ao.astype(np.float32).view(np.uint32)), f"{name}: radiance"
        assert np.array_equal(bg, bo), f"{name}: bitmap"


def test_obj_scene_exact(gpu_ctx):
    """A scene ingested from OBJ / MTL (mirror, glass, two emissive faces,
    vertex normals and colours) renders bit-exactly as the oracle renders the
    same loaded arrays (Whitted with reflection and transmission)."""
    from pathlib import Path
    d = Path(__file__).resolve().parent / "golden" / "obj"
    data = mrt.obj_scene(d / "scene.obj", d / "scene.mtl", mrt.sampler(abi.SAMPLER_CONSTANT, 0.4))
    cam = mrt.make_camera(abi.CAMERA_PERSPECTIVE, (-1.0, 1.0, -3.0), (-1.0, 1.0, 0.0), (0, 1, 0), 45.0, 45.0)
    desc = mrt.render_desc(abi.SHADER_WHITTED, 64, 64, 1, cam, depth_max=4)
    (bg, ag, st), (bo, ao, cnt) = render_both(gpu_ctx, data, desc)
    print(f"obj scene: {st.rays_camera} camera {st.rays_specular} specular rays", flush=True)
    assert (ag > 0).any()
    np.testing.assert_allclose(ag, ao, rtol=1e-3, atol=1e-5)
    assert np.max(np.abs(bg.view(np.uint8).astype(int) - bo.view(np.uint8).astype(int))) <= 1


@pytest.mark.parametrize("spp", [1, 3])
def test_aovs_against_one_hit_shaders(gpu_ctx, spp):
    """mrt_render_aovs (no reference counterpart) pinned against the
    oracle-pinned one-hit shaders on the same camera rays: the mean albedo is
    the mean DiffuseMaterial radiance (Kd, colour or Le first), the mean depth
    reproduces the DepthMap grey (maxDist - t) / maxDist, normals are unit
    vectors at 1 spp."""
    data = mrt.cornell_scene(1)
    cam = cornell_camera(96, 96)
    gpu = mrt.Scene(gpu_ctx, data)
    gpu.build(abi.BVH)
    kw = dict(pixel=rnd(3), depth_max_point=(1, 1, 1))
    aov = gpu.aovs(mrt.render_desc(abi.SHADER_DIFFUSE, 96, 96, spp, cam, **kw))
    acc = np.zeros((96, 96, 3), np.float32)
    gpu.render(mrt.render_desc(abi.SHADER_DIFFUSE, 96, 96, spp, cam, **kw), accum=acc)
    np.testing.assert_allclose(aov[..., :3], acc / np.float32(spp), rtol=1e-6, atol=1e-7)
    if spp == 1:
        assert np.array_equal(aov[..., :3], acc)
        n = np.linalg.norm(aov[..., 3:6], axis=2)
        np.testing.assert_allclose(n[aov[..., 6] > 0], 1.0, rtol=1e-5)
        depth_acc = np.zeros((96, 96, 3), np.float32)
        gpu.render(mrt.render_desc(abi.SHADER_DEPTHMAP, 96, 96, 1, cam, **kw), accum=depth_acc)
        pos = np.array(cam.position, np.float32)
        max_dist = np.float32(np.linalg.norm(np.float32(1) - pos) * 2)
        grey = np.maximum((max_dist - aov[..., 6]) / max_dist, 0)
        hit = aov[..., 6] > 0
        np.testing.assert_allclose(grey[hit], depth_acc[..., 0][hit], rtol=1e-5)


@pytest.mark.parametrize("kernel,spp,first", [(abi.KERNEL_WAVEFRONT, 4, 0), (abi.KERNEL_WAVEFRONT, 1, 0),
                                              (abi.KERNEL_WAVEFRONT, 7, 2), (abi.KERNEL_MEGAKERNEL, 2, 0)])
def test_render_with_aovs(gpu_ctx, kernel, spp, first):
    """mrt_render_with_aovs against mrt_render and mrt_render_aovs of the same
    desc: bitmap, accumulation and AOVs bit-identical.  The wavefront frames
    record the first hits while shading (4 and 5 samples: sample-parallel
    slots; 1 sample: one slot per pixel); the megakernel runs the AOV kernel
    after the frame."""
    data = mrt.soup_scene(12, 60_000, 600)
    cam = soup_camera(320, 180)
    gpu = mrt.Scene(gpu_ctx, data)
    try:
        gpu.build(abi.BVH)
        desc = mrt.render_desc(abi.SHADER_PATHTRACER, 320, 180, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, 3),
                               rr=rnd(4), light=rnd(5), hemi=rnd(6), depth_min=2, depth_max=6, kernel=kernel,
                               first_sample=first)
        acc1 = np.zeros((180, 320, 3), np.float32)
        bm1, _ = gpu.render(desc, accum=acc1)
        aov1 = gpu.aovs(desc)
        acc2 = np.zeros((180, 320, 3), np.float32)
        bm2, aov2, st = gpu.render_with_aovs(desc, accum=acc2)
    finally:
        gpu.close()
    assert np.array_equal(bm1, bm2)
    assert np.array_equal(acc1, acc2)
    assert np.array_equal(aov1, aov2)
    assert 0.05 < (aov2[..., 6] > 0).mean() < 0.95  # the sparse soup: hits and misses


@pytest.mark.parametrize("size,spp,generic", [((128, 80), 3, False), ((512, 288), 2, False), ((96, 96), 4, True),
                                              ((1280, 720), 16, False)])
def test_graph_loop_equals_launched_loop(size, spp, generic):
    """The wavefront loop replayed as captured HIP graphs (each slot group's
    batch of four iterations, captured and replayed on the group's own
    stream) gives the frame the launched loop gives: bitmap, FP32 radiance
    and ray counts identical; 512x288 runs three slot groups; 1280x720 at 16
    samples runs three groups with two samples per path slot, after uploads
    and a 177 MB clear that each group's graph must wait for behind its
    group's counter clear and start kernel."""
    W, H = size
    data = mrt.soup_scene(9, 80000, 800)
    if generic:
        data.planes = np.zeros(1, abi.PLANE_DTYPE)
        data.planes[0]["point"] = (0, -0.01, 0)
        data.planes[0]["normal"] = (0, 1, 0)
        data.planes[0]["material_id"] = -1
    cam = soup_camera(W, H)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                           rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6,
                           kernel=abi.KERNEL_WAVEFRONT)
    res = []
    for mode in (1, 2):
        ctx = mrt.Context(0)
        ctx.set_tuning(graph_loop=mode)
        sc = mrt.Scene(ctx, data)
        try:
            sc.build(abi.BVH)
            acc = np.zeros((H, W, 3), np.float32)
            bm, st = sc.render(desc, accum=acc)
            bm2, st2 = sc.render(desc, accum=np.zeros((H, W, 3), np.float32))  # a second frame reuses nothing stale
        finally:
            sc.close()
            ctx.close()
        res.append((bm, acc, (st.rays_camera, st.rays_extension, st.rays_shadow), st.trace_groups, st.slot_lanes))
        assert np.array_equal(bm, bm2)
    (b1, a1, c1, g1, l1), (b2, a2, c2, g2, l2) = res
    print(f"graph loop {size} {spp} spp: rays {c1} / {c2}, {g1} groups, {l1} slots per pixel", flush=True)
    if size == (1280, 720):
        assert g1 == 3 and l1 * 2 == spp
    assert c1 == c2 and g1 == g2
    assert np.array_equal(b1, b2)
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))


@pytest.mark.parametrize("spp", [1, 3])
def test_grid_persistent_equals_one_ray_per_lane(gpu_ctx, spp):
    """RegularGrid scenes: the persistent refilling grid walk, the one-ray-
    per-lane wavefront trace and the megakernel give identical frames and
    ray counts (512x288: three slot groups)."""
    W, H = 512, 288
    data = mrt.soup_scene(13, 120000, 1200)
    cam = soup_camera(W, H)
    cmin, cmax = mrt.camera_aabb(cam)
    gpu = mrt.Scene(gpu_ctx, data)
    res = []
    try:
        gpu.build(abi.REGULAR_GRID, cmin, cmax)
        for kernel, pers in ((abi.KERNEL_MEGAKERNEL, 0), (abi.KERNEL_WAVEFRONT, 2), (abi.KERNEL_WAVEFRONT, 1)):
            desc = mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                                   rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6, kernel=kernel,
                                   persistent=pers)
            acc = np.zeros((H, W, 3), np.float32)
            bm, st = gpu.render(desc, accum=acc)
            res.append((bm, acc, (st.rays_camera, st.rays_extension, st.rays_shadow)))
    finally:
        gpu.close()
    b0, a0, c0 = res[0]
    for b1, a1, c1 in res[1:]:
        print(f"grid frames: rays {c0} {c1}", flush=True)
        assert c0 == c1
        assert np.array_equal(b0, b1)
        assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))


@pytest.mark.parametrize("spp,tiles", [(1, None), (3, None), (9, None), (4, (0, 17, 120, 255))])
def test_persistent_megakernel_equals_per_pass(gpu_ctx, spp, tiles):
    """The persistent megakernel (one launch per frame; 1, 2 or 8 sample
    lanes per pixel, the latter two through the per-sample radiance and the
    resolve) against the per-pass megakernel: identical frames and counts,
    also on a tile subset."""
    W, H = 320, 176
    data = mrt.soup_scene(14, 100000, 1000)
    cam = soup_camera(W, H)
    gpu = mrt.Scene(gpu_ctx, data)
    res = []
    try:
        gpu.build(abi.BVH)
        for pers in (2, 0):
            desc = mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                                   rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6,
                                   kernel=abi.KERNEL_MEGAKERNEL, persistent=pers, tiles=tiles)
            acc = np.zeros((H, W, 3), np.float32)
            bm, st = gpu.render(desc, accum=acc)
            res.append((bm, acc, (st.rays_camera, st.rays_extension, st.rays_shadow), st.samples_done))
    finally:
        gpu.close()
    (b0, a0, c0, d0), (b1, a1, c1, d1) = res
    print(f"megakernel per pass / persistent: rays {c0} {c1}", flush=True)
    assert c0 == c1 and d0 == d1 == spp
    assert np.array_equal(b0, b1)
    assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
