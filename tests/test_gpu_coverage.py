"""GPU tests of the parts around the hot path that the other files do not
reach:

  * the Orthographic camera (Orthographic.cpp:23-45) on every accelerator:
    its rays, and its getAABB feeding the RegularGrid bounds
    (Shader.cpp:52-62), bit-exact against the oracle;
  * scene and render-server admission under mrt_ctx_set_budget (the memory
    check of DrawView.cpp:391-416): refused with MRT_E_BUDGET before any
    device allocation, and the context stays usable;
  * render-server jobs against the oracle itself (the server tests compare
    them with mrt_render alone);
  * the batched workload (64 scenes of 10k-4M triangles, one request each,
    all submitted at once) and the serving workload (a Poisson stream over a
    resident pool of 16 scenes of 50k-8M triangles) at their full sizes: a
    sample of the jobs, the largest included, is bit-identical to the same
    request rendered alone.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import DeviceArray, cornell_camera, oracle_scene, soup_camera
from mobileraytracer_amd import abi, serving

pytestmark = pytest.mark.gpu


def rnd(seed):
    return mrt.sampler(abi.SAMPLER_RANDOM, 0.5, seed)


def mem_free() -> int:
    if DeviceArray._hip is None:
        DeviceArray(256).free()  # loads the runtime the library uses
    free, total = C.c_size_t(), C.c_size_t()
    assert DeviceArray._hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


# ----------------------------------------------------------- Orthographic --
@pytest.mark.parametrize("accel", [abi.BVH, abi.REGULAR_GRID, abi.NAIVE])
def test_orthographic_camera_exact(gpu_ctx, accel):
    """Whitted Cornell box through an orthographic camera whose AABB widens
    the grid bounds in x and z (and whose vertical order is inverted):
    bitmap and FP32 radiance bit-exact, equal ray counts."""
    cam = mrt.make_camera(abi.CAMERA_ORTHOGRAPHIC, (0, 0, -3.4), (0, 0, 1), (0, 1, 0), 3.0, 2.5)
    cmin, cmax = mrt.camera_aabb(cam)
    assert cmin[0] < -1.0 and cmax[0] > 1.0 and cmin[1] > cmax[1]
    data = mrt.cornell_scene(1)
    desc = mrt.render_desc(abi.SHADER_WHITTED, 128, 128, 2, cam, pixel=rnd(3))
    gpu = mrt.Scene(gpu_ctx, data)
    try:
        gpu.build(accel, cmin, cmax)
        acc_g = np.zeros((128, 128, 3), np.float32)
        bm_g, st = gpu.render(desc, accum=acc_g)
    finally:
        gpu.close()
    orc = oracle_scene(data, accel, cmin, cmax)
    if accel == abi.REGULAR_GRID:
        gs, bounds, _, _ = orc.grid(abi.CAT_TRIANGLE)
        assert bounds[0] <= cmin[0] and bounds[2] <= cmin[2] and bounds[3] >= cmax[0]
    acc_o = np.zeros((128, 128, 3), np.float64)
    bm_o, cnt, _ = orc.render(desc, accum=acc_o, threads=8)
    assert (st.rays_camera, st.rays_extension, st.rays_shadow, st.rays_specular) == tuple(int(c) for c in cnt)
    assert np.array_equal(acc_g.view(np.uint32), acc_o.astype(np.float32).view(np.uint32))
    assert np.array_equal(bm_g, bm_o)
    assert (acc_g > 0).any(axis=2).mean() > 0.5


def test_orthographic_pathtracer_soup(gpu_ctx):
    """PathTracer through an orthographic camera on both kernel shapes and
    both accelerators, against the oracle (1e-3 relative, bitmap within 1)."""
    data = mrt.soup_scene(21, 40000, 400)
    cam = mrt.make_camera(abi.CAMERA_ORTHOGRAPHIC, (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (0, 1, 0), 1.2, 1.2)
    cmin, cmax = mrt.camera_aabb(cam)
    for accel, kernel in ((abi.BVH, abi.KERNEL_WAVEFRONT), (abi.REGULAR_GRID, abi.KERNEL_MEGAKERNEL)):
        desc = mrt.render_desc(abi.SHADER_PATHTRACER, 96, 96, 3, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                               rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6, kernel=kernel)
        gpu = mrt.Scene(gpu_ctx, data)
        try:
            gpu.build(accel, cmin, cmax)
            acc_g = np.zeros((96, 96, 3), np.float32)
            bm_g, st = gpu.render(desc, accum=acc_g)
        finally:
            gpu.close()
        orc = oracle_scene(data, accel, cmin, cmax)
        acc_o = np.zeros((96, 96, 3), np.float64)
        bm_o, cnt, _ = orc.render(desc, accum=acc_o, threads=16)
        assert (st.rays_camera, st.rays_extension, st.rays_shadow) == tuple(int(c) for c in cnt[:3])
        assert np.isclose(acc_g, acc_o, rtol=1e-3, atol=1e-5).all()
        d8 = np.abs(bm_g.view(np.uint8).astype(int) - bm_o.view(np.uint8).astype(int))
        assert d8.max() <= 1


# ---------------------------------------------------------------- budget --
def test_budget_refusals_allocate_nothing():
    ctx = mrt.Context(0)
    data = mrt.soup_scene(22, 20000, 200)
    mem_free()
    try:
        # Scene creation over budget.
        ctx.set_budget(1 << 20)
        f0 = mem_free()
        with pytest.raises(abi.MrtError) as e:
            mrt.Scene(ctx, data)
        assert e.value.code == abi.MRT_E_BUDGET
        assert mem_free() >= f0  # nothing allocated (the runtime may release memory meanwhile)
        # The geometry fits, its BVH does not.
        ctx.set_budget(20000 * 128 + 400 * 300 + (1 << 16))
        sc = mrt.Scene(ctx, data)
        f0 = mem_free()
        for builder in (abi.BUILDER_HOST, abi.BUILDER_GPU):
            with pytest.raises(abi.MrtError) as e:
                sc.build(abi.BVH, builder=builder)
            assert e.value.code == abi.MRT_E_BUDGET
        with pytest.raises(abi.MrtError) as e:
            sc.build(abi.REGULAR_GRID)
        assert e.value.code == abi.MRT_E_BUDGET
        assert mem_free() >= f0  # nothing allocated (the runtime may release memory meanwhile)
        # A render server larger than the budget.
        with pytest.raises(abi.MrtError) as e:
            mrt.RenderServer(ctx)
        assert e.value.code == abi.MRT_E_BUDGET
        assert mem_free() >= f0  # nothing allocated (the runtime may release memory meanwhile)
        # Lifting the budget: the same scene builds and renders as on a
        # context that never had one.
        ctx.set_budget(0)
        sc.build(abi.BVH)
        desc = mrt.render_desc(abi.SHADER_PATHTRACER, 64, 64, 2, soup_camera(64, 64), rr=rnd(1), light=rnd(2),
                               hemi=rnd(3), depth_min=2, depth_max=6, kernel=abi.KERNEL_WAVEFRONT)
        bm, _ = sc.render(desc)
        ref = mrt.Scene(mrt.Context(0), data)
        ref.build(abi.BVH)
        bm2, _ = ref.render(desc)
        assert np.array_equal(bm, bm2)
        ref.close()
        sc.close()
    finally:
        ctx.close()


# -------------------------------------------------- server against oracle --
def test_server_jobs_match_oracle(gpu_ctx):
    datas = [mrt.soup_scene(11, 40000, 400), mrt.soup_scene(12, 20000, 8), mrt.cornell_scene(1)]
    scenes = []
    for d in datas:
        s = mrt.Scene(gpu_ctx, d)
        s.build(abi.BVH)
        scenes.append(s)
    mix = [(0, 128, 80, 4, 5), (1, 256, 144, 2, 9), (2, 64, 64, 8, 13), (0, 64, 48, 3, 17), (1, 96, 64, 5, 21),
           (2, 128, 128, 2, 25), (0, 256, 144, 1, 29), (1, 160, 96, 3, 33)]
    srv = mrt.RenderServer(gpu_ctx, job_memory=64 << 20)
    try:
        ids = [srv.add_scene(s) for s in scenes]
        jobs = []
        for si, w, h, spp, seed in mix:
            cam = cornell_camera(w, h) if si == 2 else soup_camera(w, h)
            desc = mrt.render_desc(abi.SHADER_PATHTRACER, w, h, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, seed),
                                   rr=rnd(seed + 1), light=rnd(seed + 2), hemi=rnd(seed + 3), depth_min=2,
                                   depth_max=6, kernel=abi.KERNEL_WAVEFRONT)
            acc = np.zeros((h, w, 3), np.float32)
            jobs.append((si, desc, srv.submit(ids[si], desc, accum=acc)))
        orcs = [oracle_scene(d, abi.BVH) for d in datas]
        for k, (si, desc, jid) in enumerate(jobs):
            bm, acc, info = srv.wait(jid)
            acc_o = np.zeros((desc.height, desc.width, 3), np.float64)
            bm_o, cnt, _ = orcs[si].render(desc, accum=acc_o, threads=16)
            print(f"job {k}: scene {si} {desc.width}x{desc.height} {desc.samples_pixel} spp, {info.lanes} slots "
                  f"per pixel", flush=True)
            assert np.isclose(acc, acc_o, rtol=1e-3, atol=1e-5).all(), f"job {k}: radiance"
            d8 = np.abs(bm.view(np.uint8).astype(int) - bm_o.view(np.uint8).astype(int))
            assert d8.max() <= 1, f"job {k}: bitmap"
    finally:
        srv.close()
        for s in scenes:
            s.close()


# --------------------------------------------- full-size batched / serving --
def alone_equal(tag, scene, desc, bm, acc):
    wb = np.zeros((desc.height, desc.width), np.uint32)
    wa = np.zeros((desc.height, desc.width, 3), np.float32)
    scene.render(desc, bitmap=wb, accum=wa)
    print(f"{tag}: {desc.width}x{desc.height} {desc.samples_pixel} spp", flush=True)
    assert np.array_equal(bm, wb), f"{tag}: bitmap"
    assert np.array_equal(acc.view(np.uint32), wa.view(np.uint32)), f"{tag}: accumulation"


def test_batched_64_scenes_full_size(gpu_ctx):
    """The batched workload as the benchmark runs it (64 resident scenes,
    36M triangles, every request submitted at once, largest first): sampled
    jobs equal their requests rendered alone."""
    t0 = time.time()
    pool = serving.build_pool(gpu_ctx, serving.pool_sizes(64, 10_000, 4_000_000, seed=3), seed=300)
    print(f"pool of {len(pool)} scenes, {sum(p.triangles for p in pool)} triangles, {time.time() - t0:.1f} s",
          flush=True)
    reqs = serving.batch_requests(pool, seed=3)
    order = sorted(range(len(reqs)), key=lambda i: -int(reqs[i][1].width) * int(reqs[i][1].height) *
                   int(reqs[i][1].samples_pixel))
    srv = mrt.RenderServer(gpu_ctx, job_memory=4 << 30)
    try:
        for p in pool:
            p.server_id = srv.add_scene(p.scene)
        sample = {order[0], order[1], order[len(order) // 2], order[-1], 7, 30}
        out = {}
        t0 = time.time()
        for i in order:
            si, desc = reqs[i]
            bm = np.zeros((desc.height, desc.width), np.uint32)
            acc = np.zeros((desc.height, desc.width, 3), np.float32) if i in sample else None
            out[i] = (desc, srv.submit(pool[si].server_id, desc, bitmap=bm, accum=acc), bm, acc)
        results = {}
        for i in order:
            desc, jid, _, _ = out[i]
            bm, acc, info = srv.wait(jid)
            assert info.state == abi.JOB_DONE
            if i in sample:
                results[i] = (bm, acc)
        print(f"64 jobs in {time.time() - t0:.2f} s", flush=True)
    finally:
        srv.close()
    for i in sorted(results):
        si, desc = reqs[i]
        alone_equal(f"batched job {i} (scene {si})", pool[si].scene, desc, *results[i])
    for p in pool:
        p.scene.close()


def test_serving_pool_poisson_full_size(gpu_ctx):
    """The serving workload: 16 resident scenes (50k-8M triangles), requests
    arriving as a Poisson stream while others run; sampled jobs equal their
    requests rendered alone."""
    t0 = time.time()
    pool = serving.build_pool(gpu_ctx, serving.pool_sizes(16, 50_000, 8_000_000, seed=4), seed=400)
    print(f"pool of {len(pool)} scenes, {sum(p.triangles for p in pool)} triangles, {time.time() - t0:.1f} s",
          flush=True)
    srv = mrt.RenderServer(gpu_ctx)
    rng = np.random.default_rng(4)
    try:
        for p in pool:
            p.server_id = srv.add_scene(p.scene)
        n = 48
        gaps = rng.exponential(1.0 / 20.0, n)
        reqs = [(int(rng.integers(len(pool))), serving.request_desc(rng, 4 * 7919 + 16 * k)) for k in range(n)]
        sizes = [int(d.width) * int(d.height) * int(d.samples_pixel) for _, d in reqs]
        sample = set(range(0, n, 8)) | {int(np.argmax(sizes))}
        jobs = []
        t0 = time.perf_counter()
        for k, (si, desc) in enumerate(reqs):
            time.sleep(float(gaps[k]))
            bm = np.zeros((desc.height, desc.width), np.uint32)
            acc = np.zeros((desc.height, desc.width, 3), np.float32) if k in sample else None
            jobs.append(srv.submit(pool[si].server_id, desc, bitmap=bm, accum=acc))
        results = {}
        overlapped = 0
        for k, jid in enumerate(jobs):
            bm, acc, info = srv.wait(jid)
            assert info.state == abi.JOB_DONE
            if k in sample:
                results[k] = (bm, acc)
        stt = srv.stats()
        print(f"{n} requests at 20/s served in {time.perf_counter() - t0:.2f} s, {stt.batches} batches",
              flush=True)
    finally:
        srv.close()
    for k in sorted(results):
        si, desc = reqs[k]
        alone_equal(f"served request {k} (scene {si})", pool[si].scene, desc, *results[k])
    for p in pool:
        p.scene.close()
