"""Render server (continuous batching, mrt_server_*) through the C ABI.

Every job's RGBA8 bitmap and FP32 accumulation must be bit-identical to
mrt_render of the same request alone (which the render tests pin to the
oracle): a pixel's samples take the same draws and the same operations
whatever else shares the launches.  Mixes cover several scenes (light tree and
linear light scan), resolutions, sample counts and progressive resumes, jobs
admitted together and while others run, and a slot pool small enough that
jobs wait for slots to free up; these run every job on the pipelines.
The idle-direct option (a lone small job rendered through mrt_render by an
idle server) is checked the same way.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import DeviceArray, cornell_camera, soup_camera
from mobileraytracer_amd import abi

pytestmark = pytest.mark.gpu

RANDOM = abi.SAMPLER_RANDOM


def rnd(seed):
    return mrt.sampler(RANDOM, 0.5, seed)


@pytest.fixture(scope="module")
def scenes(gpu_ctx):
    # The first scene's triangle trees are built and written on the device
    # (GPU builder), the others' on the host: the server relocates both
    # kinds into its arena.
    datas = [(mrt.soup_scene(11, 40000, 400), abi.BUILDER_GPU),   # 400 area lights: light tree
             (mrt.soup_scene(12, 20000, 8), abi.BUILDER_HOST),    # 8 area lights: linear scan
             (mrt.cornell_scene(1), abi.BUILDER_HOST)]
    out = []
    for d, builder in datas:
        s = mrt.Scene(gpu_ctx, d)
        s.build(abi.BVH, builder=builder)
        out.append(s)
    return out


def make_desc(scene_idx, w, h, spp, seed, first=0, pixel_kind=abi.SAMPLER_HALTON):
    cam = cornell_camera(w, h) if scene_idx == 2 else soup_camera(w, h)
    return mrt.render_desc(abi.SHADER_PATHTRACER, w, h, spp, cam, pixel=mrt.sampler(pixel_kind, 0, seed),
                           rr=rnd(seed + 1), light=rnd(seed + 2), hemi=rnd(seed + 3), depth_min=2, depth_max=6,
                           first_sample=first, kernel=abi.KERNEL_WAVEFRONT)


def alone(scene, desc, bitmap=None, accum=None):
    bm = np.zeros((desc.height, desc.width), np.uint32) if bitmap is None else bitmap.copy()
    acc = np.zeros((desc.height, desc.width, 3), np.float32) if accum is None else accum.copy()
    bm, st = scene.render(desc, bitmap=bm, accum=acc)
    return bm, acc, st


def check_same(tag, got_bm, got_acc, want_bm, want_acc):
    print(f"{tag}: comparing {got_bm.shape}", flush=True)
    assert np.array_equal(got_bm, want_bm), f"{tag}: bitmap differs in {int((got_bm != want_bm).sum())} pixels"
    assert np.array_equal(got_acc.view(np.uint32), want_acc.view(np.uint32)), f"{tag}: accumulation differs"


# (scene, width, height, spp, seed)
MIX = [(0, 128, 80, 4, 5), (1, 256, 144, 2, 9), (2, 64, 64, 8, 13), (0, 64, 48, 3, 17), (1, 96, 64, 5, 21),
       (2, 128, 128, 2, 25), (0, 256, 144, 1, 29), (1, 160, 96, 3, 33)]


@pytest.mark.parametrize("pipelines,max_slots", [(3, 0), (3, 64 * 3 * 200), (1, 64 * 600)])
def test_server_mix_equals_render_alone(gpu_ctx, scenes, pipelines, max_slots):
    """All jobs submitted at once; with the small pools most of them wait
    for waves that finished jobs free up."""
    srv = mrt.RenderServer(gpu_ctx, max_slots=max_slots, pipelines=pipelines, job_memory=64 << 20,
                           idle_direct=False)
    try:
        ids = [srv.add_scene(s) for s in scenes]
        jobs = []
        for (si, w, h, spp, seed) in MIX:
            desc = make_desc(si, w, h, spp, seed)
            acc = np.zeros((h, w, 3), np.float32)
            jobs.append((si, desc, srv.submit(ids[si], desc, accum=acc)))
        print(f"submitted {len(jobs)} jobs (pipelines {pipelines}, slots {max_slots})", flush=True)
        rays = np.zeros(3, np.int64)
        lanes = []
        for k, (si, desc, jid) in enumerate(jobs):
            bm, acc, info = srv.wait(jid)
            assert info.state == abi.JOB_DONE and info.status == abi.MRT_OK
            assert 1 <= info.lanes <= min(8, desc.samples_pixel)
            lanes.append(info.lanes)
            assert info.t_submit_us <= info.t_admit_us <= info.t_gpu_done_us <= info.t_done_us
            wb, wa, st = alone(scenes[si], desc)
            rays += (st.rays_camera, st.rays_extension, st.rays_shadow)
            check_same(f"job {k} scene {si} {desc.width}x{desc.height} spp {desc.samples_pixel}", bm, acc, wb, wa)
        stt = srv.stats()
        print(f"server rays {stt.rays_camera} {stt.rays_extension} {stt.rays_shadow}, alone {rays.tolist()}, "
              f"batches {stt.batches}", flush=True)
        assert (stt.rays_camera, stt.rays_extension, stt.rays_shadow) == tuple(int(r) for r in rays)
        assert stt.jobs_done == len(MIX) and stt.jobs_failed == 0 and stt.slots_used == 0
        print(f"slots per pixel of the jobs: {lanes}", flush=True)
        if max_slots == 0:
            # The default pool: multi-sample jobs ran sample-parallel.
            assert max(lanes) > 1
    finally:
        srv.close()


def test_server_staggered_and_progressive(gpu_ctx, scenes):
    """Jobs arriving while others run, and a frame rendered as two
    progressive requests (samples 0-1, then 2-4 on the same buffers), equal
    the frame rendered alone in one request."""
    srv = mrt.RenderServer(gpu_ctx, max_slots=64 * 3 * 300, job_memory=64 << 20, idle_direct=False)
    try:
        ids = [srv.add_scene(s) for s in scenes]
        pending = []
        for (si, w, h, spp, seed) in MIX[:5]:
            desc = make_desc(si, w, h, spp, seed + 100)
            acc = np.zeros((h, w, 3), np.float32)
            pending.append((si, desc, srv.submit(ids[si], desc, accum=acc)))
            time.sleep(0.003)
        # Progressive: two requests on the same buffers.
        W, H = 128, 80
        first = make_desc(0, W, H, 2, 41)
        rest = make_desc(0, W, H, 5, 41, first=2)
        bm = np.zeros((H, W), np.uint32)
        acc = np.zeros((H, W, 3), np.float32)
        bm, acc, info = srv.wait(srv.submit(ids[0], first, bitmap=bm, accum=acc))
        bm, acc, info = srv.wait(srv.submit(ids[0], rest, bitmap=bm, accum=acc))
        wb, wa, _ = alone(scenes[0], make_desc(0, W, H, 5, 41))
        check_same("progressive 0-1 + 2-4", bm, acc, wb, wa)
        for k, (si, desc, jid) in enumerate(pending):
            gb, ga, info = srv.wait(jid)
            wb, wa, _ = alone(scenes[si], desc)
            check_same(f"staggered job {k}", gb, ga, wb, wa)
    finally:
        srv.close()


def test_server_device_buffers(gpu_ctx, scenes):
    """Jobs writing into caller device memory give the host-memory result."""
    srv = mrt.RenderServer(gpu_ctx, job_memory=16 << 20, idle_direct=False)
    dbm = dacc = None
    try:
        sid = srv.add_scene(scenes[1])
        desc = make_desc(1, 256, 144, 3, 51)
        host_bm = np.zeros((144, 256), np.uint32)
        host_acc = np.zeros((144, 256, 3), np.float32)
        dbm = DeviceArray(host_bm.nbytes)
        dacc = DeviceArray(host_acc.nbytes)
        dbm.upload(host_bm)
        dacc.upload(host_acc)
        jid = srv.submit(sid, desc, bitmap=dbm, accum=dacc, mem=abi.MEM_DEVICE)
        srv.wait(jid)
        got_bm = dbm.download(host_bm)
        got_acc = dacc.download(host_acc)
        wb, wa, _ = alone(scenes[1], desc)
        check_same("device buffers", got_bm, got_acc, wb, wa)
    finally:
        srv.close()
        for d in (dbm, dacc):
            if d is not None:
                d.free()


def test_server_cancel_and_errors(gpu_ctx, scenes):
    srv = mrt.RenderServer(gpu_ctx, max_slots=64 * 3 * 100, job_memory=64 << 20, idle_direct=False)
    try:
        sid = srv.add_scene(scenes[0])
        # Unknown scene, a Whitted request, a frame larger than the pool.
        with pytest.raises(abi.MrtError) as e:
            srv.submit(sid + 5, make_desc(0, 64, 48, 1, 1))
        assert e.value.code == abi.MRT_E_INVALID
        whitted = make_desc(0, 64, 48, 1, 1)
        whitted.shader = abi.SHADER_WHITTED
        with pytest.raises(abi.MrtError) as e:
            srv.submit(sid, whitted)
        assert e.value.code == abi.MRT_E_UNSUPPORTED
        with pytest.raises(abi.MrtError) as e:
            srv.submit(sid, make_desc(0, 512, 288, 1, 1))
        assert e.value.code == abi.MRT_E_BUDGET
        # A resident scene cannot be destroyed.
        assert abi.lib().mrt_scene_destroy(scenes[0].h) == abi.MRT_E_INVALID
        # Cancel: a long job and one queued behind it.
        long_desc = make_desc(0, 128, 80, 2000, 3)
        j1 = srv.submit(sid, long_desc)
        j2 = srv.submit(sid, make_desc(0, 128, 80, 2000, 4))
        time.sleep(0.05)
        srv.cancel(j2)
        srv.cancel(j1)
        for j in (j1, j2):
            with pytest.raises(abi.MrtError) as e:
                srv.wait(j, timeout_us=60_000_000)
            assert e.value.code == abi.MRT_E_STOPPED
        # The server keeps serving.
        desc = make_desc(0, 64, 48, 2, 61)
        bm, acc, info = srv.wait(srv.submit(sid, desc, accum=np.zeros((48, 64, 3), np.float32)))
        wb, wa, _ = alone(scenes[0], desc)
        check_same("after cancel", bm, acc, wb, wa)
        st = srv.stats()
        assert st.jobs_failed == 2 and st.jobs_running == 0 and st.jobs_queued == 0
    finally:
        srv.close()


def test_server_idle_direct(gpu_ctx, scenes):
    """With idle_direct on: a lone small job at an idle server
    is rendered alone through mrt_render (info.direct), host and device
    buffers; a job at or above 6 x 2^20 pixel samples, and jobs queued
    together, run on the pipelines; every result and the server's ray counts
    equal mrt_render alone; a cancel during a direct render completes it and
    fails it with MRT_E_STOPPED."""
    srv = mrt.RenderServer(gpu_ctx, job_memory=256 << 20, idle_direct=True)
    dbm = None
    try:
        ids = [srv.add_scene(s) for s in scenes]
        rays = np.zeros(3, np.int64)
        # Lone jobs, one at a time: direct.
        for k, (si, w, h, spp, seed) in enumerate(MIX[:4]):
            desc = make_desc(si, w, h, spp, seed + 200)
            bm, acc, info = srv.wait(srv.submit(ids[si], desc, accum=np.zeros((h, w, 3), np.float32)))
            assert info.state == abi.JOB_DONE and info.direct == 1 and info.lanes == 1
            assert info.t_submit_us <= info.t_admit_us <= info.t_gpu_done_us <= info.t_done_us
            wb, wa, st = alone(scenes[si], desc)
            rays += (st.rays_camera, st.rays_extension, st.rays_shadow)
            check_same(f"direct job {k}", bm, acc, wb, wa)
        # Device buffers, direct.
        desc = make_desc(1, 256, 144, 3, 251)
        host_bm = np.zeros((144, 256), np.uint32)
        dbm = DeviceArray(host_bm.nbytes)
        dbm.upload(host_bm)
        info = srv.wait(srv.submit(ids[1], desc, bitmap=dbm, mem=abi.MEM_DEVICE))[2]
        assert info.direct == 1
        wb, _, st = alone(scenes[1], desc)
        rays += (st.rays_camera, st.rays_extension, st.rays_shadow)
        assert np.array_equal(dbm.download(host_bm), wb)
        # 2048x1024 at 3 samples is 6.3M pixel samples: the pipelines.
        big = make_desc(0, 2048, 1024, 3, 261)
        bm, acc, info = srv.wait(srv.submit(ids[0], big, accum=np.zeros((1024, 2048, 3), np.float32)))
        assert info.direct == 0
        wb, wa, st = alone(scenes[0], big)
        rays += (st.rays_camera, st.rays_extension, st.rays_shadow)
        check_same("above the megakernel size", bm, acc, wb, wa)
        # A burst: the jobs that queue together go to the pipelines.
        jobs = []
        for (si, w, h, spp, seed) in MIX:
            desc = make_desc(si, w, h, spp, seed + 300)
            jobs.append((si, desc, srv.submit(ids[si], desc, accum=np.zeros((h, w, 3), np.float32))))
        direct = []
        for k, (si, desc, jid) in enumerate(jobs):
            bm, acc, info = srv.wait(jid)
            direct.append(info.direct)
            wb, wa, st = alone(scenes[si], desc)
            rays += (st.rays_camera, st.rays_extension, st.rays_shadow)
            check_same(f"burst job {k}", bm, acc, wb, wa)
        print(f"burst: direct flags {direct}", flush=True)
        assert sum(direct) < len(direct)
        stt = srv.stats()
        assert (stt.rays_camera, stt.rays_extension, stt.rays_shadow) == tuple(int(r) for r in rays)
        # Cancel while a direct render runs: it completes, then fails.
        desc = make_desc(0, 256, 144, 160, 271)  # 5.9M pixel samples
        jid = srv.submit(ids[0], desc, accum=np.zeros((144, 256, 3), np.float32))
        time.sleep(0.002)
        srv.cancel(jid)
        with pytest.raises(abi.MrtError) as e:
            srv.wait(jid, timeout_us=60_000_000)
        assert e.value.code == abi.MRT_E_STOPPED
        st = srv.stats()
        assert st.jobs_running == 0 and st.jobs_queued == 0 and st.slots_used == 0
    finally:
        srv.close()
        if dbm is not None:
            dbm.free()
