"""Serving workload for the render server: a resident pool of seeded scenes
and a Poisson stream of render requests, with per-request latency.

Pool: `n` scenes with triangle counts log-uniform in [lo, hi], alternately a
uniform soup (as in the 1M-triangle benchmark scene) and a cluster of
displaced icospheres, 1% of the triangles emissive (area lights), BVH.

Request: a scene of the pool (uniform), a camera pose on a shell around the
unit cube looking at its centre, a resolution of {256x256, 640x480, 1280x720,
1920x1080} and a sample count of {1, 2, 4, ..., 64} (uniform), PathTracer with
maximum depth 6 and Russian roulette from depth 3, FP32 accumulation kept.
Latency is the server's own clock from submission to the result being in
the caller's buffers.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import RenderServer, Scene, abi, icosphere_scene, make_camera, render_desc, sampler, soup_scene

RESOLUTIONS = [(256, 256), (640, 480), (1280, 720), (1920, 1080)]
SPPS = [1, 2, 4, 8, 16, 32, 64]


@dataclass
class PoolScene:
    scene: Scene
    kind: str
    triangles: int
    server_id: int = -1


def pool_sizes(n: int, lo: int, hi: int, seed: int) -> list[int]:
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n)]


def build_pool(ctx, sizes: list[int], seed: int, threads: int = 8, times: dict | None = None) -> list[PoolScene]:
    """Scenes of the given triangle counts (icosphere clusters round theirs
    to whole spheres), built with the BVH.  The generators run on `threads`
    host threads.  `times`, when given, receives the seconds spent
    generating, creating (packing and upload) and building the scenes."""
    from concurrent.futures import ThreadPoolExecutor

    def generate(i, n):
        if i % 2 == 0:
            return soup_scene(seed + i, n, max(1, n // 100)), "soup", n
        sub = int(min(8, max(2, round(math.log(max(n, 160) / 160.0, 4)))))
        spheres = max(1, int(round(n / (20 * 4 ** sub))))
        total = spheres * 20 * 4 ** sub
        return icosphere_scene(seed + i, spheres, sub, max(1, total // 100)), "icospheres", total

    t = {"generate_s": 0.0, "create_s": 0.0, "build_s": 0.0}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        datas = list(ex.map(lambda a: generate(*a), enumerate(sizes)))
    t["generate_s"] = time.perf_counter() - t0
    out = []
    for data, kind, n in datas:
        t0 = time.perf_counter()
        s = Scene(ctx, data)
        t1 = time.perf_counter()
        s.build(abi.BVH)
        t2 = time.perf_counter()
        t["create_s"] += t1 - t0
        t["build_s"] += t2 - t1
        s.data = None  # the library keeps its own copy
        out.append(PoolScene(s, kind, n))
    if times is not None:
        times.update(t)
    return out


def request_desc(rng: np.random.Generator, seed: int):
    w, h = RESOLUTIONS[rng.integers(len(RESOLUTIONS))]
    spp = SPPS[rng.integers(len(SPPS))]
    while True:
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        if abs(u[1]) < 0.9:
            break
    c = np.array([0.5, 0.5, 0.5])
    pos = c + u * rng.uniform(0.9, 1.5)
    at = c + rng.uniform(-0.1, 0.1, size=3)
    ratio = np.float32(w) / np.float32(h)
    fh = float(np.float32(45.0) * np.float32(max(float(ratio), 1.0)))
    fv = float(np.float32(45.0) * np.float32(max(float(np.float32(1) / ratio), 1.0)))
    cam = make_camera(abi.CAMERA_PERSPECTIVE, tuple(pos.tolist()), tuple(at.tolist()), (0, 1, 0), fh, fv)
    return render_desc(abi.SHADER_PATHTRACER, w, h, spp, cam, pixel=sampler(abi.SAMPLER_HALTON, 0.0, seed),
                       rr=sampler(abi.SAMPLER_RANDOM, 0.0, seed + 1), light=sampler(abi.SAMPLER_RANDOM, 0.0, seed + 2),
                       hemi=sampler(abi.SAMPLER_RANDOM, 0.0, seed + 3), depth_min=2, depth_max=6,
                       kernel=abi.KERNEL_WAVEFRONT)


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else float("nan")


def plain_fifo(pool: list[PoolScene], arrivals: list[float], reqs: list, accum: bool = True) -> dict:
    """The plain GPU path for the same stream: each request rendered alone
    (mrt_render, AUTO kernel, host buffers, wall clock) on an idle GPU, and
    the latencies of serving them one at a time in arrival order with those
    service times (first in, first out, no overlap)."""
    alone = []
    for si, desc in reqs:
        d = abi.RenderDesc.from_buffer_copy(desc)
        d.kernel = abi.KERNEL_AUTO
        bm = np.zeros((d.height, d.width), np.uint32)
        acc = np.zeros((d.height, d.width, 3), np.float32) if accum else None
        t0 = time.perf_counter()
        pool[si].scene.render(d, bitmap=bm, accum=acc)
        alone.append((time.perf_counter() - t0) * 1e3)
    free, lat = 0.0, []
    for ta, ms in zip(arrivals, alone):
        start = max(ta * 1e3, free)
        free = start + ms
        lat.append(free - ta * 1e3)
    return {"alone_p50_ms": round(pct(alone, 50), 2), "alone_p99_ms": round(pct(alone, 99), 2),
            "fifo_p50_ms": round(pct(lat, 50), 2), "fifo_p99_ms": round(pct(lat, 99), 2),
            "busy_s": round(sum(alone) / 1e3, 2)}


def run_stream(srv: RenderServer, pool: list[PoolScene], rate: float, duration: float, seed: int,
               accum: bool = True, max_requests: int = 100000, plain: bool = False) -> dict:
    """Submits Poisson(rate) arrivals for `duration` seconds, waits for all of
    them and returns latency percentiles (ms) and the throughput of the
    window from the first submission to the last result; with `plain`, also
    the same stream on the plain GPU path (plain_fifo)."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = rng.exponential(1.0 / rate)
    while t < duration and len(arrivals) < max_requests:
        arrivals.append(t)
        t += rng.exponential(1.0 / rate)
    reqs = [(int(rng.integers(len(pool))), request_desc(rng, seed * 7919 + 16 * k)) for k in range(len(arrivals))]
    pending: queue.Queue = queue.Queue()
    infos = []
    errors = []

    def waiter():
        while True:
            item = pending.get()
            if item is None:
                return
            jid, meta = item
            try:
                _, _, info = srv.wait(jid)
                infos.append((meta, info.t_submit_us, info.t_admit_us, info.t_gpu_done_us, info.t_done_us))
            except Exception as e:  # a failed job is reported, not hidden
                errors.append(str(e))

    th = threading.Thread(target=waiter)
    th.start()
    st0 = srv.stats()
    t0 = time.perf_counter()
    try:
        for k, ta in enumerate(arrivals):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            si, desc = reqs[k]
            bm = np.zeros((desc.height, desc.width), np.uint32)
            acc = np.zeros((desc.height, desc.width, 3), np.float32) if accum else None
            jid = srv.submit(pool[si].server_id, desc, bitmap=bm, accum=acc)
            pending.put((jid, (si, desc.width, desc.height, desc.samples_pixel)))
    finally:
        pending.put(None)
        th.join()
    st1 = srv.stats()
    rays = (st1.rays_camera + st1.rays_extension + st1.rays_shadow) - (
        st0.rays_camera + st0.rays_extension + st0.rays_shadow)
    lat = [(d - s) / 1000.0 for (_, s, _, _, d) in infos]
    wait = [(a - s) / 1000.0 for (_, s, a, _, _) in infos]
    window = (max(d for (*_, d) in infos) - min(s for (_, s, *_) in infos)) / 1e6 if infos else float("nan")
    px_samples = sum(m[1] * m[2] * m[3] for (m, *_) in infos)
    out = {"rate_rps": rate, "duration_s": duration, "requests": len(arrivals), "done": len(infos),
           "failed": len(errors), "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
           "mean_ms": round(float(np.mean(lat)) if lat else float("nan"), 2),
           "max_ms": round(max(lat) if lat else float("nan"), 2),
           "queue_p50_ms": round(pct(wait, 50), 2), "queue_p99_ms": round(pct(wait, 99), 2),
           "window_s": round(window, 3), "achieved_rps": round(len(infos) / window, 2) if infos else 0.0,
           "mrays_per_s": round(rays / window / 1e6, 2) if infos else 0.0, "rays": int(rays),
           "pixel_samples": int(px_samples), "errors": errors[:3]}
    if plain:
        out["plain"] = plain_fifo(pool, arrivals, reqs, accum)
    return out


# ---------------------------------------------------------------- batched --
BATCH_SPPS = [1, 4, 16, 64]


def batch_requests(pool: list[PoolScene], seed: int) -> list:
    """One request per pool scene: random camera pose, resolution of
    RESOLUTIONS and spp of {1, 4, 16, 64}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(len(pool)):
        d = request_desc(rng, seed * 104729 + 16 * i)
        d.samples_pixel = BATCH_SPPS[rng.integers(len(BATCH_SPPS))]
        out.append((i, d))
    return out


def run_batch(srv: RenderServer, pool: list[PoolScene], reqs: list) -> dict:
    """Submits every request at once, largest first (pixels x samples: the
    longest jobs start while the pool is emptiest, the usual makespan
    heuristic; no pixel depends on the order), and waits for all: makespan
    and Mrays/s."""
    st0 = srv.stats()
    order = sorted(reqs, key=lambda r: -int(r[1].width) * int(r[1].height) * int(r[1].samples_pixel))
    t0 = time.perf_counter()
    ids = []
    for si, desc in order:
        bm = np.zeros((desc.height, desc.width), np.uint32)
        acc = np.zeros((desc.height, desc.width, 3), np.float32)
        ids.append(srv.submit(pool[si].server_id, desc, bitmap=bm, accum=acc))
    lat = []
    for jid in ids:
        _, _, info = srv.wait(jid)
        lat.append((info.t_done_us - info.t_submit_us) / 1000.0)
    el = time.perf_counter() - t0
    st1 = srv.stats()
    rays = (st1.rays_camera + st1.rays_extension + st1.rays_shadow) - (
        st0.rays_camera + st0.rays_extension + st0.rays_shadow)
    return {"jobs": len(reqs), "makespan_s": round(el, 3), "rays": int(rays), "mrays_per_s": round(rays / el / 1e6, 2),
            "p50_ms": round(pct(lat, 50), 2), "max_ms": round(max(lat), 2)}


def tune_server(ctx, pool: list[PoolScene], reqs: list, candidates=None) -> tuple[dict, list]:
    """Autotuner of the server shape (pipelines, slot-pool size) on a probe
    batch: the requests at min(spp, 4) samples (enough for jobs to run
    sample-parallel where the pool allows).  These choices change speed only,
    never a pixel.  Returns the best configuration and every probe's
    result."""
    if candidates is None:
        candidates = [dict(pipelines=p, max_slots=s) for p in (1, 2, 3)
                      for s in (2_097_152, 6_291_456, 12_582_912, 25_165_824, 33_554_432)]
    probe = []
    for si, d in reqs:
        q = abi.RenderDesc.from_buffer_copy(d)
        q.samples_pixel = min(int(d.samples_pixel), 4)
        probe.append((si, q))
    results = []
    for cand in candidates:
        srv = RenderServer(ctx, job_memory=4 << 30, **cand)
        try:
            for p in pool:
                p.server_id = srv.add_scene(p.scene)
            r = run_batch(srv, pool, probe)
        finally:
            srv.close()
        results.append(dict(cand, probe_mrays_per_s=r["mrays_per_s"], probe_makespan_s=r["makespan_s"]))
    best = max(results, key=lambda r: r["probe_mrays_per_s"])
    return {k: best[k] for k in candidates[0]}, results
