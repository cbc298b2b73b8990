"""Benchmark of the ray-scene hot path on one MI355X.

Workload (the headline configuration): a random soup of 1,000,000 FP32
triangles in the unit cube (990,000 Lambertian with vertex-colour albedo
U(0.2, 0.8), 10,000 emissive area lights with radiance 5), 1920x1080,
PathTracer at 16 samples per pixel, maximum depth 6 with Russian roulette
from depth 3, per-pixel Halton pixel sampler, BVH accelerator, FP32.

One step = one Renderer::renderFrame of the full 16-spp frame.  The scene,
the BVH and the output buffers are resident in HBM before the timed region.
Mrays/s counts every traced camera, extension, reflection/transmission and
shadow ray (rays the reference traces beyond the maximum depth, whose result
it discards, are not traced and not counted).

Serving leg (after the timed frames, outside their timing): a render server
holding a resident pool of 16 seeded scenes (50k-8M triangles, soups and
displaced-icosphere clusters, 1% emissive) receives a Poisson stream of
requests (random camera pose, 256x256 to 1920x1080, 1-64 spp) for a fixed
window; p50/p99 request latency and the window's Mrays/s go into the line's
"serving" object.  --serve-plain adds ("plain") the same stream on the plain
GPU path: each request rendered alone with mrt_render, and the latencies of
serving them one at a time in arrival order with those times (off by
default: those renders run the trace kernel of the timed frames at other
sizes, and the profile of the default command then mixes them into that
kernel's average).  --serve-sweep runs only the serving workload over a
list of rates and prints one JSON line per rate.

Denoiser leg (after the timed frames, outside their timing): the BF16 U-Net
(mrt_denoiser_run, trained weights) over a 1920x1080 frame of FP32 radiance
and AOVs resident in HBM; device time per frame and the MFMA roofline of the
network (algorithmic FLOPs / time) go into the line's "denoiser" object.
--denoise runs only this leg, plus the whole denoised frame of the benchmark
scene (4-sample render with its AOVs, then the network), and prints its own
line.  The denoiser object also carries "training": Adam steps of the
network on the GPU (8 crops of 256x256 per step, synthetic pairs in HBM),
ms per step and TFLOP/s; --denoise-train runs only that leg.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

W, H, SPP = 1920, 1080, 16
N_TRIS, N_LIGHTS = 1_000_000, 10_000
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


KERNELS = {"auto": 0, "megakernel": 1, "wavefront": 2}


def make_desc(mrt, abi, cam, spp=SPP, instrument=False, kernel="auto", ray_sort=0, persistent=0):
    return mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, kernel=KERNELS[kernel], ray_sort=ray_sort, persistent=persistent,
                           pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, 5),
                           rr=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 6),
                           light=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 7),
                           hemi=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, 8),
                           depth_min=2, depth_max=6, instrument=instrument)


COPY_BYTES = 1 << 30


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel`: the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the counter
    reads for a plain copy of known size, see --copy-calibration); None when
    no summary exists for it."""
    p = ROOT / "profiles" / f"{kernel}_soup1m_pmc.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            return int(float(line.split("=")[1]))
    return None


def copy_calibration() -> int:
    """A plain device copy of COPY_BYTES (reads and writes COPY_BYTES each):
    the reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {COPY_BYTES} bytes")
    return 0


def copy_bandwidth(torch) -> float:
    """Achievable device copy bandwidth (GB/s, read + write bytes): the median
    of 5 timed copies of COPY_BYTES after one warmup."""
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b) / 1000.0)
    del src, dst
    return 2.0 * COPY_BYTES / float(np.median(ts)) / 1e9


def soup_camera(mrt, abi):
    ratio = np.float32(W) / np.float32(H)
    return mrt.make_camera(abi.CAMERA_PERSPECTIVE, (0.5, 0.5, -1.0), (0.5, 0.5, 0.5), (0, 1, 0),
                           float(np.float32(45.0) * ratio), 45.0)


# Tile ids (row * 16 + column) of the CPU baseline's sample: a 4 x 4 lattice
# over the frame (the cube's image and the background beside it).  The
# reference's tile loop hands whole tiles to its threads, so the sample needs
# at least as many tiles as threads.
BASELINE_TILES = tuple(r * 16 + c for r in (2, 6, 9, 13) for c in (3, 6, 9, 12))
BASELINE_SPP = 2
ONE_THREAD_TILE = 7 * 16 + 7


def host_threads() -> int:
    """Host threads of the CPU baseline: the logical CPUs this process may
    run on, at most 16 (one GPU's share of an MI355X host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(data, cam, abi, mrt, accel: str = "bvh", runs: int = 3) -> dict:
    """The oracle (reference-faithful CPU restatement: AoS triangles,
    unordered un-culled BVH walk, linear light scan, std::thread tile loop)
    on a bounded sample of the same workload and accelerator: the first 2 of
    the 16 samples of 16 of the 256 tiles (a lattice over the frame), median
    of `runs` runs on every host thread available (at most 16); and one
    thread on one tile at 2 samples.  All 16 samples of 16 tiles would take
    about 45 s per run on 16 threads; the sample passes differ only in their
    random numbers, so the rate per ray is that of the whole frame."""
    import oracle

    threads = host_threads()
    sc = oracle.OracleScene(data.desc())
    p = abi.AccelParams()
    p.leaf_size = 2
    if accel == "grid":
        cmin, cmax = mrt.camera_aabb(cam)
        p.camera_min[:] = list(cmin)
        p.camera_max[:] = list(cmax)
    sc.build(abi.REGULAR_GRID if accel == "grid" else abi.BVH, p)

    def mask(desc, tiles):
        ids = np.rint(np.frombuffer(bytes(desc.tile_values), np.float32) * 256).astype(int)
        return np.array([1 if t in tiles else 0 for t in ids], np.uint8)

    desc = make_desc(mrt, abi, cam, spp=BASELINE_SPP)
    rates, rays = [], 0
    for _ in range(runs):
        _, counters, secs = sc.render(desc, threads=threads, tile_mask=mask(desc, BASELINE_TILES))
        rays = int(counters.sum())
        rates.append(rays / secs / 1e6)
        log(f"cpu baseline run: {rays} rays in {secs:.2f} s on {threads} threads")
    d1 = make_desc(mrt, abi, cam, spp=2)
    _, c1, s1 = sc.render(d1, threads=1, tile_mask=mask(d1, (ONE_THREAD_TILE,)))
    one = int(c1.sum()) / s1 / 1e6
    name = "RegularGrid" if accel == "grid" else "BVH"
    return {"value": float(np.median(rates)), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle render ({name}), 16 tiles (a 4x4 lattice of the 16x16 grid) x samples 0-1 of the "
                      f"1920x1080 {SPP}-spp frame ({rays} rays), median of {runs} runs",
            "runs_mrays_per_s": [round(r, 4) for r in rates], "cpu_model": cpu_model(),
            "one_thread": {"value": round(one, 4), "unit": "Mrays/s",
                           "sample": f"tile {ONE_THREAD_TILE} x 2 samples, {int(c1.sum())} rays in {s1:.2f} s"},
            "build": "g++ -O3 -march=x86-64-v3 (built on another host than the one it runs on)"}


def ordered_traffic(data, cam, abi, mrt) -> dict:
    """Algorithmic bytes per ray of the benchmark frame, independent of this
    implementation: the oracle's ordered (near-first, culled) traversal of
    the same reference BVH over the rays of every 256th pixel (all 16
    samples): 32 bytes of ray in, 16 of hit out (4 for a shadow ray), 64 per
    interior visit (the node's child pair) and 36 per triangle test (A, AB,
    AC)."""
    import oracle

    sc = oracle.OracleScene(data.desc())
    p = abi.AccelParams()
    p.leaf_size = 2
    sc.build(abi.BVH, p)
    t0 = time.time()
    r = sc.ordered_traffic(make_desc(mrt, abi, cam, spp=SPP), 256, host_threads())
    rc, rs = r["rays_closest"], r["rays_shadow"]
    pairs = r["pairs_closest"] + r["pairs_shadow"] + r["pairs_lights"]
    tests = r["tests_closest"] + r["tests_shadow"] + r["tests_lights"]
    per_ray = (48 * rc + 36 * rs + 64 * pairs + 36 * tests) / (rc + rs)
    log(f"ordered traversal over 1/256 of the pixels ({time.time() - t0:.1f} s): {rc:.0f} closest-hit rays "
        f"({r['pairs_closest'] / rc:.1f} pairs, {r['tests_closest'] / rc:.1f} tests; lights "
        f"{r['pairs_lights'] / rc:.1f} pairs, {r['tests_lights'] / rc:.1f} tests), {rs:.0f} shadow rays "
        f"({r['pairs_shadow'] / rs:.1f} pairs, {r['tests_shadow'] / rs:.1f} tests): {per_ray:.1f} bytes per ray")
    return {"bytes_per_ray": round(per_ray, 1), "rays": int(rc + rs),
            "pairs_per_ray": round(pairs / (rc + rs), 2), "tests_per_ray": round(tests / (rc + rs), 2),
            "sample": "rays of every 256th pixel, all 16 samples"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-bytes", action="store_true",
                    help="roofline bytes from the GPU's own visit counts instead of the oracle's ordered traversal")
    ap.add_argument("--kernel", choices=sorted(KERNELS), default="auto")
    ap.add_argument("--sort", choices=["default", "on", "off"], default="default",
                    help="wavefront: reorder extension rays by origin cell (default: off)")
    ap.add_argument("--no-persistent", action="store_true", help="no persistent kernels: the wavefront with one ray per lane, no refill")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for the HBM byte counters)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-rate", type=float, default=5.0, help="serving leg: requests per second")
    ap.add_argument("--serve-duration", type=float, default=20.0, help="serving leg: arrival window, seconds")
    ap.add_argument("--serve-plain", action="store_true",
                    help="serving leg: also the same stream on the plain path (each request alone, FIFO)")
    ap.add_argument("--serve-scenes", type=int, default=16)
    ap.add_argument("--serve-min-tris", type=int, default=50_000)
    ap.add_argument("--serve-max-tris", type=int, default=8_000_000)
    ap.add_argument("--serve-sweep", type=str, default="",
                    help="comma-separated rates: run only the serving workload, one JSON line per rate")
    ap.add_argument("--accel", choices=["bvh", "grid"], default="bvh",
                    help="accelerator of the scene: the reference BVH (default) or the 128^3 RegularGrid")
    ap.add_argument("--builder", choices=["auto", "host", "gpu"], default="auto",
                    help="BVH builder of the scene (auto: GPU from 65536 triangles)")
    ap.add_argument("--denoise", action="store_true", help="run only the denoiser leg (1920x1080, trained weights)")
    ap.add_argument("--denoise-reps", type=int, default=20)
    ap.add_argument("--denoise-train", action="store_true",
                    help="run only the denoiser training leg (Adam steps, 8 crops of 256x256)")
    ap.add_argument("--denoise-train-steps", type=int, default=40)
    ap.add_argument("--batched", action="store_true",
                    help="run only the batched workload: 64 scenes (10k-4M triangles), one request each, all at once")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a frame would shard by tiles or by sample index; only one GPU's share is built)", file=sys.stderr)
        return 2

    import torch

    import mobileraytracer_amd as mrt
    from mobileraytracer_amd import abi

    if args.batched:
        return batched(mrt, abi)

    if args.denoise_train:
        ctx = mrt.Context(0)
        res = denoiser_train_leg(mrt, ctx, args.denoise_train_steps)
        print(json.dumps({"metric": "denoiser training ms per step (BF16 U-Net on MFMA, 8 crops of 256x256)",
                          "value": res["ms_per_step"], "unit": "ms", "higher_is_better": False, "training": res}),
              flush=True)
        return 0

    if args.denoise:
        ctx = mrt.Context(0)
        data = mrt.soup_scene(2, N_TRIS, N_LIGHTS, 5.0, 0.005, abi.SAMPLER_RANDOM)
        cam = soup_camera(mrt, abi)
        scene = mrt.Scene(ctx, data)
        scene.build(abi.BVH)
        res = denoiser_leg(mrt, abi, ctx, torch, args.denoise_reps, scene=scene,
                           desc4=make_desc(mrt, abi, cam, spp=4, kernel=args.kernel))
        print(json.dumps({"metric": "denoiser ms per 1920x1080 frame (BF16 U-Net on MFMA)",
                          "value": res["ms_per_frame"], "unit": "ms", "higher_is_better": False, "denoiser": res}),
              flush=True)
        return 0

    if args.serve_sweep:
        ctx = mrt.Context(0)
        srv, pool = serving_setup(mrt, ctx, args)
        for rate in [float(r) for r in args.serve_sweep.split(",")]:
            # Bounded request count at high rates (the host buffers of queued
            # requests are held until they are served).
            dur = min(args.serve_duration, 600.0 / rate)
            res = serving_run(mrt, srv, pool, rate, dur)
            print(json.dumps({"metric": "render-request latency (serving sweep)", "serving": res}), flush=True)
        srv.close()
        return 0

    t0 = time.time()
    data = mrt.soup_scene(2, N_TRIS, N_LIGHTS, 5.0, 0.005, abi.SAMPLER_RANDOM)
    cam = soup_camera(mrt, abi)
    log(f"scene generated in {time.time() - t0:.1f} s")
    ctx = mrt.Context(0)
    scene = mrt.Scene(ctx, data)
    builder = {"auto": abi.BUILDER_AUTO, "host": abi.BUILDER_HOST, "gpu": abi.BUILDER_GPU}[args.builder]
    t0 = time.perf_counter()
    if args.accel == "grid":
        cmin, cmax = mrt.camera_aabb(cam)
        scene.build(abi.REGULAR_GRID, cmin, cmax, builder=builder)
    else:
        scene.build(abi.BVH, builder=builder)
    build_s = time.perf_counter() - t0
    log(f"{'RegularGrid' if args.accel == 'grid' else 'BVH'} built and uploaded in {build_s:.3f} s "
        f"({args.builder} builder), footprint {scene.footprint() / 1e6:.1f} MB")

    dev = torch.device("cuda:0")
    bitmap = torch.zeros((H, W), dtype=torch.int32, device=dev)
    accum = torch.zeros((H, W, 3), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    # Instrumented frame (the timed frame's configuration with the traversal
    # counters on, outside the timed region): node visits and primitive tests
    # per ray, lane occupancy of the persistent trace.
    sort = {"default": 0, "on": 1, "off": 2}[args.sort]
    pers = 2 if args.no_persistent else 0
    _, st = scene.render(make_desc(mrt, abi, cam, instrument=True, kernel=args.kernel, ray_sort=sort, persistent=pers),
                         bitmap=bitmap,
                         accum=accum,
                         mem=abi.MEM_DEVICE)
    st_instr = st
    rays1 = st.rays_camera + st.rays_extension + st.rays_shadow + st.rays_specular
    closest1 = st.rays_camera + st.rays_extension + st.rays_specular
    pairs_per_ray = st.node_visits / rays1
    prims_per_ray = st.prim_tests / rays1
    shadow1 = max(1, st.rays_shadow)
    log(f"instrumented pass: {rays1} rays, {pairs_per_ray:.1f} node pairs and {prims_per_ray:.1f} "
        f"primitive tests per ray; per closest-hit ray: "
        f"{(st.node_visits - st.light_node_visits - st.shadow_node_visits) / closest1:.1f} pairs "
        f"{(st.prim_tests - st.light_prim_tests - st.shadow_prim_tests) / closest1:.1f} tests, lights "
        f"{st.light_node_visits / closest1:.1f} pairs {st.light_prim_tests / closest1:.1f} tests; per shadow ray: "
        f"{st.shadow_node_visits / shadow1:.1f} pairs {st.shadow_prim_tests / shadow1:.1f} tests")
    occupancy = None
    if st.trace_lane_slots:
        occupancy = st.trace_lane_iters / st.trace_lane_slots
        log(f"persistent trace: {st.trace_lane_iters / rays1:.1f} lane iterations per ray, lanes holding a ray "
            f"{100.0 * occupancy:.1f}% of lane iterations, visits+tests per lane "
            f"iteration {(st.node_visits + st.prim_tests) / max(1, st.trace_lane_iters):.2f}")

    desc = make_desc(mrt, abi, cam, kernel=args.kernel, ray_sort=sort, persistent=pers)
    for _ in range(args.warmup):
        scene.render(desc, bitmap=bitmap, accum=accum, mem=abi.MEM_DEVICE)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    rays = 0
    closest = 0
    kernel_ms = 0.0
    trace_ms = 0.0
    busy_ms = 0.0
    launches = 0
    groups = 0
    mega_p = False
    for _ in range(args.steps):
        _, st = scene.render(desc, bitmap=bitmap, accum=accum, mem=abi.MEM_DEVICE)
        rays += st.rays_camera + st.rays_extension + st.rays_shadow + st.rays_specular
        closest += st.rays_camera + st.rays_extension + st.rays_specular
        kernel_ms += st.ms_kernel
        trace_ms += st.ms_trace
        busy_ms += st.ms_trace_busy
        launches += st.trace_launches
        groups = st.trace_groups
        mega_p = st.persistent_megakernel != 0
        if mega_p:
            log(f"step: {st.ms_render:.1f} ms, {st.rays_camera} camera / {st.rays_extension} extension / "
                f"{st.rays_shadow} shadow rays; persistent megakernel {st.ms_kernel:.1f} ms "
                f"({st.slot_lanes} sample lanes per pixel)")
        else:
            log(f"step: {st.ms_render:.1f} ms, {st.rays_camera} camera / {st.rays_extension} extension / "
                f"{st.rays_shadow} shadow rays; trace {st.ms_trace:.1f} ms in {st.trace_launches} launches "
                f"({st.trace_groups} slot groups, some trace launch running {st.ms_trace_busy:.1f} ms)")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    mrays = rays / elapsed / 1e6

    # Roofline of the dominant kernel.  Algorithmic bytes per ray (BVH): the
    # oracle's ordered traversal of the same reference tree over a 1/256 pixel
    # subsample of the frame (ordered_traffic), independent of this
    # implementation; beside it the GPU's own visits over the compressed walk
    # tree: 64 bytes per node visit + 64 per triangle test (a 128-byte leaf
    # record holds the exact leaf box and two triangles) + 48 of ray in and
    # result out.  These bytes count every node and triangle a ray touches;
    # the frame's scene (about 250 MB) stays in the 256 MB Infinity Cache and
    # the L2s, so they can exceed what HBM could deliver (frac > 1), and
    # traffic_frac gives the bytes the counters saw leave L2.
    # Wavefront (the AUTO kernel of the 16-spp 1080p frame): the persistent
    # trace kernel, timed with HIP events around each of its launches on the
    # stream of its slot group.  Slot groups run their trace launches
    # concurrently on separate streams, so a launch's duration includes time
    # shared with the other groups' launches; device_achieved divides all
    # trace bytes by the time at least one trace launch ran, and
    # frame_achieved all bytes (trace and 256 bytes of shading per closest-hit
    # ray) by the frame time.
    # Persistent megakernel: one launch per frame, timed with HIP events
    # around it, its bytes the rays' traversal bytes plus the shading bytes.
    # Per-pass megakernel: one launch per sample pass, likewise.
    # RegularGrid scenes: 8 bytes (a cell's two CSR offsets) per DDA step and
    # 52 per cell primitive test (its 4-byte reference and the triangle);
    # the light tree's pairs and tests as above.
    wavefront = launches > 0
    persistent = (wavefront or mega_p) and not args.no_persistent
    bytes_per_ray = 48 + 64 * pairs_per_ray + (64 if persistent and args.accel == "bvh" else 48) * prims_per_ray
    gpu_bytes_per_ray = bytes_per_ray
    ordered = None
    if args.accel == "bvh" and not args.no_oracle_bytes:
        ordered = ordered_traffic(data, cam, abi, mrt)
        bytes_per_ray = ordered["bytes_per_ray"]
    if args.accel == "grid":
        light_pairs = st_instr.light_node_visits / rays1
        light_tests = st_instr.light_prim_tests / rays1
        bytes_per_ray = (48 + 8 * (pairs_per_ray - light_pairs) + 64 * light_pairs +
                         52 * (prims_per_ray - light_tests) + 48 * light_tests)
    if mega_p:
        kernel_name = "k_mega_p"
        launches = args.steps
        bytes_total = rays * bytes_per_ray + closest * 256
        avg_launch_s = kernel_ms / 1000.0 / launches
    elif wavefront:
        kernel_name = ("k_wf_extend + k_wf_shadow" if not persistent else
                       "k_wf_trace_grid_p" if args.accel == "grid" else "k_wf_trace_p")
        bytes_total = rays * bytes_per_ray
        avg_launch_s = trace_ms / 1000.0 / launches
    else:
        kernel_name = "k_render_pass"
        launches = args.steps * SPP
        bytes_total = rays * bytes_per_ray + closest * 256
        avg_launch_s = kernel_ms / 1000.0 / launches
    bytes_per_launch = bytes_total / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    device_achieved = bytes_total / (busy_ms / 1000.0) / 1e9 if wavefront and busy_ms > 0 else achieved
    frame_achieved = (rays * bytes_per_ray + closest * 256) / elapsed / 1e9
    traffic = measured_traffic(kernel_name) if persistent and args.accel == "bvh" else None
    copy_gbs = copy_bandwidth(torch)

    out = {
        "metric": "Mrays/s (1M-tri path trace, 16 spp) vs CPU; p50/p99 render-request latency",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 1M-triangle soup (10k emissive), deterministic per-pixel sampling",
        "config": {"workload": f"soup-1M 1920x1080 PathTracer 16spp depth6 RR3 Halton "
                               f"{'RegularGrid128' if args.accel == 'grid' else 'BVH'} {args.kernel}",
                   "triangles": N_TRIS, "lights": N_LIGHTS, "width": W, "height": H, "spp": SPP,
                   "rays_per_step": rays // args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     # Counted bytes leaving L2 (the committed rocprofv3
                     # FETCH_SIZE + WRITE_SIZE summary) over this run's launch
                     # time: what the HBM / Infinity Cache side actually moved.
                     "traffic_achieved": round(traffic / avg_launch_s / 1e9, 1) if traffic else None,
                     "traffic_frac": round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "copy_achievable": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
                     "kernel": kernel_name, "launches": launches,
                     "avg_launch_ms": round(avg_launch_s * 1000.0, 4),
                     "concurrent_groups": groups, "device_achieved": round(device_achieved, 1),
                     "device_frac": round(device_achieved / HBM_PEAK_GBS, 4),
                     "frame_achieved": round(frame_achieved, 1),
                     "frame_frac": round(frame_achieved / HBM_PEAK_GBS, 4),
                     "bytes_per_launch": round(bytes_per_launch), "bytes_per_ray": round(bytes_per_ray, 1),
                     "bytes_per_ray_source": "oracle ordered traversal" if ordered else "GPU visit counts",
                     "oracle_ordered": ordered, "gpu_bytes_per_ray": round(gpu_bytes_per_ray, 1),
                     "node_visits_per_ray": round(pairs_per_ray, 2), "tests_per_ray": round(prims_per_ray, 2),
                     "lane_occupancy": round(occupancy, 4) if occupancy is not None else None},
    }
    if not args.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(data, cam, abi, mrt, args.accel)
        log(f"cpu baseline took {time.time() - t0:.1f} s")
        out["vs_cpu"] = round(mrays / out["cpu_baseline"]["value"], 1)
    out["scene_build"] = build_leg(mrt, abi, ctx, data, cam, args.builder, build_s)
    # The network alone (synthetic inputs: no trace launch besides the timed
    # frames'); the denoised-frame pipeline is measured by --denoise.
    out["denoiser"] = denoiser_leg(mrt, abi, ctx, torch, args.denoise_reps)
    out["denoiser"]["training"] = denoiser_train_leg(mrt, ctx, args.denoise_train_steps)
    if not args.no_serve:
        scene.close()
        del bitmap, accum
        srv, pool = serving_setup(mrt, ctx, args)
        out["serving"] = serving_run(mrt, srv, pool, args.serve_rate, args.serve_duration, plain=args.serve_plain)
        srv.close()
    print(json.dumps(out), flush=True)
    return 0


def build_leg(mrt, abi, ctx, data, cam, builder: str, build_s: float) -> dict:
    """Reference BVH of the benchmark scene's 990k triangles built on the
    host (mrt_bvh_build) and on the GPU (mrt_bvh_build_gpu, boxes and nodes
    in host memory, so the copies are included): wall clock and whether the
    node arrays are identical; the whole timed scene build above; and whole
    RegularGrid scene builds with each builder."""
    import ctypes as C

    n = len(data.triangles)
    boxes = np.zeros((n, 6), np.float32)
    abi.check(abi.lib().mrt_primitive_aabbs(abi.CAT_TRIANGLE, data.triangles.ctypes.data, n, boxes.ctypes.data),
              "mrt_primitive_aabbs")
    cnt = C.c_uint32()
    host = np.zeros((2 * n - 1, 8), np.uint32)
    gpu = np.zeros_like(host)
    t0 = time.perf_counter()
    abi.check(abi.lib().mrt_bvh_build(boxes.ctypes.data, n, 2, host.ctypes.data, len(host), C.byref(cnt)),
              "mrt_bvh_build")
    host_ms = (time.perf_counter() - t0) * 1e3
    nh = cnt.value
    gpu_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        abi.check(abi.lib().mrt_bvh_build_gpu(ctx.h, boxes.ctypes.data, n, 2, gpu.ctypes.data, len(gpu), C.byref(cnt),
                                              abi.MEM_HOST), "mrt_bvh_build_gpu")
        gpu_ms.append((time.perf_counter() - t0) * 1e3)
    same = cnt.value == nh and np.array_equal(host[:nh], gpu[:nh])
    log(f"reference BVH of {n} triangles: host {host_ms:.1f} ms, GPU {min(gpu_ms):.1f} ms, identical {same}")
    # RegularGrid of the same scene (128^3 cells), whole mrt_scene_build
    # with each builder.
    cmin, cmax = mrt.camera_aabb(cam)
    grid_ms = {}
    for name, b in (("host", abi.BUILDER_HOST), ("gpu", abi.BUILDER_GPU)):
        sc = mrt.Scene(ctx, data)
        t0 = time.perf_counter()
        sc.build(abi.REGULAR_GRID, cmin, cmax, builder=b)
        grid_ms[name] = (time.perf_counter() - t0) * 1e3
        sc.close()
    log(f"RegularGrid scene build: host {grid_ms['host']:.0f} ms, GPU {grid_ms['gpu']:.0f} ms")
    return {"builder": builder, "scene_build_ms": round(build_s * 1e3, 1), "bvh_nodes": nh,
            "bvh_host_ms": round(host_ms, 1), "bvh_gpu_ms": round(min(gpu_ms), 1), "bvh_identical": bool(same),
            "grid_scene_build_host_ms": round(grid_ms["host"], 1), "grid_scene_build_gpu_ms": round(grid_ms["gpu"], 1)}


def denoiser_leg(mrt, abi, ctx, torch, reps: int, scene=None, desc4=None) -> dict:
    """The trained U-Net over a synthetic 1920x1080 frame whose FP32
    radiance and AOVs are resident in HBM; device time from the library's
    events around the network's kernels (median of `reps` runs).  With the
    benchmark scene: the whole denoised frame (4-sample render with its AOVs
    in the same call, then the network; device buffers), against the
    4-sample render alone."""
    from mobileraytracer_amd import denoise

    flat, levels, channels, ds = denoise.load_weights()
    dn = mrt.Denoiser(ctx, flat, mrt.denoiser_desc(levels, channels, ds))
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    rad = (torch.rand((H, W, 3), generator=g) * 2.0).to(dev)
    aov = torch.rand((H, W, 7), generator=g).to(dev)
    out = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    for _ in range(3):
        dn.run(rad, aov, 1.0, out=out, mem=abi.MEM_DEVICE, width=W, height=H)
    times = []
    for _ in range(reps):
        dn.run(rad, aov, 1.0, out=out, mem=abi.MEM_DEVICE, width=W, height=H)
        times.append(dn.last_ms())
    pipeline = None
    if scene is not None:
        bm = torch.zeros((H, W), dtype=torch.int32, device=dev)
        acc = torch.zeros((H, W, 3), dtype=torch.float32, device=dev)
        aovs = torch.zeros((H, W, 7), dtype=torch.float32, device=dev)
        t_full, t_render = [], []
        for k in range(4):
            bm.zero_()
            acc.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            scene.render_with_aovs(desc4, bitmap=bm, accum=acc, aov=aovs, mem=abi.MEM_DEVICE)
            dn.run(acc, aovs, 0.25, out=out, mem=abi.MEM_DEVICE, width=W, height=H)
            torch.cuda.synchronize()
            t_full.append((time.perf_counter() - t0) * 1e3)
            bm.zero_()
            acc.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            scene.render(desc4, bitmap=bm, accum=acc, mem=abi.MEM_DEVICE)
            torch.cuda.synchronize()
            t_render.append((time.perf_counter() - t0) * 1e3)
        pipeline = {"ms_denoised_frame": round(float(np.median(t_full[1:])), 2),
                    "ms_render_4spp": round(float(np.median(t_render[1:])), 2),
                    "workload": "benchmark soup 1920x1080, PathTracer 4 spp with AOVs in the same render, then the "
                                "U-Net; device buffers, wall clock, median of 3 after one warm-up"}
        log(f"denoised 1080p frame: {pipeline['ms_denoised_frame']:.1f} ms (4-spp render alone "
            f"{pipeline['ms_render_4spp']:.1f} ms)")
    dn.close()
    ms = float(np.median(times))
    flops = denoise.algorithmic_flops(H, W, levels, channels)
    tflops = flops / (ms / 1000.0) / 1e12
    # Each layer at its own bound (the full-resolution layers move more bytes
    # than MFMA time covers): the network's time at the layer-wise roofline.
    roof_ms = denoise.layer_roofline_s(H, W, HBM_PEAK_GBS * 1e9, BF16_PEAK_TFLOPS * 1e12, levels, channels) * 1e3
    log(f"denoiser: {ms:.3f} ms per {W}x{H} frame (median of {reps}), {tflops:.1f} TFLOP/s algorithmic; "
        f"layer-wise roofline {roof_ms:.3f} ms ({roof_ms / ms:.3f} of it)")
    return {"ms_per_frame": round(ms, 4), "ms_min": round(min(times), 4), "reps": reps,
            "workload": f"U-Net levels={levels} channels={list(channels)} {W}x{H}, BF16 weights/activations, FP32 "
                        "accumulation, inputs resident in HBM",
            "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / BF16_PEAK_TFLOPS, 4), "traffic": None,
                         "flops_per_frame": flops, "kernel": "whole network (k_conv3x3 + pack/pool/up/unpack)",
                         "layer_roofline_ms": round(roof_ms, 4), "frac_of_layer_roofline": round(roof_ms / ms, 4),
                         "layer_roofline": "sum over the network's kernels of max(least HBM bytes / 8 TB/s, "
                                           "FLOPs / 2.5 PFLOP/s)"},
            "pipeline": pipeline}


def denoiser_train_leg(mrt, ctx, steps: int, batch: int = 8, crop: int = 256) -> dict:
    """Training steps of the U-Net on the GPU (mrt_denoiser_trainer_step:
    gather of `batch` crops, forward, backward-data and weight-gradient
    products on BF16 MFMA, Adam) over synthetic pairs resident in HBM (16
    frames of 512x512); device time from the library's events, algorithmic
    FLOPs per step from the network's shapes (forward, weight gradient and,
    but for the first layer, backward-data: 2 * 9 * cin * cout per pixel
    each)."""
    from mobileraytracer_amd import denoise

    rng = np.random.default_rng(0)
    n, side = 16, 512
    x = np.zeros((n, side, side, 10), np.float32)
    x[..., 0:3] = np.log1p(rng.exponential(0.5, size=(n, side, side, 3)))
    x[..., 3:9] = rng.uniform(-1, 1, size=(n, side, side, 6))
    x[..., 9] = rng.uniform(0, 3, size=(n, side, side))
    t = np.log1p(rng.exponential(0.5, size=(n, side, side, 3))).astype(np.float32)
    flat, levels, channels, ds = denoise.load_weights()
    tr = mrt.DenoiserTrainer(ctx, flat, mrt.denoiser_desc(levels, channels, ds),
                             mrt.denoiser_train_desc(batch=batch, crop=crop, seed=1))
    tr.set_pairs(x, t)
    del x, t
    tr.step(3, 1e-4)
    per = []
    chunk = max(1, steps // 4)
    for _ in range(4):
        tr.step(chunk, 1e-4)
        per.append(tr.stats().ms_per_step)
    st = tr.stats()
    tr.close()
    ms = float(np.median(per))
    tflops = st.flops_per_step / (ms / 1000.0) / 1e12
    log(f"denoiser training: {ms:.3f} ms per step ({batch} crops of {crop}^2), {tflops:.1f} TFLOP/s algorithmic")
    return {"ms_per_step": round(ms, 4), "steps": 4 * chunk, "batch": batch, "crop": crop,
            "workload": f"U-Net levels={levels} channels={list(channels)}: Adam steps on {batch} crops of "
                        f"{crop}x{crop} from 16 synthetic 512x512 pairs in HBM; BF16 MFMA forward, backward-data "
                        "and weight-gradient kernels, FP32 master weights",
            "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / BF16_PEAK_TFLOPS, 4), "traffic": None,
                         "flops_per_step": st.flops_per_step,
                         "kernel": "whole step (k_conv3x3, k_octet, k_wgrad, k_sum_splits, pool/upsample "
                                   "gradients, k_mse, k_gather, k_adam)"}}


def batched(mrt, abi) -> int:
    """64 independent scenes (log-uniform 10k-4M triangles, alternately soups
    and displaced-icosphere clusters) rendered as one batch through the
    render server, whose shape the autotuner picks on a probe batch (the
    requests at up to 4 spp);
    compared with the same requests rendered one after another by mrt_render."""
    from mobileraytracer_amd import serving

    ctx = mrt.Context(0)
    t0 = time.time()
    sizes = serving.pool_sizes(64, 10_000, 4_000_000, seed=3)
    pool = serving.build_pool(ctx, sizes, seed=300)
    log(f"batched: {len(pool)} scenes, {sum(p.triangles for p in pool)} triangles, built in {time.time() - t0:.1f} s")
    reqs = serving.batch_requests(pool, seed=3)
    t0 = time.time()
    best, probes = serving.tune_server(ctx, pool, reqs)
    log(f"autotuner ({time.time() - t0:.1f} s): {probes}; picked {best}")
    srv = mrt.RenderServer(ctx, job_memory=4 << 30, **best)
    for p in pool:
        p.server_id = srv.add_scene(p.scene)
    res = serving.run_batch(srv, pool, reqs)
    srv.close()
    log(f"batch: {res}")
    # The same requests one at a time through mrt_render.
    t0 = time.perf_counter()
    rays = 0
    for si, desc in reqs:
        bm = np.zeros((desc.height, desc.width), np.uint32)
        acc = np.zeros((desc.height, desc.width, 3), np.float32)
        _, st = pool[si].scene.render(desc, bitmap=bm, accum=acc)
        rays += st.rays_camera + st.rays_extension + st.rays_shadow
    seq = time.perf_counter() - t0
    log(f"one at a time: {seq:.2f} s, {rays / seq / 1e6:.1f} Mrays/s")
    out = {"metric": "Mrays/s (batched: 64 scenes in one pipeline)", "value": res["mrays_per_s"], "unit": "Mrays/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "config": {"workload": "64 scenes 10k-4M triangles (soups, icosphere clusters), one request each: "
                                  "256x256..1920x1080, spp {1,4,16,64}, PathTracer depth 6 RR 3",
                      "triangles": sum(p.triangles for p in pool), "server": best},
           "batch": res, "autotuner": probes,
           "one_at_a_time": {"seconds": round(seq, 3), "rays": int(rays), "mrays_per_s": round(rays / seq / 1e6, 2)}}
    print(json.dumps(out), flush=True)
    return 0


def serving_setup(mrt, ctx, args):
    """Render server with the resident pool (see the module docstring)."""
    from mobileraytracer_amd import serving

    t0 = time.time()
    sizes = serving.pool_sizes(args.serve_scenes, args.serve_min_tris, args.serve_max_tris, seed=4)
    times = {}
    pool = serving.build_pool(ctx, sizes, seed=400, times=times)
    srv = mrt.RenderServer(ctx)
    t1 = time.time()
    for p in pool:
        p.server_id = srv.add_scene(p.scene)
    times["add_s"] = time.time() - t1
    st = srv.stats()
    log(f"serving pool: {len(pool)} scenes, {sum(p.triangles for p in pool)} triangles "
        f"({', '.join(f'{p.kind} {p.triangles}' for p in pool)}), arena {st.scene_bytes / 1e6:.0f} MB, "
        f"built in {time.time() - t0:.1f} s (" + ", ".join(f"{k[:-2]} {v:.2f} s" for k, v in times.items()) + ")")
    return srv, pool


def serving_run(mrt, srv, pool, rate, duration, plain=False):
    from mobileraytracer_amd import serving

    t0 = time.time()
    res = serving.run_stream(srv, pool, rate, duration, seed=4, plain=plain)
    res["config"] = (f"{len(pool)} resident scenes {min(p.triangles for p in pool)}-"
                     f"{max(p.triangles for p in pool)} triangles; Poisson arrivals; 256x256..1920x1080; "
                     f"1-64 spp; PathTracer depth 6 RR 3; FP32 accumulation")
    log(f"serving at {rate} req/s: {res['done']} of {res['requests']} requests in {time.time() - t0:.1f} s, "
        f"p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms, {res['mrays_per_s']} Mrays/s")
    if "plain" in res:
        pl = res["plain"]
        log(f"plain GPU path, one request at a time in arrival order: p50 {pl['fifo_p50_ms']} ms p99 "
            f"{pl['fifo_p99_ms']} ms (alone p50 {pl['alone_p50_ms']} ms p99 {pl['alone_p99_ms']} ms)")
    return res


if __name__ == "__main__":
    sys.exit(main())
