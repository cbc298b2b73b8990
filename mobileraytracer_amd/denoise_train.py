"""Training and evaluation of the denoiser on frames rendered by the engine.

    python -m mobileraytracer_amd.denoise_train --out weights.npz --eval-json eval.json

Pairs: regions of 1920x1080 PathTracer frames (maximum depth 6, Russian
roulette from depth 3, Halton pixel jitter) of seeded triangle soups and
displaced icosphere clusters (1% of the triangles emissive) from random poses
around the unit cube.  A pair is a block of 3 x 4 of the renderer's 16 x 16
tiles (360 x 268 pixels) at a random place in the frame, rendered alone
(tile_skip), so its pixels see the scene at the 1080p frame's angular
resolution and jitter.  Input: the mean of samples [0, 4) and the AOVs of
the same camera rays; target: the mean of samples [4, 4 + T) (T = 1024 by
default), independent of the input.  Training: mrt_denoiser_trainer_* on the
GPU (BF16 MFMA forward, backward-data and weight-gradient kernels, FP32
master weights, Adam) on random 256x256 crops with flips, loss = mean squared
error of log(1 + radiance), for a fixed wall-clock budget with a cosine
learning-rate decay.

Evaluation, on scenes and poses not in the training set (the 1M-triangle
soup of the benchmark with its camera, and an icosphere cluster): relative
MSE, mean over pixels and channels of (x - ref)^2 / (ref^2 + 0.01), against
the mean of 1024 independent samples (from sample 1024), for the plain
frames of 4, 64, 128, 256 and 512 samples and the denoised 4-sample frame
(mrt_denoiser_run, BF16 on MFMA), with the time each took, and the plain
render's equal-quality point (the samples and time at which it matches the
denoised frame's relMSE).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import denoise

W_FULL, H_FULL = 1920, 1080
SPP_IN = 4
REGION_TILES = (3, 4)  # tiles across, down


def log(msg: str) -> None:
    print(f"[denoise] {msg}", file=sys.stderr, flush=True)


def pose_camera(mrt, abi, rng, w, h):
    """Camera on a shell of radius 0.9-1.5 around the unit cube's centre,
    looking near it; 45 degrees on the shorter side."""
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
    return mrt.make_camera(abi.CAMERA_PERSPECTIVE, tuple(pos.tolist()), tuple(at.tolist()), (0, 1, 0), fh, fv)


def frame_desc(mrt, abi, cam, w, h, first, end, seed, tiles=None):
    return mrt.render_desc(abi.SHADER_PATHTRACER, w, h, end, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, seed),
                           rr=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 1),
                           light=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 2),
                           hemi=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 3), depth_min=2, depth_max=6,
                           kernel=abi.KERNEL_WAVEFRONT, first_sample=first, tiles=tiles)


def render_mean(scene, desc) -> tuple[np.ndarray, float]:
    acc = np.zeros((desc.height, desc.width, 3), np.float32)
    t0 = time.perf_counter()
    scene.render(desc, accum=acc)
    ms = (time.perf_counter() - t0) * 1e3
    return acc / np.float32(desc.samples_pixel - desc.first_sample), ms


def make_scene(mrt, kind, seed, n):
    if kind == "soup":
        return mrt.soup_scene(seed, n, max(1, n // 100))
    sub = int(min(8, max(2, round(math.log(max(n, 160) / 160.0, 4)))))
    spheres = max(1, int(round(n / (20 * 4 ** sub))))
    return mrt.icosphere_scene(seed, spheres, sub, max(1, spheres * 20 * 4 ** sub // 100))


def render_pairs(mrt, abi, ctx, n_pairs, n_scenes, spp_tgt, seed, tri_lo, tri_hi, budget_s):
    """(inputs (N, 268, 360, 10), log-space targets (N, 268, 360, 3)), FP32:
    up to n_pairs regions over n_scenes scenes, stopping early when the
    rendering time passes budget_s."""
    rng = np.random.default_rng(seed)
    tw, th = W_FULL // 16, H_FULL // 16
    ta, td = REGION_TILES
    rw, rh = ta * tw, td * th
    xs = np.zeros((n_pairs, rh, rw, 10), np.float32)
    ts = np.zeros((n_pairs, rh, rw, 3), np.float32)
    # Frame buffers in device memory (torch allocates them; the renders
    # write them in place); only each region is copied to the host.
    import torch

    dev = torch.device("cuda:0")
    bm = torch.zeros((H_FULL, W_FULL), dtype=torch.int32, device=dev)
    acc = torch.zeros((H_FULL, W_FULL, 3), dtype=torch.float32, device=dev)
    aov = torch.zeros((H_FULL, W_FULL, 7), dtype=torch.float32, device=dev)
    per_scene = -(-n_pairs // n_scenes)
    done = 0
    t_start = time.perf_counter()
    for i in range(n_scenes):
        if done >= n_pairs or time.perf_counter() - t_start > budget_s:
            break
        kind = "soup" if i % 2 == 0 else "icospheres"
        n = int(round(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))))
        t0 = time.perf_counter()
        scene = mrt.Scene(ctx, make_scene(mrt, kind, 1000 + seed + i, n))
        scene.build(abi.BVH)
        k0 = done
        for _ in range(per_scene):
            if done >= n_pairs or time.perf_counter() - t_start > budget_s:
                break
            cam = pose_camera(mrt, abi, rng, W_FULL, H_FULL)
            s = int(rng.integers(1 << 20)) * 16
            cx, cy = int(rng.integers(0, 16 - ta + 1)), int(rng.integers(0, 16 - td + 1))
            tiles = [(cy + r) * 16 + cx + c for r in range(td) for c in range(ta)]
            x0, y0 = cx * tw, cy * th
            acc.zero_()
            bm.zero_()
            d_in = frame_desc(mrt, abi, cam, W_FULL, H_FULL, 0, SPP_IN, s, tiles)
            scene.render_with_aovs(d_in, bitmap=bm, accum=acc, aov=aov, mem=abi.MEM_DEVICE)
            rad = acc[y0:y0 + rh, x0:x0 + rw].cpu().numpy() / np.float32(SPP_IN)
            a = aov[y0:y0 + rh, x0:x0 + rw].cpu().numpy()
            xs[done, ..., 0:3] = np.log1p(np.maximum(rad, 0))
            xs[done, ..., 3:9] = a[..., 0:6]
            xs[done, ..., 9] = a[..., 6]
            acc.zero_()
            # The target in passes of up to 64 samples, each adding to acc:
            # a frame's per-sample radiance (12 bytes per pixel and sample)
            # must stay within 2 GiB for sample-parallel slots.
            for f in range(SPP_IN, SPP_IN + spp_tgt, 64):
                bm.zero_()
                scene.render(frame_desc(mrt, abi, cam, W_FULL, H_FULL, f, min(f + 64, SPP_IN + spp_tgt), s, tiles),
                             bitmap=bm, accum=acc, mem=abi.MEM_DEVICE)
            tgt = acc[y0:y0 + rh, x0:x0 + rw].cpu().numpy()
            ts[done] = np.log1p(np.maximum(tgt / np.float32(spp_tgt), 0))
            done += 1
        scene.close()
        log(f"scene {i + 1}/{n_scenes} ({kind}, {n} triangles): {done - k0} pairs in {time.perf_counter() - t0:.1f} s "
            f"({done} of {n_pairs})")
    return xs[:done], ts[:done]


def train(mrt, ctx, x_all, t_all, seconds, batch, crop, lr, seed, levels=5, channels=denoise.DEFAULT_CHANNELS):
    """Adam steps of mrt_denoiser_trainer on the pairs for `seconds` of wall
    clock, learning rate lr with a cosine decay to 2%; returns (flat FP32
    weights, steps, last loss, ms per step, algorithmic TFLOP/s)."""
    w0 = denoise.init_weights(levels, channels, seed)
    tr = mrt.DenoiserTrainer(ctx, w0, mrt.denoiser_desc(levels, channels),
                             mrt.denoiser_train_desc(batch=batch, crop=crop, seed=seed))
    tr.set_pairs(x_all, t_all)
    t0 = time.perf_counter()
    last = t0
    loss = float("nan")
    ema = None
    while True:
        el = time.perf_counter() - t0
        if el >= seconds:
            break
        cur = lr * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * min(1.0, el / seconds))))
        loss = tr.step(50, cur)
        ema = loss if ema is None else 0.9 * ema + 0.1 * loss
        if time.perf_counter() - last > 15:
            last = time.perf_counter()
            st = tr.stats()
            log(f"step {st.steps}, {el:.0f} s, loss {ema:.5f}, {st.ms_per_step:.2f} ms per step")
    st = tr.stats()
    w = tr.weights()
    tr.close()
    tflops = st.flops_per_step / (st.ms_per_step / 1e3) / 1e12
    log(f"trained {st.steps} steps in {time.perf_counter() - t0:.1f} s, loss {ema:.5f}, "
        f"{st.ms_per_step:.2f} ms per step ({tflops:.0f} TFLOP/s)")
    return w, int(st.steps), float(ema), float(st.ms_per_step), float(tflops)


def rel_mse(x, ref):
    return float(np.mean((x.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 0.01)))


PLAIN_SPP = (4, 64, 128, 256, 512)
REF_FIRST = 1024  # plain frames use samples below 512; 1920 x 1080 x 2048 < 2^32


def render_plain(scene, mrt, abi, cam, spp, seed):
    """Mean of samples [0, spp) rendered in passes of up to 64 samples (a
    frame's per-sample radiance must stay within 2 GiB for sample-parallel
    slots, so one 512-sample call would run one slot per pixel), and the
    wall-clock milliseconds of the passes (host buffers)."""
    acc = np.zeros((H_FULL, W_FULL, 3), np.float32)
    t0 = time.perf_counter()
    for f in range(0, spp, 64):
        scene.render(frame_desc(mrt, abi, cam, W_FULL, H_FULL, f, min(f + 64, spp), seed), accum=acc)
    return acc / np.float32(spp), (time.perf_counter() - t0) * 1e3


def equal_quality(spps, rels, times, target):
    """Samples and milliseconds at which the plain render reaches relMSE
    `target`: log-log interpolation of relMSE between the two measured sample
    counts that bracket it, time linear in samples; None above the largest."""
    for k in range(len(spps) - 1):
        r1, r2 = rels[k], rels[k + 1]
        if r1 >= target >= r2 and r1 > r2:
            f = (math.log(target) - math.log(r1)) / (math.log(r2) - math.log(r1))
            spp = math.exp(math.log(spps[k]) + f * (math.log(spps[k + 1]) - math.log(spps[k])))
            ms = times[k] + (spp - spps[k]) * (times[k + 1] - times[k]) / (spps[k + 1] - spps[k])
            return spp, ms
    return None, None


def evaluate(mrt, abi, ctx, dn, spp_ref, seed):
    """relMSE of the plain frames of 4 to 512 samples and of the denoised
    4-sample frame against `spp_ref` further, independent samples (from
    sample 1024), for the benchmark soup and an icosphere cluster, and the
    plain render's equal-quality point: the samples and time it needs to
    match the denoised frame's relMSE."""
    out = []
    cases = [("soup 1M (benchmark scene and camera)", "soup", 2, 1_000_000, None),
             ("icosphere cluster 1.6M", "icospheres", 5, 1_638_400, 77)]
    for name, kind, sseed, n, pose_seed in cases:
        if kind == "soup":
            data = mrt.soup_scene(sseed, n, n // 100, 5.0, 0.005, abi.SAMPLER_RANDOM)
            ratio = np.float32(W_FULL) / np.float32(H_FULL)
            cam = mrt.make_camera(abi.CAMERA_PERSPECTIVE, (0.5, 0.5, -1.0), (0.5, 0.5, 0.5), (0, 1, 0),
                                  float(np.float32(45.0) * ratio), 45.0)
        else:
            data = mrt.icosphere_scene(sseed, 20, 6, n // 100)  # 20 spheres of 20 * 4^6 triangles
            cam = pose_camera(mrt, abi, np.random.default_rng(pose_seed), W_FULL, H_FULL)
        scene = mrt.Scene(ctx, data)
        scene.build(abi.BVH)
        s = 16 * (seed + 7)
        d4 = frame_desc(mrt, abi, cam, W_FULL, H_FULL, 0, SPP_IN, s)
        render_mean(scene, d4)  # warm-up
        x4, ms4 = render_mean(scene, d4)
        t0 = time.perf_counter()
        aov = scene.aovs(d4)
        ms_aov = (time.perf_counter() - t0) * 1e3
        # The 4-sample frame and its AOVs in one render.
        acc = np.zeros((H_FULL, W_FULL, 3), np.float32)
        t0 = time.perf_counter()
        _, aov1, _ = scene.render_with_aovs(d4, accum=acc)
        ms4_aov = (time.perf_counter() - t0) * 1e3
        if not (np.array_equal(aov1, aov) and np.array_equal(acc / np.float32(SPP_IN), x4)):
            raise RuntimeError("render_with_aovs differs from render + aovs")
        dn.run(x4, aov)
        t0 = time.perf_counter()
        den = dn.run(x4, aov)
        ms_dn_host = (time.perf_counter() - t0) * 1e3
        ms_dn = dn.last_ms()
        ref = np.zeros((H_FULL, W_FULL, 3), np.float32)
        t0 = time.perf_counter()
        for f in range(REF_FIRST, REF_FIRST + spp_ref, 64):
            scene.render(frame_desc(mrt, abi, cam, W_FULL, H_FULL, f, min(f + 64, REF_FIRST + spp_ref), s), accum=ref)
        ms_ref = (time.perf_counter() - t0) * 1e3
        ref /= np.float32(spp_ref)
        plain_rel, plain_ms = [], []
        for k in PLAIN_SPP:
            xk, msk = render_plain(scene, mrt, abi, cam, k, s)
            plain_rel.append(rel_mse(xk, ref))
            plain_ms.append(round(msk, 2))
        scene.close()
        rel_den = rel_mse(den, ref)
        ms_den_frame = ms4_aov + ms_dn
        eq_spp, eq_ms = equal_quality(PLAIN_SPP, plain_rel, plain_ms, rel_den)
        r = {"scene": name, "triangles": n, "relmse_4spp": plain_rel[0], "relmse_denoised_4spp": rel_den,
             "relmse_64spp": plain_rel[1], "plain_spp": list(PLAIN_SPP), "plain_relmse": plain_rel,
             "plain_ms": plain_ms, "ms_render_4spp": round(ms4, 2), "ms_aov": round(ms_aov, 2),
             "ms_render_4spp_with_aovs": round(ms4_aov, 2), "ms_denoise_device": round(ms_dn, 3),
             "ms_denoise_call": round(ms_dn_host, 2), "ms_denoised_frame": round(ms_den_frame, 2),
             "ms_render_64spp": plain_ms[1], "equal_quality_spp": None if eq_spp is None else round(eq_spp, 1),
             "equal_quality_ms": None if eq_ms is None else round(eq_ms, 1),
             "equal_quality_speedup": None if eq_ms is None else round(eq_ms / ms_den_frame, 1),
             "ms_render_ref": round(ms_ref, 1), "spp_ref": spp_ref, "ref_first_sample": REF_FIRST}
        log(json.dumps(r))
        out.append(r)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="weights file (.npz) to write")
    ap.add_argument("--eval-json", default=None, help="write the evaluation here")
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--scenes", type=int, default=20)
    ap.add_argument("--spp-target", type=int, default=1024)
    ap.add_argument("--render-budget", type=float, default=660.0,
                    help="seconds of pair rendering at most (fewer pairs if it runs out)")
    ap.add_argument("--spp-ref", type=int, default=1024)
    ap.add_argument("--tri-lo", type=int, default=50_000)
    ap.add_argument("--tri-hi", type=int, default=1_000_000)
    ap.add_argument("--train-seconds", type=float, default=150.0)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--eval-only", default=None, metavar="WEIGHTS",
                    help="skip training: evaluate this weights file (--out is ignored)")
    args = ap.parse_args()

    # torch (device buffers of the pair renders) is initialised before the
    # library is loaded: its wheel carries a HIP runtime of the same soname
    # as the system one, and the first one loaded serves both.
    import torch

    torch.cuda.init()
    import mobileraytracer_amd as mrt
    from mobileraytracer_amd import abi

    ctx = mrt.Context(0)
    if args.eval_only:
        flat_bf16, levels, channels, ds = denoise.load_weights(args.eval_only)
        dn = mrt.Denoiser(ctx, flat_bf16, mrt.denoiser_desc(levels, channels, ds))
        res = evaluate(mrt, abi, ctx, dn, args.spp_ref, args.seed)
        dn.close()
        if args.eval_json:
            with np.load(args.eval_only, allow_pickle=False) as z:
                meta = str(z["meta"])
            with open(args.eval_json, "w") as f:
                json.dump({"training": json.loads(meta) if meta else {}, "evaluation": res}, f, indent=1)
        ctx.close()
        return 0
    t0 = time.perf_counter()
    x_all, t_all = render_pairs(mrt, abi, ctx, args.pairs, args.scenes, args.spp_target, args.seed, args.tri_lo,
                                args.tri_hi, args.render_budget)
    render_s = time.perf_counter() - t0
    log(f"{len(x_all)} pairs rendered in {render_s:.1f} s")
    flat, steps, loss, ms_step, tflops = train(mrt, ctx, x_all, t_all, args.train_seconds, args.batch, args.crop,
                                               args.lr, args.seed)
    n_pairs = len(x_all)
    del x_all, t_all
    meta = json.dumps({"pairs": n_pairs, "pair_region": [REGION_TILES[0] * (W_FULL // 16),
                                                          REGION_TILES[1] * (H_FULL // 16)],
                       "frame": [W_FULL, H_FULL], "spp_in": SPP_IN, "spp_target": args.spp_target,
                       "render_seconds": round(render_s, 1), "steps": steps, "loss": loss, "batch": args.batch,
                       "crop": args.crop, "train_seconds": args.train_seconds, "ms_per_step": round(ms_step, 3),
                       "tflops": round(tflops, 1), "trainer": "mrt_denoiser_trainer (BF16 MFMA kernels)",
                       "seed": args.seed})
    denoise.save_weights(args.out, flat, 5, denoise.DEFAULT_CHANNELS, 1.0, meta)
    log(f"weights written to {args.out}")
    flat_bf16, levels, channels, ds = denoise.load_weights(args.out)
    dn = mrt.Denoiser(ctx, flat_bf16, mrt.denoiser_desc(levels, channels, ds))
    res = evaluate(mrt, abi, ctx, dn, args.spp_ref, args.seed)
    dn.close()
    if args.eval_json:
        with open(args.eval_json, "w") as f:
            json.dump({"training": json.loads(meta), "evaluation": res}, f, indent=1)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
