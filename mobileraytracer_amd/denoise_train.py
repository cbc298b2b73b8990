"""Training and evaluation of the denoiser on frames rendered by the engine.

    python -m mobileraytracer_amd.denoise_train --out weights.npz --eval-json eval.json

Pairs: 1920x1080 PathTracer frames (maximum depth 6, Russian roulette from
depth 3, Halton pixel jitter) of seeded triangle soups and displaced
icosphere clusters (1% of the triangles emissive) from random poses around
the unit cube.  Input: the mean of samples [0, 4) and the AOVs of the same
camera rays; target: the mean of samples [4, 4 + T), independent of the
input, so the L2 optimum on noisy targets is the clean one.  Training: Adam
on random 256x256 crops with flips, loss = mean squared error of
log(1 + radiance), BF16 autocast, for a fixed wall-clock budget with a
cosine learning-rate decay.

Evaluation, on scenes and poses not in the training set (the 1M-triangle
soup of the benchmark with its camera, and an icosphere cluster): relative
MSE, mean over pixels and channels of (x - ref)^2 / (ref^2 + 0.01), against
the mean of 1024 further samples, for the plain 4-sample frame, the denoised
4-sample frame (mrt_denoiser_run, BF16 on MFMA) and the plain 64-sample
frame, with the time each took.

torch is initialised before the library is loaded: its wheel carries a HIP
runtime of the same soname as the system one, and the first one loaded
serves both.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

os.environ.setdefault("MIOPEN_FIND_MODE", "FAST")
import torch  # noqa: E402

from . import denoise  # noqa: E402

W_FULL, H_FULL = 1920, 1080
SPP_IN = 4


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


def frame_desc(mrt, abi, cam, w, h, first, end, seed):
    return mrt.render_desc(abi.SHADER_PATHTRACER, w, h, end, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0.0, seed),
                           rr=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 1),
                           light=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 2),
                           hemi=mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed + 3), depth_min=2, depth_max=6,
                           kernel=abi.KERNEL_WAVEFRONT, first_sample=first)


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


def render_pairs(mrt, abi, ctx, n_scenes, poses, spp_tgt, seed, tri_lo, tri_hi):
    """(inputs (N, 10, H, W), targets (N, 3, H, W)) as FP32 numpy arrays."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for i in range(n_scenes):
        kind = "soup" if i % 2 == 0 else "icospheres"
        n = int(round(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))))
        t0 = time.perf_counter()
        scene = mrt.Scene(ctx, make_scene(mrt, kind, 1000 + seed + i, n))
        scene.build(abi.BVH)
        for k in range(poses):
            cam = pose_camera(mrt, abi, rng, W_FULL, H_FULL)
            s = int(rng.integers(1 << 20)) * 16
            d_in = frame_desc(mrt, abi, cam, W_FULL, H_FULL, 0, SPP_IN, s)
            rad, _ = render_mean(scene, d_in)
            aov = scene.aovs(d_in)
            tgt, _ = render_mean(scene, frame_desc(mrt, abi, cam, W_FULL, H_FULL, SPP_IN, SPP_IN + spp_tgt, s))
            xs.append(denoise.pack_input(rad, aov, 1.0, 1.0, 5)[:, :H_FULL, :W_FULL])
            ts.append(tgt.transpose(2, 0, 1).copy())
        scene.close()
        log(f"scene {i + 1}/{n_scenes} ({kind}, {n} triangles): {poses} pairs in {time.perf_counter() - t0:.1f} s")
    return np.stack(xs), np.stack(ts)


def train(x_all, t_all, seconds, batch, crop, lr, seed, levels=5, channels=denoise.DEFAULT_CHANNELS,
          device="cuda:0"):
    dev = torch.device(device)
    torch.manual_seed(seed)
    model = denoise.make_unet(levels, channels).to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    x_all = torch.from_numpy(x_all).to(dev)
    t_all = torch.from_numpy(np.log1p(np.maximum(t_all, 0))).to(dev)
    N, _, H, W = x_all.shape
    g = torch.Generator(device="cpu").manual_seed(seed)
    t0 = time.perf_counter()
    last = t0
    it = 0
    ema = None
    while True:
        el = time.perf_counter() - t0
        if el >= seconds:
            break
        for pg in opt.param_groups:
            pg["lr"] = lr * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * min(1.0, el / seconds))))
        idx = torch.randint(N, (batch,), generator=g)
        oy = torch.randint(H - crop + 1, (batch,), generator=g)
        ox = torch.randint(W - crop + 1, (batch,), generator=g)
        xb = torch.stack([x_all[i, :, y:y + crop, x:x + crop] for i, y, x in zip(idx, oy, ox)])
        tb = torch.stack([t_all[i, :, y:y + crop, x:x + crop] for i, y, x in zip(idx, oy, ox)])
        if torch.rand(1, generator=g).item() < 0.5:
            xb, tb = xb.flip(3), tb.flip(3)
        if torch.rand(1, generator=g).item() < 0.5:
            xb, tb = xb.flip(2), tb.flip(2)
        with torch.autocast(dev.type, dtype=torch.bfloat16):
            y = model(xb)
        loss = torch.mean((y.float() - tb) ** 2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        it += 1
        ema = loss.item() if ema is None else 0.98 * ema + 0.02 * loss.item()
        if time.perf_counter() - last > 15:
            last = time.perf_counter()
            log(f"iteration {it}, {el:.0f} s, loss {ema:.5f}")
    log(f"trained {it} iterations in {time.perf_counter() - t0:.1f} s, final loss {ema:.5f}")
    return model.cpu().float(), it, ema


def rel_mse(x, ref):
    return float(np.mean((x.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 0.01)))


def evaluate(mrt, abi, ctx, dn, spp_ref, seed):
    """relMSE of the plain 4-sample, denoised 4-sample and plain 64-sample
    frames against 1024 further samples, for the benchmark soup and an
    icosphere cluster."""
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
        x64, ms64 = render_mean(scene, frame_desc(mrt, abi, cam, W_FULL, H_FULL, 0, 64, s))
        ref, ms_ref = render_mean(scene, frame_desc(mrt, abi, cam, W_FULL, H_FULL, 64, 64 + spp_ref, s))
        scene.close()
        r = {"scene": name, "triangles": n, "relmse_4spp": rel_mse(x4, ref), "relmse_denoised_4spp": rel_mse(den, ref),
             "relmse_64spp": rel_mse(x64, ref), "ms_render_4spp": round(ms4, 2), "ms_aov": round(ms_aov, 2),
             "ms_render_4spp_with_aovs": round(ms4_aov, 2),
             "ms_denoise_device": round(ms_dn, 3), "ms_denoise_call": round(ms_dn_host, 2),
             "ms_render_64spp": round(ms64, 2), "ms_render_ref": round(ms_ref, 1), "spp_ref": spp_ref}
        log(json.dumps(r))
        out.append(r)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="weights file (.npz) to write")
    ap.add_argument("--eval-json", default=None, help="write the evaluation here")
    ap.add_argument("--scenes", type=int, default=10)
    ap.add_argument("--poses", type=int, default=4)
    ap.add_argument("--spp-target", type=int, default=128)
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
    x_all, t_all = render_pairs(mrt, abi, ctx, args.scenes, args.poses, args.spp_target, args.seed, args.tri_lo,
                                args.tri_hi)
    log(f"{len(x_all)} pairs rendered in {time.perf_counter() - t0:.1f} s")
    model, iters, loss = train(x_all, t_all, args.train_seconds, args.batch, args.crop, args.lr, args.seed)
    del x_all, t_all
    torch.cuda.empty_cache()
    flat = denoise.flat_weights(model)
    meta = json.dumps({"pairs": args.scenes * args.poses, "spp_in": SPP_IN, "spp_target": args.spp_target,
                       "iterations": iters, "loss": loss, "batch": args.batch, "crop": args.crop,
                       "train_seconds": args.train_seconds, "seed": args.seed})
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
