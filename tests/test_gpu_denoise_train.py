"""The denoiser's trainer on the GPU (mrt_denoiser_trainer_*) against its
statements: the gradient of one batch against the PyTorch graph with the
trainer's BF16 rounding points (denoise.training_output, FP32 on the CPU),
Adam against the numpy statement, the crop schedule against its published
definition, a step against the gradient of the batch it gathers, and a
short training run that lowers the loss."""
from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mobileraytracer_amd as mrt  # noqa: E402
from mobileraytracer_amd import denoise  # noqa: E402

pytestmark = pytest.mark.gpu


def packed_inputs(rng, n, h, w):
    """(n, h, w, 10) network inputs of the trainer's form."""
    x = np.zeros((n, h, w, 10), np.float32)
    x[..., 0:3] = np.log1p(rng.exponential(0.5, size=(n, h, w, 3)))
    x[..., 3:6] = rng.uniform(0, 1, size=(n, h, w, 3))
    x[..., 6:9] = rng.uniform(-1, 1, size=(n, h, w, 3))
    x[..., 9] = rng.uniform(0, 3, size=(n, h, w))
    return x


def smooth_targets(x):
    """A learnable target: a 3x3 box blur of the input radiance channels."""
    r = x[..., 0:3]
    p = np.pad(r, ((0, 0), (1, 1), (1, 1), (0, 0)), mode="edge")
    out = np.zeros_like(r)
    for dy in range(3):
        for dx in range(3):
            out += p[:, dy:dy + r.shape[1], dx:dx + r.shape[2]]
    return (out / 9.0).astype(np.float32)


def flat_grads(model):
    return np.concatenate([np.concatenate([m.weight.grad.float().numpy().ravel(), m.bias.grad.float().numpy().ravel()])
                           for m in model.convs])


def statement_gradient(model, x, t):
    for p in model.parameters():
        p.grad = None
    xt = torch.from_numpy(x.transpose(0, 3, 1, 2).copy())
    tt = torch.from_numpy(t.transpose(0, 3, 1, 2).copy())
    loss = torch.mean((denoise.training_output(model, xt) - tt) ** 2)
    loss.backward()
    return loss.item(), flat_grads(model)


@pytest.mark.parametrize("levels,channels,crop,batch,seed", [
    (3, (32, 64, 32), 32, 3, 1),
    (2, (64, 96), 48, 2, 2),
    (5, denoise.DEFAULT_CHANNELS, 128, 2, 3),
])
def test_gradient_matches_statement(gpu_ctx, levels, channels, crop, batch, seed):
    print(f"trainer gradient: levels {levels} channels {channels} crop {crop} batch {batch}", flush=True)
    torch.manual_seed(seed)
    model = denoise.make_unet(levels, channels).float()
    w = denoise.flat_weights(model)
    rng = np.random.default_rng(seed)
    x = packed_inputs(rng, batch, crop, crop)
    t = np.log1p(rng.exponential(0.5, size=(batch, crop, crop, 3))).astype(np.float32)
    tr = mrt.DenoiserTrainer(gpu_ctx, w, mrt.denoiser_desc(levels, channels),
                             mrt.denoiser_train_desc(batch=batch, crop=crop))
    try:
        loss, g = tr.gradient(x, t)
        # No update: the master weights are the initial ones.
        assert np.array_equal(tr.weights(), w)
    finally:
        tr.close()
    ref_loss, ref_g = statement_gradient(model, x, t)
    # The output y is rounded to BF16; accumulation order can move a value
    # across a rounding boundary, so the loss agrees to BF16 precision.
    assert abs(loss - ref_loss) <= 2e-3 * ref_loss
    at = 0
    for k, (ci, co) in enumerate(denoise.layer_shapes(levels, channels)):
        n = co * ci * 9 + co
        a, b = g[at:at + n], ref_g[at:at + n]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err <= 2e-2, (k, ci, co, err)
        # Bias part separately (the ones-operand MFMA).
        ab, bb = a[-co:], b[-co:]
        assert np.linalg.norm(ab - bb) <= 2e-2 * np.linalg.norm(bb) + 1e-7, (k, "bias")
        at += n
    assert at == g.size


def test_adam_matches_statement(gpu_ctx):
    levels, channels, crop, batch = 3, (32, 64, 32), 32, 2
    torch.manual_seed(4)
    w = denoise.flat_weights(denoise.make_unet(levels, channels))
    rng = np.random.default_rng(4)
    tr = mrt.DenoiserTrainer(gpu_ctx, w, mrt.denoiser_desc(levels, channels),
                             mrt.denoiser_train_desc(batch=batch, crop=crop))
    try:
        p, m, v = w.copy(), np.zeros_like(w), np.zeros_like(w)
        for k in range(1, 4):
            x = packed_inputs(rng, batch, crop, crop)
            t = np.log1p(rng.exponential(0.5, size=(batch, crop, crop, 3))).astype(np.float32)
            _, g = tr.gradient(x, t, lr=1e-3)
            p, m, v = denoise.adam_step(p, g, m, v, k, 1e-3)
            got = tr.weights()
            assert np.all(np.abs(got - p) <= 2 * np.spacing(np.abs(p))), k
            p = got
        assert tr.stats().steps == 3
    finally:
        tr.close()


def reference_crops(seed, step, batch, n, h, w, crop):
    out = []
    for i in range(batch):
        k = step * batch + i
        r1 = mrt.splitmix64(seed ^ mrt.splitmix64(2 * k))
        r2 = mrt.splitmix64(seed ^ mrt.splitmix64(2 * k + 1))
        out.append((r1 % n, (r1 >> 32) % (h - crop + 1), r2 % (w - crop + 1), (r2 >> 40) & 1, (r2 >> 41) & 1))
    return np.array(out, np.uint32)


def cut(x, c, crop):
    n, oy, ox, fx, fy = (int(v) for v in c)
    a = x[n, oy:oy + crop, ox:ox + crop]
    if fx:
        a = a[:, ::-1]
    if fy:
        a = a[::-1]
    return np.ascontiguousarray(a)


def test_step_equals_gradient_of_its_crops(gpu_ctx):
    levels, channels, crop, batch, seed = 3, (32, 64, 32), 32, 4, 0xC0FFEE
    torch.manual_seed(6)
    w = denoise.flat_weights(denoise.make_unet(levels, channels))
    rng = np.random.default_rng(6)
    n, h, wd = 3, 45, 70
    x = packed_inputs(rng, n, h, wd)
    t = smooth_targets(x)
    desc = mrt.denoiser_desc(levels, channels)
    tdesc = mrt.denoiser_train_desc(batch=batch, crop=crop, seed=seed)
    a = mrt.DenoiserTrainer(gpu_ctx, w, desc, tdesc)
    b = mrt.DenoiserTrainer(gpu_ctx, w, desc, tdesc)
    try:
        a.set_pairs(x, t)
        for step in range(3):
            crops = a.crops(step)
            assert np.array_equal(crops, reference_crops(seed, step, batch, n, h, wd, crop)), step
            xb = np.stack([cut(x, c, crop) for c in crops])
            tb = np.stack([cut(t, c, crop) for c in crops])
            la = a.step(1, 2e-3)
            lb, _ = b.gradient(xb, tb, lr=2e-3)
            assert la == lb
            assert np.array_equal(a.weights(), b.weights()), step
    finally:
        a.close()
        b.close()


def test_training_lowers_the_loss(gpu_ctx):
    levels, channels, crop, batch = 3, (32, 64, 64), 32, 8
    torch.manual_seed(8)
    w = denoise.flat_weights(denoise.make_unet(levels, channels))
    rng = np.random.default_rng(8)
    x = packed_inputs(rng, 4, 64, 64)
    t = smooth_targets(x)
    tr = mrt.DenoiserTrainer(gpu_ctx, w, mrt.denoiser_desc(levels, channels),
                             mrt.denoiser_train_desc(batch=batch, crop=crop, seed=1))
    try:
        tr.set_pairs(x, t)
        first = tr.step(10, 1e-3)
        for _ in range(15):
            last = tr.step(20, 1e-3)
        st = tr.stats()
        print(f"loss {first:.5f} -> {last:.5f} over {st.steps} steps, {st.ms_per_step:.3f} ms per step", flush=True)
        assert st.steps == 310
        assert np.isfinite(last) and last < 0.25 * first
    finally:
        tr.close()


def test_trainer_refusals(gpu_ctx):
    w = denoise.flat_weights(denoise.make_unet(3, (32, 64, 32)))
    desc = mrt.denoiser_desc(3, (32, 64, 32))
    with pytest.raises(mrt.abi.MrtError):
        mrt.DenoiserTrainer(gpu_ctx, w, desc, mrt.denoiser_train_desc(batch=2, crop=40))  # not a multiple of 32
    with pytest.raises(mrt.abi.MrtError):
        mrt.DenoiserTrainer(gpu_ctx, w, desc, mrt.denoiser_train_desc(batch=65, crop=32))
    with pytest.raises(mrt.abi.MrtError):
        mrt.DenoiserTrainer(gpu_ctx, w[:-1], desc, mrt.denoiser_train_desc(batch=2, crop=32))
    tr = mrt.DenoiserTrainer(gpu_ctx, w, desc, mrt.denoiser_train_desc(batch=2, crop=32))
    try:
        with pytest.raises(mrt.abi.MrtError):
            tr.step(1, 1e-3)  # no pairs
        with pytest.raises(mrt.abi.MrtError):
            tr.set_pairs(np.zeros((1, 16, 16, 10), np.float32), np.zeros((1, 16, 16, 3), np.float32))  # < crop
        x = np.zeros((3, 32, 32, 10), np.float32)
        with pytest.raises(mrt.abi.MrtError):
            tr.gradient(x, np.zeros((3, 32, 32, 3), np.float32))  # batch above the trainer's
    finally:
        tr.close()
