"""The statements the denoiser's trainer is checked against, pinned on the CPU:
the crop schedule's splitmix64 against its published output, the numpy Adam
against torch.optim.Adam, and the trainer's forward graph
(denoise.training_output) against the inference network and, with its BF16
rounding, against the FP64 gradient of the plain network."""
from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import mobileraytracer_amd as mrt  # noqa: E402
from mobileraytracer_amd import denoise  # noqa: E402


def test_splitmix64_published_sequence():
    # splitmix64 seeded with 0: its first three outputs (the state advances by
    # 0x9E3779B97F4A7C15 before each finalisation).
    g = 0x9E3779B97F4A7C15
    assert mrt.splitmix64(0) == 0xE220A8397B1DCDAF
    assert mrt.splitmix64(g) == 0x6E789E6AA1B965F4
    assert mrt.splitmix64((2 * g) % (1 << 64)) == 0x06C45D188009454F


def test_adam_statement_matches_torch():
    """Each step of the statement, from torch's state, equals torch's step to
    an ulp or two (torch's CPU kernels round a few operations differently,
    e.g. a fused multiply-add in its lerp)."""
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=1000).astype(np.float32)
    grads = [rng.normal(size=1000).astype(np.float32) * s for s in (1.0, 0.01, 3.0, 1e-4)]
    t = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([t], lr=2e-3)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for k, g in enumerate(grads, 1):
        t.grad = torch.from_numpy(g.copy())
        opt.step()
        p1, m1, v1 = denoise.adam_step(p, g, m, v, k, 2e-3)
        pt, mt = t.detach().numpy().copy(), opt.state[t]["exp_avg"].numpy().copy()
        vt = opt.state[t]["exp_avg_sq"].numpy().copy()
        assert np.all(np.abs(m1 - mt) <= np.spacing(np.abs(m) + np.abs(g) + np.abs(mt)))
        assert np.all(np.abs(v1 - vt) <= 2 * np.spacing(np.abs(vt)))
        assert np.all(np.abs(p1 - pt) <= 2 * np.spacing(np.abs(pt)) + 1e-5 * np.abs(pt - p))
        p, m, v = pt, mt, vt


def _inputs(rng, n, h, w):
    x = np.zeros((n, 10, h, w), np.float32)
    x[:, 0:3] = np.log1p(rng.exponential(0.5, size=(n, 3, h, w)))
    x[:, 3:6] = rng.uniform(0, 1, size=(n, 3, h, w))
    x[:, 6:9] = rng.uniform(-1, 1, size=(n, 3, h, w))
    x[:, 9] = rng.uniform(0, 3, size=(n, h, w))
    return x.astype(np.float32)


@pytest.mark.parametrize("levels,channels", [(2, (32, 64)), (3, (32, 64, 32))])
def test_training_graph_is_the_inference_network(levels, channels):
    torch.manual_seed(5)
    model = denoise.make_unet(levels, channels).float()
    x = torch.from_numpy(_inputs(np.random.default_rng(1), 2, 32, 32))
    with torch.no_grad():
        plain = denoise.training_output(model, x, rounding=False)
        assert torch.equal(plain, model(x))
        rounded = denoise.training_output(model, x, rounding=True)
        assert torch.equal(rounded, model(x, bf16=True).to(torch.bfloat16).float())


def _flat_grads(model):
    return np.concatenate([np.concatenate([m.weight.grad.double().numpy().ravel(), m.bias.grad.double().numpy().ravel()])
                           for m in model.convs])


def test_rounded_gradient_close_to_fp64_gradient():
    """The BF16-rounded graph's gradient is the FP64 gradient up to BF16
    rounding; its weight order is the flat layout's."""
    torch.manual_seed(7)
    model = denoise.make_unet(2, (32, 64))
    rng = np.random.default_rng(2)
    x = torch.from_numpy(_inputs(rng, 2, 32, 32))
    t = torch.from_numpy(np.log1p(rng.exponential(0.5, size=(2, 3, 32, 32))).astype(np.float32))
    loss = torch.mean((denoise.training_output(model, x) - t) ** 2)
    loss.backward()
    g32 = _flat_grads(model)
    m64 = model.double()
    for p in m64.parameters():
        p.grad = None
    loss64 = torch.mean((denoise.training_output(m64, x.double(), rounding=False) - t.double()) ** 2)
    loss64.backward()
    g64 = _flat_grads(m64)
    assert abs(loss.item() - loss64.item()) <= 2e-2 * loss64.item()
    at = 0
    for ci, co in denoise.layer_shapes(2, (32, 64)):
        n = co * ci * 9 + co
        a, b = g32[at:at + n], g64[at:at + n]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b), (ci, co)
        at += n
    assert at == g32.size == denoise.weight_count(2, (32, 64))
    # The FP64 gradient is the loss's derivative: central differences on a
    # few weights of the first and last layers.
    flat64 = denoise.flat_weights(m64).astype(np.float64)
    for idx in (0, 5, flat64.size - 1, flat64.size - 10):
        h = 1e-6
        vals = []
        for s in (1, -1):
            f = flat64.copy()
            f[idx] += s * h
            with torch.no_grad():
                at = 0
                for mm in m64.convs:
                    for p in (mm.weight, mm.bias):
                        p.copy_(torch.from_numpy(f[at:at + p.numel()].reshape(p.shape)))
                        at += p.numel()
                vals.append(torch.mean((denoise.training_output(m64, x.double(), rounding=False) - t.double()) ** 2)
                            .item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g64[idx]) <= 1e-6 * max(1.0, abs(g64[idx])) + 1e-9, idx
