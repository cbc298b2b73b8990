"""The denoiser's FP32 reference (mobileraytracer_amd.denoise, PyTorch on the
CPU) pinned against a loop-by-loop numpy statement of the network, BF16
rounding against known answers and torch's own cast, and the library's
weight layout count against the torch module's parameters.  No GPU."""
from __future__ import annotations

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from mobileraytracer_amd import abi, denoise


@pytest.fixture(scope="module")
def torch():
    # Imported by the tests that use it, not at collection: a GPU test
    # session then keeps the HIP runtime the library itself loads.
    return pytest.importorskip("torch")


def conv_np(x, w, b):
    """3x3 convolution, zero padding 1, by explicit sums: x (cin, H, W), w
    (cout, cin, 3, 3)."""
    cin, H, W = x.shape
    xp = np.zeros((cin, H + 2, W + 2), np.float64)
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((w.shape[0], H, W), np.float64)
    for o in range(w.shape[0]):
        acc = np.full((H, W), float(b[o]))
        for i in range(cin):
            for dy in range(3):
                for dx in range(3):
                    acc += w[o, i, dy, dx] * xp[i, dy:dy + H, dx:dx + W]
        out[o] = acc
    return out


def unet_np(x, convs, levels):
    """The network as written in the module docstring, in float64."""
    relu = lambda t: np.maximum(t, 0)  # noqa: E731
    k = 0
    skips = []
    h = x
    for l in range(levels):
        if l > 0:
            s = skips[-1]
            c, H, W = s.shape
            h = s.reshape(c, H // 2, 2, W // 2, 2).max(axis=(2, 4))
        h = relu(conv_np(h, *convs[k]))
        h = relu(conv_np(h, *convs[k + 1]))
        k += 2
        skips.append(h)
    d = skips[-1]
    for l in range(levels - 2, -1, -1):
        u = d.repeat(2, axis=1).repeat(2, axis=2)
        h = relu(conv_np(np.concatenate([u, skips[l]], 0), *convs[k]))
        d = relu(conv_np(h, *convs[k + 1]))
        k += 2
    return conv_np(d, *convs[k])


@pytest.mark.parametrize("levels,channels", [(1, (4,)), (2, (4, 6)), (3, (3, 5, 4))])
def test_unet_matches_loop_statement(torch, levels, channels):
    torch.manual_seed(levels)
    model = denoise.make_unet(levels, channels)
    rng = np.random.default_rng(levels)
    m = 1 << (levels - 1)
    x = rng.normal(size=(denoise.IN_CHANNELS, 2 * m, 3 * m)).astype(np.float32)
    with torch.no_grad():
        y = model(torch.from_numpy(x)[None])[0].numpy()
    convs = [(c.weight.detach().double().numpy(), c.bias.detach().double().numpy()) for c in model.convs]
    ref = unet_np(x.astype(np.float64), convs, levels)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-5)


def test_layer_order_and_counts(torch):
    assert denoise.layer_shapes(2, (32, 64)) == [(10, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
                                                 (32, 3)]
    for levels, ch in [(5, denoise.DEFAULT_CHANNELS), (1, (32,)), (3, (32, 64, 96)), (4, (64, 32, 128, 256))]:
        model = denoise.make_unet(levels, ch)
        n = sum(p.numel() for p in model.parameters())
        assert denoise.weight_count(levels, ch) == n
        assert mrt.denoiser_weight_count(mrt.denoiser_desc(levels, ch)) == n
        flat = denoise.flat_weights(model)
        assert flat.size == n
        other = denoise.make_unet(levels, ch)
        denoise.set_flat_weights(other, flat)
        assert np.array_equal(denoise.flat_weights(other), flat)
    # All-zero channels select the default network.
    d = abi.DenoiserDesc()
    assert mrt.denoiser_weight_count(d) == denoise.weight_count()


@pytest.mark.parametrize("levels,channels", [(6, (32,) * 5), (2, (32, 48)), (2, (32, 0)), (2, (288, 32))])
def test_invalid_network_shapes(levels, channels):
    d = abi.DenoiserDesc()
    d.levels = levels
    for i, c in enumerate(channels[:5]):
        d.channels[i] = c
    with pytest.raises(abi.MrtError):
        mrt.denoiser_weight_count(d)


def test_bf16_rounding(torch):
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    # Ties go to the even mantissa; above the tie rounds up.
    cases = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -2.5, 0.0], np.float32)
    got = denoise.bf16_round(cases)
    np.testing.assert_array_equal(got, want)
    assert got[0] == one
    rng = np.random.default_rng(0)
    a = (rng.normal(size=10000) * np.exp(rng.normal(size=10000) * 5)).astype(np.float32)
    t = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(denoise.bf16_round(a), t)
    np.testing.assert_array_equal(denoise.from_bf16_bits(denoise.bf16_bits(a)), t)


def test_pack_input_known_answer():
    rad = np.array([[[3.0, 0.5, -1.0]]], np.float32)
    aov = np.array([[[0.1, 0.2, 0.3, 0.0, 1.0, 0.0, 2.0]]], np.float32)
    x = denoise.pack_input(rad, aov, radiance_scale=0.5, depth_scale=0.25, levels=3)
    assert x.shape == (10, 4, 4)
    np.testing.assert_allclose(x[:, 0, 0], [np.log(2.5), np.log(1.25), 0.0, 0.1, 0.2, 0.3, 0.0, 1.0, 0.0, 0.5],
                               rtol=1e-6)
    assert not x[:, 1:, :].any() and not x[:, :, 1:].any()
    y = np.zeros((3, 4, 4), np.float32)
    y[:, 0, 0] = [np.log(3.0), -1.0, 0.0]
    np.testing.assert_allclose(denoise.unpack_output(y, 1, 1)[0, 0], [2.0, 0.0, 0.0], rtol=1e-6)


def test_layer_bounds_known_answer():
    """Per-kernel FLOPs and least HBM bytes of a 2-level network on a 30x32
    frame (padded to 30x32: even sides), written out by hand."""
    ks = denoise.layer_bounds(30, 32, 2, (32, 64))
    assert [k["kernel"] for k in ks] == ["enc0a", "enc0b", "pool0", "enc1a", "enc1b", "dec0a", "dec0b", "out"]
    p0, p1 = 30 * 32, 15 * 16
    wb = lambda ci, co: 9 * ci * co * 2 + 4 * co  # noqa: E731
    expect = {
        "enc0a": (2 * 9 * 10 * 32 * p0, 40 * p0 + wb(10, 32) + 64 * p0),
        "enc0b": (2 * 9 * 32 * 32 * p0, 64 * p0 + wb(32, 32) + 64 * p0),
        "pool0": (0, 64 * p0 + 64 * p1),
        "enc1a": (2 * 9 * 32 * 64 * p1, 64 * p1 + wb(32, 64) + 128 * p1),
        "enc1b": (2 * 9 * 64 * 64 * p1, 128 * p1 + wb(64, 64) + 128 * p1),
        "dec0a": (2 * 9 * 96 * 32 * p0, 128 * p1 + 64 * p0 + wb(96, 32) + 64 * p0),
        "dec0b": (2 * 9 * 32 * 32 * p0, 64 * p0 + wb(32, 32) + 64 * p0),
        "out": (2 * 9 * 32 * 3 * p0, 64 * p0 + wb(32, 3) + 12 * p0),
    }
    for k in ks:
        assert (k["flops"], k["bytes"]) == expect[k["kernel"]], k["kernel"]
    # The FLOPs are the network's algorithmic count; 1080 rows pad to 1088.
    for h, w in [(1080, 1920), (64, 48)]:
        assert sum(k["flops"] for k in denoise.layer_bounds(h, w)) == denoise.algorithmic_flops(h, w)
    t = denoise.layer_roofline_s(30, 32, 1e12, 1e15, 2, (32, 64))
    assert t == pytest.approx(sum(max(b / 1e12, f / 1e15) for f, b in expect.values()))
