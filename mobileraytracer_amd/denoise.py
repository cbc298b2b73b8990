"""The denoiser's U-Net in PyTorch (training, and the FP32 statement of the
network that mrt_denoiser_* runs in BF16 on MFMA) and its weights file.

Network (levels L, channels c[0..L-1]):
  x0 = [log(1 + max(radiance, 0)) (3), albedo (3), normal (3), depth (1)],
       zero-padded right and bottom to multiples of 2^(L-1);
  e_l = relu(conv(relu(conv(in_l)))), in_0 = x0, in_l = maxpool2(e_{l-1});
  d_l = relu(conv(relu(conv(cat(up2(d_{l+1}), e_l))))), d_{L-1} = e_{L-1};
  y = conv(d_0) (3 channels), radiance = max(exp(y) - 1, 0).
Every conv is 3x3 with zero padding 1 and a bias.  The library rounds the
input, the weights and every activation after a ReLU to BF16 (nearest even)
and accumulates in FP32; `forward(..., bf16=True)` does the same rounding.

torch is imported lazily: the inference path is the C ABI, which needs no
torch.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

IN_CHANNELS = 10
OUT_CHANNELS = 3
DEFAULT_CHANNELS = (32, 64, 96, 128, 128)
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "denoiser_unet5.npz"


def layer_shapes(levels: int = 5, channels=DEFAULT_CHANNELS) -> list[tuple[int, int]]:
    """(cin, cout) of every conv in weight order: encoder 0..L-1 (two each),
    decoder L-2..0 (two each), output."""
    c = list(channels)[:levels]
    out = []
    for l in range(levels):
        out.append((IN_CHANNELS if l == 0 else c[l - 1], c[l]))
        out.append((c[l], c[l]))
    for l in range(levels - 2, -1, -1):
        out.append((c[l + 1] + c[l], c[l]))
        out.append((c[l], c[l]))
    out.append((c[0], OUT_CHANNELS))
    return out


def weight_count(levels: int = 5, channels=DEFAULT_CHANNELS) -> int:
    return sum(co * ci * 9 + co for ci, co in layer_shapes(levels, channels))


def algorithmic_flops(height: int, width: int, levels: int = 5, channels=DEFAULT_CHANNELS) -> int:
    """Multiply-adds x 2 of the network's convolutions over a frame padded to
    multiples of 2^(levels-1), with the logical channel counts (10 in, 3
    out; the library pads both to 32)."""
    m = 1 << (levels - 1)
    H, W = (height + m - 1) // m * m, (width + m - 1) // m * m
    shapes = layer_shapes(levels, channels)
    lvl = [l for l in range(levels) for _ in range(2)] + [l for l in range(levels - 2, -1, -1) for _ in range(2)] + [0]
    return sum(2 * 9 * ci * co * (H >> l) * (W >> l) for (ci, co), l in zip(shapes, lvl))


def layer_bounds(height: int, width: int, levels: int = 5, channels=DEFAULT_CHANNELS) -> list[dict]:
    """Per kernel of one network run (convolutions and 2x2 max pools, in run
    order): algorithmic FLOPs (logical channels, as algorithmic_flops) and the
    least HBM bytes the kernel can move: every tensor it reads once and every
    tensor it writes once, BF16 activations (2 bytes per channel), the FP32
    network input (radiance 3 + AOVs 7 floats per frame pixel) and FP32
    radiance out, weights and biases.  The decoder's first convolution reads
    the coarser level's tensor at its own size (the upsampling happens while
    staging)."""
    m = 1 << (levels - 1)
    H, W = (height + m - 1) // m * m, (width + m - 1) // m * m
    c = list(channels)[:levels]
    frame = height * width
    px = [(H >> l) * (W >> l) for l in range(levels)]
    out = []

    def conv(name, l, ci, co, rd):
        wb = 9 * ci * co * 2 + 4 * co
        out.append({"kernel": name, "flops": 2 * 9 * ci * co * px[l], "bytes": rd + wb + 2 * co * px[l]})

    for l in range(levels):
        ci = IN_CHANNELS if l == 0 else c[l - 1]
        conv(f"enc{l}a", l, ci, c[l], 40 * frame if l == 0 else 2 * ci * px[l])
        conv(f"enc{l}b", l, c[l], c[l], 2 * c[l] * px[l])
        if l + 1 < levels:
            out.append({"kernel": f"pool{l}", "flops": 0, "bytes": 2 * c[l] * (px[l] + px[l + 1])})
    for l in range(levels - 2, -1, -1):
        conv(f"dec{l}a", l, c[l + 1] + c[l], c[l], 2 * c[l + 1] * px[l + 1] + 2 * c[l] * px[l])
        conv(f"dec{l}b", l, c[l], c[l], 2 * c[l] * px[l])
    wb = 9 * c[0] * OUT_CHANNELS * 2 + 4 * OUT_CHANNELS
    out.append({"kernel": "out", "flops": 2 * 9 * c[0] * OUT_CHANNELS * px[0],
                "bytes": 2 * c[0] * px[0] + wb + 12 * frame})
    return out


def layer_roofline_s(height: int, width: int, hbm_bytes_per_s: float, flops_per_s: float, levels: int = 5,
                     channels=DEFAULT_CHANNELS) -> float:
    """Seconds one network run takes at the roofline of each of its kernels:
    the sum over layer_bounds of max(bytes / HBM bandwidth, FLOPs / peak)."""
    return sum(max(k["bytes"] / hbm_bytes_per_s, k["flops"] / flops_per_s)
               for k in layer_bounds(height, width, levels, channels))


def bf16_round(a: np.ndarray) -> np.ndarray:
    """FP32 -> nearest-even BF16, returned as FP32 (finite inputs)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def bf16_bits(a: np.ndarray) -> np.ndarray:
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def make_unet(levels: int = 5, channels=DEFAULT_CHANNELS):
    import torch
    import torch.nn.functional as F

    class UNet(torch.nn.Module):
        def __init__(self):
            super().__init__()
            self.levels = levels
            self.channels = tuple(channels)[:levels]
            self.convs = torch.nn.ModuleList(
                [torch.nn.Conv2d(ci, co, 3, padding=1) for ci, co in layer_shapes(levels, channels)])

        def forward(self, x, bf16: bool = False):
            """x: (N, 10, H, W) with H, W multiples of 2^(levels-1)."""
            if bf16:
                def q(t):
                    return t.to(torch.bfloat16).to(torch.float32)

                def conv(k, t):
                    m = self.convs[k]
                    return F.conv2d(t, q(m.weight), m.bias, padding=1)
                x = q(x)
            else:
                def q(t):
                    return t

                def conv(k, t):
                    return self.convs[k](t)
            k = 0
            skips = []
            h = x
            for l in range(self.levels):
                if l > 0:
                    h = F.max_pool2d(skips[-1], 2)
                h = q(F.relu(conv(k, h)))
                h = q(F.relu(conv(k + 1, h)))
                k += 2
                skips.append(h)
            d = skips[-1]
            for l in range(self.levels - 2, -1, -1):
                u = F.interpolate(d, scale_factor=2, mode="nearest")
                h = q(F.relu(conv(k, torch.cat([u, skips[l]], 1))))
                d = q(F.relu(conv(k + 1, h)))
                k += 2
            return conv(k, d)

    return UNet()


def init_weights(levels: int = 5, channels=DEFAULT_CHANNELS, seed: int = 0) -> np.ndarray:
    """Initial flat FP32 weights: every weight and bias uniform in
    +-1/sqrt(9 cin) (the bound of torch.nn.Conv2d's default initialisation),
    from a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in layer_shapes(levels, channels):
        bound = 1.0 / np.sqrt(9.0 * ci)
        parts.append(rng.uniform(-bound, bound, size=co * ci * 9))
        parts.append(rng.uniform(-bound, bound, size=co))
    return np.concatenate(parts).astype(np.float32)


def training_output(model, x, rounding: bool = True):
    """The trainer's forward pass as an autograd graph (mrt_denoiser_trainer_*):
    x (N, 10, H, W) packed inputs, H and W multiples of 8 * 2^(levels-1);
    returns y (N, 3, H, W).  With `rounding`, BF16 rounding (nearest even) at
    the trainer's points, both ways: the input, the weights (straight
    through: their gradient is not rounded), every ReLU output, the pooled
    tensors, the decoder's concatenated input and the output y are rounded
    in the forward pass, and the gradient arriving at each of them (but the
    input and the weights) is rounded in the backward pass, as the trainer
    stores it in BF16.  Without it, the plain FP32 (or FP64) network."""
    import torch
    import torch.nn.functional as F

    class RoundBoth(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            return t.to(torch.bfloat16).to(t.dtype)

        @staticmethod
        def backward(ctx, g):
            return g.to(torch.bfloat16).to(g.dtype)

    class RoundForward(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            return t.to(torch.bfloat16).to(t.dtype)

        @staticmethod
        def backward(ctx, g):
            return g

    q = RoundBoth.apply if rounding else (lambda t: t)
    qw = RoundForward.apply if rounding else (lambda t: t)

    def conv(k, t):
        m = model.convs[k]
        return F.conv2d(t, qw(m.weight), m.bias, padding=1)

    L = model.levels
    k = 0
    skips = []
    h = qw(x)
    for l in range(L):
        if l > 0:
            h = q(F.max_pool2d(skips[-1], 2))
        h = q(F.relu(conv(k, h)))
        h = q(F.relu(conv(k + 1, h)))
        k += 2
        skips.append(h)
    d = skips[-1]
    for l in range(L - 2, -1, -1):
        u = F.interpolate(d, scale_factor=2, mode="nearest")
        h = q(F.relu(conv(k, q(torch.cat([u, skips[l]], 1)))))
        d = q(F.relu(conv(k + 1, h)))
        k += 2
    return q(conv(k, d))


def adam_step(p, g, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One update of the trainer's Adam in FP32 numpy, torch.optim.Adam's
    arithmetic: the hyperparameters are doubles, rounded to FP32 where they
    meet the tensors (1 - b1, b2, 1 - b2, eps, lr / (1 - b1^t),
    sqrt(1 - b2^t)); returns (p, m, v); step counts from 1."""
    f = np.float32
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    m = (m + f(1.0 - b1) * (g - m)).astype(np.float32)
    v = (v * f(b2) + (f(1.0 - b2) * g) * g).astype(np.float32)
    step_size, bc2_sqrt = f(lr / (1.0 - b1 ** step)), f(np.sqrt(1.0 - b2 ** step))
    denom = (np.sqrt(v) / bc2_sqrt + f(eps)).astype(np.float32)
    p = (p + (-step_size) * (m / denom)).astype(np.float32)
    return p, m, v


def flat_weights(model) -> np.ndarray:
    """The library's flat FP32 weight vector of a UNet."""
    parts = []
    for m in model.convs:
        parts.append(m.weight.detach().float().cpu().numpy().ravel())
        parts.append(m.bias.detach().float().cpu().numpy().ravel())
    return np.concatenate(parts).astype(np.float32)


def set_flat_weights(model, flat: np.ndarray) -> None:
    import torch

    flat = np.asarray(flat, np.float32)
    at = 0
    with torch.no_grad():
        for m in model.convs:
            for p in (m.weight, m.bias):
                n = p.numel()
                p.copy_(torch.from_numpy(flat[at:at + n].reshape(p.shape)))
                at += n
    if at != flat.size:
        raise ValueError(f"weight vector has {flat.size} entries, the network {at}")


def pack_input(radiance: np.ndarray, aov: np.ndarray, radiance_scale: float = 1.0, depth_scale: float = 1.0,
               levels: int = 5) -> np.ndarray:
    """(10, H, W) FP32 network input of a frame, padded to multiples of
    2^(levels-1) (the library's pack kernel)."""
    h, w = radiance.shape[:2]
    m = 1 << (levels - 1)
    H, W = (h + m - 1) // m * m, (w + m - 1) // m * m
    x = np.zeros((IN_CHANNELS, H, W), np.float32)
    r = np.maximum(np.asarray(radiance, np.float32) * np.float32(radiance_scale), 0).astype(np.float32)
    x[0:3, :h, :w] = np.log1p(r).transpose(2, 0, 1)
    x[3:9, :h, :w] = np.asarray(aov, np.float32)[..., :6].transpose(2, 0, 1)
    x[9, :h, :w] = np.asarray(aov, np.float32)[..., 6] * np.float32(depth_scale)
    return x


def unpack_output(y: np.ndarray, h: int, w: int) -> np.ndarray:
    """(h, w, 3) radiance from the (3, H, W) network output."""
    return np.maximum(np.expm1(y[:, :h, :w].transpose(1, 2, 0)), 0).astype(np.float32)


def reference_denoise(model, radiance, aov, radiance_scale: float = 1.0, depth_scale: float = 1.0,
                      bf16: bool = True) -> np.ndarray:
    """The network on the CPU in FP32 (BF16 rounding at the library's points
    when bf16 is set)."""
    import torch

    h, w = radiance.shape[:2]
    x = pack_input(radiance, aov, radiance_scale, depth_scale, model.levels)
    with torch.no_grad():
        y = model.cpu().float()(torch.from_numpy(x)[None], bf16=bf16)[0].numpy()
    return unpack_output(y, h, w)


# ------------------------------------------------------------ weights file --
def save_weights(path, flat: np.ndarray, levels: int, channels, depth_scale: float, meta: str = "") -> None:
    """Weights as BF16 bit patterns (the precision the library runs at) with
    the network shape."""
    np.savez(path, weights_bf16=bf16_bits(flat), levels=np.int32(levels),
             channels=np.asarray(list(channels)[:levels], np.int32), depth_scale=np.float32(depth_scale),
             meta=np.asarray(meta))


def load_weights(path=DEFAULT_WEIGHTS):
    """(flat FP32 weights, levels, channels, depth_scale) of a weights file."""
    with np.load(path, allow_pickle=False) as z:
        flat = from_bf16_bits(z["weights_bf16"])
        levels = int(z["levels"])
        channels = tuple(int(c) for c in z["channels"])
        depth_scale = float(z["depth_scale"])
    if flat.size != weight_count(levels, channels):
        raise ValueError(f"{path}: {flat.size} weights, the network needs {weight_count(levels, channels)}")
    return flat, levels, channels, depth_scale
