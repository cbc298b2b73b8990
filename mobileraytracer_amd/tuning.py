"""Autotuner of the render launch shape.

The reference renders every frame the same way (Renderer::renderFrame's
tile loop, Renderer.cpp:67-116); this library has two GPU shapes for the
PathTracer frames that loop covers, and launch knobs for each, none of which
changes a pixel:

  * the persistent megakernel (one launch per frame, paths in registers, a
    wave shades once `service_lanes` of its lanes finished their walks);
  * the persistent wavefront (trace / regen / shade launches over up to
    `max_groups` slot groups on their own streams, each launch holding
    `launch_share_pct` of the chip, optionally replayed as a HIP graph);

both with `max_slot_lanes` sample lanes per pixel, `refill_lanes` and
`fetch_chunk` for the refilling walk.  Which is fastest depends on the scene
size, the frame size and the samples per pixel (the megakernel wins small
frames by not iterating; wide launches amortise better on big ones), so the
tuner measures: per workload class (accelerator, primitive count, pixel count
and samples per pixel, each bucketed to a power of two) it runs a coordinate
descent over the knobs of each shape from the library defaults, keeps a
candidate only when it is faster by `min_gain`, and checks that every frame it
renders is bit-identical to the default frame.  The leaf size is not a knob:
the BVH keeps the reference's leaf size (BVH.hpp:19), so traversal order and
results stay the reference's.

Frames the wavefront cannot run (Whitted, several light samples, mirrors or
glass) and RegularGrid megakernel frames run one megakernel launch per sample
pass, which has no knobs: the tuner records them as such without a search.
"""
from __future__ import annotations

import json
import math
import statistics
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable

import numpy as np

from . import abi

MEGAKERNEL = "megakernel"
WAVEFRONT = "wavefront"
PER_PASS = "per_pass"

# Candidate values per knob and kernel shape (the library default first).
SEARCH_SPACE: dict[str, dict[str, tuple[int, ...]]] = {
    MEGAKERNEL: {
        "max_slot_lanes": (8, 4, 16),
        "service_lanes": (16, 8, 24, 32),
        "refill_lanes": (32, 16, 48),
        "fetch_chunk": (64, 128, 256),
    },
    WAVEFRONT: {
        "max_slot_lanes": (8, 4, 16),
        "max_groups": (3, 1, 2),
        "launch_share_pct": (0, 30, 60, 100),
        "refill_lanes": (12, 8, 16, 32),
        "fetch_chunk": (64, 128, 256),
        "overlap_loads": (0, 1, 2),
        "graph_loop": (0, 1),
    },
}

KNOBS = tuple(k for k, _ in abi.Tuning._fields_ if k != "reserved")


def _bucket(n: int) -> int:
    """Power-of-two class of n (0 for n <= 1)."""
    return 0 if n <= 1 else int(math.ceil(math.log2(n)))


def workload_key(accel: int, n_prims: int, width: int, height: int, spp: int) -> str:
    """Workload class: accelerator, then the power-of-two class of the
    primitive count, of the pixel count and of the samples per pixel (16 and
    more share a class: the launch shape stops changing there)."""
    return f"a{accel}-p{_bucket(n_prims)}-px{_bucket(width * height)}-s{_bucket(min(int(spp), 16))}"


def _cfg_id(cfg: dict) -> tuple:
    return tuple(sorted(cfg.items()))


def coordinate_descent(measure: Callable[[dict], float], space: dict[str, tuple], start: dict,
                       passes: int = 2, min_gain: float = 0.02) -> tuple[dict, float, list]:
    """Minimise measure(cfg) one knob at a time from `start`.  A value
    replaces the current one only when it is faster by more than `min_gain`
    (relative), so timing noise does not move the result; a configuration is
    measured once.  Returns (best configuration, its time, every trial as
    (configuration, time))."""
    seen: dict[tuple, float] = {}
    trials: list[tuple[dict, float]] = []

    def timed(cfg: dict) -> float:
        k = _cfg_id(cfg)
        if k not in seen:
            seen[k] = float(measure(dict(cfg)))
            trials.append((dict(cfg), seen[k]))
        return seen[k]

    best = dict(start)
    t_best = timed(best)
    for _ in range(max(1, passes)):
        improved = False
        for knob, values in space.items():
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best)
                cand[knob] = v
                t = timed(cand)
                if t < t_best * (1.0 - min_gain):
                    best, t_best, improved = cand, t, True
        if not improved:
            break
    return best, t_best, trials


@dataclass
class TuneResult:
    """Outcome of one workload class's search."""
    key: str
    shape: str                      # MEGAKERNEL, WAVEFRONT or PER_PASS
    knobs: dict = field(default_factory=dict)
    ms: float = 0.0                 # the chosen configuration's frame time
    default_ms: float = 0.0         # the AUTO kernel with default knobs
    trials: list = field(default_factory=list)

    def to_json(self) -> dict:
        return {"key": self.key, "shape": self.shape, "knobs": self.knobs, "ms": self.ms,
                "default_ms": self.default_ms, "trials": self.trials}

    @staticmethod
    def from_json(d: dict) -> "TuneResult":
        return TuneResult(d["key"], d["shape"], dict(d.get("knobs", {})), float(d.get("ms", 0.0)),
                          float(d.get("default_ms", 0.0)), list(d.get("trials", [])))


def _kernel_of(shape: str) -> int:
    return abi.KERNEL_WAVEFRONT if shape == WAVEFRONT else abi.KERNEL_MEGAKERNEL


class Autotuner:
    """Per-workload launch shapes for one context.

    tune(scene, desc) searches and records the best shape of the desc's
    workload class; apply(scene, desc) sets the context's knobs (mrt_tuning)
    for that class and returns a copy of the desc with its kernel; render()
    does both and renders.  The table can be saved as JSON and loaded by
    another process on the same GPU model."""

    def __init__(self, ctx, table: dict[str, TuneResult] | None = None, reps: int = 2, passes: int = 1,
                 min_gain: float = 0.02, auto: bool = True):
        self.ctx = ctx
        self.table: dict[str, TuneResult] = dict(table or {})
        self.reps = reps
        self.passes = passes
        self.min_gain = min_gain
        self.auto = auto

    # -- workload classes ---------------------------------------------------
    @staticmethod
    def key(scene, desc: abi.RenderDesc) -> str:
        d = scene.data
        n = len(d.triangles) + len(d.spheres) + len(d.planes)
        accel = getattr(scene, "accel", abi.BVH)
        return workload_key(accel, n, int(desc.width), int(desc.height),
                            int(desc.samples_pixel) - int(desc.first_sample))

    # -- measurement ----------------------------------------------------------
    def _frame_buffers(self, desc):
        try:
            import torch

            if torch.cuda.is_available():
                dev = torch.device("cuda", 0)
                return (torch.zeros((desc.height, desc.width), dtype=torch.int32, device=dev), abi.MEM_DEVICE, torch)
        except ImportError:
            pass
        return np.zeros((desc.height, desc.width), np.uint32), abi.MEM_HOST, None

    def _render_as(self, scene, desc, shape: str, knobs: dict, bitmap, mem):
        self.ctx.set_tuning(**knobs)
        d = abi.RenderDesc.from_buffer_copy(desc)
        d.kernel = _kernel_of(shape)
        d.persistent = 2 if shape == PER_PASS else 0
        return scene.render(d, bitmap=bitmap, mem=mem)

    def _measure(self, scene, desc, shape, knobs, bitmap, mem, torch, ref) -> float:
        times = []
        for r in range(self.reps + 1):
            _, st = self._render_as(scene, desc, shape, knobs, bitmap, mem)
            if r == 0:
                img = bitmap.cpu().numpy() if torch is not None else bitmap.copy()
                if ref is not None and not np.array_equal(img.view(np.uint32), ref.view(np.uint32)):
                    raise RuntimeError(f"tuning changed the frame: {shape} {knobs}")
                continue
            times.append(st.ms_render)
        return statistics.median(times)

    # -- search ---------------------------------------------------------------
    def tune(self, scene, desc: abi.RenderDesc) -> TuneResult:
        key = self.key(scene, desc)
        saved = self.ctx.tuning()
        bitmap, mem, torch = self._frame_buffers(desc)
        try:
            self.ctx.set_tuning()
            d0 = abi.RenderDesc.from_buffer_copy(desc)
            d0.kernel = abi.KERNEL_AUTO
            d0.persistent = 0
            scene.render(d0, bitmap=bitmap, mem=mem)
            ref = bitmap.cpu().numpy().copy() if torch is not None else bitmap.copy()
            t_default = []
            for _ in range(self.reps):
                _, st = scene.render(d0, bitmap=bitmap, mem=mem)
                t_default.append(st.ms_render)
            default_ms = statistics.median(t_default)
            shapes = []
            # The wavefront runs only the frames it covers; the megakernel
            # is persistent only where the wavefront could run and the scene
            # has BVH walk trees.
            try:
                self._render_as(scene, desc, WAVEFRONT, {}, bitmap, mem)
                shapes.append(WAVEFRONT)
            except abi.MrtError as e:
                if e.code != abi.MRT_E_UNSUPPORTED:
                    raise
            _, st = self._render_as(scene, desc, MEGAKERNEL, {}, bitmap, mem)
            shapes.insert(0, MEGAKERNEL if st.persistent_megakernel else PER_PASS)
            best: TuneResult | None = None
            trials = []
            for shape in shapes:
                space = SEARCH_SPACE.get(shape, {})
                start = {k: v[0] for k, v in space.items()}
                cfg, ms, tr = coordinate_descent(
                    lambda c, s=shape: self._measure(scene, desc, s, c, bitmap, mem, torch, ref),
                    space, start, self.passes, self.min_gain)
                trials += [{"shape": shape, "knobs": c, "ms": round(t, 4)} for c, t in tr]
                if best is None or ms < best.ms * (1.0 - self.min_gain):
                    best = TuneResult(key, shape, cfg, ms)
            best.default_ms = default_ms
            best.trials = trials
            self.table[key] = best
            return best
        finally:
            self.ctx.set_tuning(**saved)

    # -- use ------------------------------------------------------------------
    def apply(self, scene, desc: abi.RenderDesc) -> abi.RenderDesc:
        key = self.key(scene, desc)
        res = self.table.get(key)
        if res is None:
            if not self.auto:
                self.ctx.set_tuning()
                return abi.RenderDesc.from_buffer_copy(desc)
            res = self.tune(scene, desc)
        self.ctx.set_tuning(**res.knobs)
        d = abi.RenderDesc.from_buffer_copy(desc)
        d.kernel = _kernel_of(res.shape)
        d.persistent = 2 if res.shape == PER_PASS else 0
        return d

    def render(self, scene, desc: abi.RenderDesc, **kw):
        return scene.render(self.apply(scene, desc), **kw)

    # -- persistence ----------------------------------------------------------
    def save(self, path) -> None:
        Path(path).write_text(json.dumps({k: r.to_json() for k, r in self.table.items()}, indent=1))

    def load(self, path) -> None:
        data = json.loads(Path(path).read_text())
        for k, r in data.items():
            res = TuneResult.from_json(r)
            if res.shape not in (MEGAKERNEL, WAVEFRONT, PER_PASS) or any(n not in KNOBS for n in res.knobs):
                raise ValueError(f"bad tuning record {k}")
            self.table[k] = res
