"""The launch-shape autotuner on the GPU: it searches both kernel shapes of a
BVH PathTracer frame and the wavefront's knobs of a RegularGrid frame, every
frame it renders equal to the default frame (it raises otherwise), and a
render through the tuned table gives the bitmap and accumulation of the
default AUTO render bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from helpers import soup_camera
from mobileraytracer_amd import abi, tuning

pytestmark = pytest.mark.gpu


def rnd(seed):
    return mrt.sampler(abi.SAMPLER_RANDOM, 0.0, seed)


@pytest.mark.parametrize("accel", [abi.BVH, abi.REGULAR_GRID])
def test_autotuner_picks_a_shape_and_keeps_pixels(gpu_ctx, accel):
    W, H, spp = 320, 176, 4
    data = mrt.soup_scene(21, 100000, 1000)
    cam = soup_camera(W, H)
    desc = mrt.render_desc(abi.SHADER_PATHTRACER, W, H, spp, cam, pixel=mrt.sampler(abi.SAMPLER_HALTON, 0, 5),
                           rr=rnd(6), light=rnd(7), hemi=rnd(8), depth_min=2, depth_max=6)
    scene = mrt.Scene(gpu_ctx, data)
    tuner = tuning.Autotuner(gpu_ctx, reps=1, passes=1)
    try:
        if accel == abi.REGULAR_GRID:
            cmin, cmax = mrt.camera_aabb(cam)
            scene.build(accel, cmin, cmax)
        else:
            scene.build(accel)
        acc0 = np.zeros((H, W, 3), np.float32)
        bm0, st0 = scene.render(desc, accum=acc0)
        before = gpu_ctx.tuning()
        res = tuner.tune(scene, desc)
        print(f"tuned {res.key}: {res.shape} {res.knobs} {res.ms:.2f} ms (default {res.default_ms:.2f} ms, "
              f"{len(res.trials)} trials)", flush=True)
        shapes = {t["shape"] for t in res.trials}
        if accel == abi.BVH:
            assert shapes == {tuning.MEGAKERNEL, tuning.WAVEFRONT}
        else:
            assert shapes == {tuning.PER_PASS, tuning.WAVEFRONT}
        # The chosen configuration is within the noise threshold of the
        # fastest trial of its shape, and no slower than that shape's
        # defaults (the search's starting point).
        own = [t["ms"] for t in res.trials if t["shape"] == res.shape]
        assert res.ms * (1.0 - tuner.min_gain) <= min(own) + 1e-3
        assert res.ms <= own[0] + 1e-3
        assert gpu_ctx.tuning() == before, "the tuner restores the context's knobs"
        acc1 = np.zeros((H, W, 3), np.float32)
        bm1, st1 = tuner.render(scene, desc, accum=acc1)
        assert (st1.rays_camera, st1.rays_extension, st1.rays_shadow) == \
            (st0.rays_camera, st0.rays_extension, st0.rays_shadow)
        assert np.array_equal(bm1, bm0)
        assert np.array_equal(acc1.view(np.uint32), acc0.view(np.uint32))
    finally:
        gpu_ctx.set_tuning()
        scene.close()
