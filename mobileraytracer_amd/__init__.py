"""MI355X (gfx950) ray-scene path of MobileRT behind a C ABI.

Python-side conveniences over libmrt_gpu.so (include/mobilert_gpu.h): scene
generation, upload, accelerator builds, traversal and the tile renderer.
Every compute call runs HIP kernels; when the library or the GPU is missing
the calls raise instead of computing anything on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import (BVH, NAIVE, REGULAR_GRID, MEM_DEVICE, MEM_HOST, SHADER_PATHTRACER, SHADER_WHITTED,  # noqa: F401
                  check, lib, ptr)

__all__ = ["Context", "Scene", "cornell_scene", "soup_scene", "make_camera", "tile_values", "render_desc",
           "sampler", "abi"]


def sampler(kind: int = abi.SAMPLER_CONSTANT, value: float = 0.5, seed: int = 0, domain_size: int = 0) -> abi.Sampler:
    return abi.Sampler(kind, value, seed, domain_size)


def make_camera(kind, position, look_at, up, a, b) -> abi.Camera:
    """Camera ctor fields (perspective: fovs in degrees; orthographic: sizes)."""
    cam = abi.Camera()
    p, l, u = (np.asarray(v, np.float32) for v in (position, look_at, up))
    check(lib().mrt_camera_make(kind, p.ctypes.data, l.ctypes.data, u.ctypes.data, a, b, C.byref(cam)),
          "mrt_camera_make")
    return cam


def camera_aabb(cam: abi.Camera) -> tuple[np.ndarray, np.ndarray]:
    """Camera::getAABB (Camera.cpp:39-45, Orthographic.cpp:33-45): (min, max)."""
    out = np.zeros(6, np.float32)
    check(lib().mrt_camera_aabb(C.byref(cam), out.ctypes.data), "mrt_camera_aabb")
    return out[:3].copy(), out[3:].copy()


def tile_values(seed: int = 0) -> np.ndarray:
    out = np.zeros(256, np.float32)
    check(lib().mrt_tile_values(seed, out.ctypes.data), "mrt_tile_values")
    return out


def render_desc(shader: int, width: int, height: int, spp: int, camera: abi.Camera, *, samples_light: int = 1,
                depth_min: int = 4, depth_max: int = 8, pixel=None, rr=None, light=None, hemi=None,
                tile_seed: int = 0, kernel: int = abi.KERNEL_MEGAKERNEL, first_sample: int = 0,
                instrument: bool = False, ray_sort: int = 0, persistent: int = 0,
                depth_max_point=(1.0, 1.0, 1.0), tiles=None) -> abi.RenderDesc:
    """Render descriptor.  tiles: None renders every tile of the 16x16
    grid; else an iterable of tile ids (row * 16 + column) to render, the
    others are skipped (tile_skip)."""
    d = abi.RenderDesc()
    d.shader = shader
    d.width, d.height, d.samples_pixel, d.samples_light = width, height, spp, samples_light
    d.ray_depth_min, d.ray_depth_max = depth_min, depth_max
    d.camera = camera
    d.sampler_pixel = pixel or sampler()
    d.sampler_rr = rr or sampler()
    d.sampler_light = light or sampler()
    d.sampler_hemi = hemi or sampler()
    vals = tile_values(tile_seed)
    C.memmove(d.tile_values, vals.ctypes.data, 256 * 4)
    d.kernel = kernel
    d.first_sample = first_sample
    d.instrument = 1 if instrument else 0
    d.ray_sort = ray_sort
    d.persistent = persistent
    d.depth_max_point[:] = [float(x) for x in depth_max_point]
    if tiles is not None:
        keep = set(int(t) for t in tiles)
        for k in range(8):
            d.tile_skip[k] = sum(1 << b for b in range(32) if 32 * k + b not in keep)
    return d


class SceneData:
    """Host arrays of a scene in the C ABI layouts."""

    def __init__(self, triangles=None, materials=None, lights=None, spheres=None, planes=None):
        self.triangles = triangles if triangles is not None else np.zeros(0, abi.TRIANGLE_DTYPE)
        self.materials = materials if materials is not None else np.zeros(0, abi.MATERIAL_DTYPE)
        self.lights = lights if lights is not None else np.zeros(0, abi.LIGHT_DTYPE)
        self.spheres = spheres if spheres is not None else np.zeros(0, abi.SPHERE_DTYPE)
        self.planes = planes if planes is not None else np.zeros(0, abi.PLANE_DTYPE)

    def desc(self) -> abi.SceneDesc:
        d = abi.SceneDesc()

        def put(arr, name):
            setattr(d, name, arr.ctypes.data if len(arr) else None)
            setattr(d, "n_" + name, len(arr))
        put(self.triangles, "triangles")
        put(self.spheres, "spheres")
        put(self.planes, "planes")
        put(self.materials, "materials")
        put(self.lights, "lights")
        return d


def cornell_scene(seed: int = 1) -> SceneData:
    tris = np.zeros(abi.CORNELL_TRIANGLES, abi.TRIANGLE_DTYPE)
    mats = np.zeros(abi.CORNELL_MATERIALS, abi.MATERIAL_DTYPE)
    lights = np.zeros(abi.CORNELL_LIGHTS, abi.LIGHT_DTYPE)
    check(lib().mrt_scene_cornell(seed, tris.ctypes.data, mats.ctypes.data, lights.ctypes.data), "mrt_scene_cornell")
    return SceneData(tris, mats, lights)


def soup_scene(seed: int, n: int, n_emissive: int, le: float = 5.0, half_extent: float = 0.005,
               light_sampler: int = abi.SAMPLER_RANDOM) -> SceneData:
    tris = np.zeros(n - n_emissive, abi.TRIANGLE_DTYPE)
    lights = np.zeros(n_emissive, abi.LIGHT_DTYPE)
    check(lib().mrt_scene_soup(seed, n, n_emissive, le, half_extent, tris.ctypes.data,
                               lights.ctypes.data if n_emissive else None, light_sampler), "mrt_scene_soup")
    return SceneData(tris, None, lights)


def icosphere_scene(seed: int, n_spheres: int, subdiv: int, n_emissive: int, displace: float = 0.3, le: float = 5.0,
                    light_sampler: int = abi.SAMPLER_RANDOM) -> SceneData:
    n = n_spheres * 20 * 4 ** subdiv
    tris = np.zeros(n - n_emissive, abi.TRIANGLE_DTYPE)
    lights = np.zeros(n_emissive, abi.LIGHT_DTYPE)
    check(lib().mrt_scene_icospheres(seed, n_spheres, subdiv, displace, n_emissive, le, tris.ctypes.data,
                                     lights.ctypes.data if n_emissive else None, light_sampler),
          "mrt_scene_icospheres")
    return SceneData(tris, None, lights)


def obj_scene(obj_path: str, mtl_path: str | None = None, light_sampler: abi.Sampler | None = None) -> SceneData:
    """OBJ / MTL files as a scene (OBJLoader::process + fillScene)."""
    h = C.c_void_p()
    check(lib().mrt_obj_load(str(obj_path).encode(), str(mtl_path).encode() if mtl_path else None, C.byref(h)),
          "mrt_obj_load")
    try:
        nt, nm, nl, ns = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().mrt_obj_counts(h, C.byref(nt), C.byref(nm), C.byref(nl), C.byref(ns)), "mrt_obj_counts")
        tris = np.zeros(nt.value, abi.TRIANGLE_DTYPE)
        mats = np.zeros(nm.value, abi.MATERIAL_DTYPE)
        lights = np.zeros(nl.value, abi.LIGHT_DTYPE)
        check(lib().mrt_obj_fill(h, tris.ctypes.data if nt.value else None, mats.ctypes.data if nm.value else None,
                                 lights.ctypes.data if nl.value else None,
                                 C.byref(light_sampler) if light_sampler is not None else None), "mrt_obj_fill")
    finally:
        lib().mrt_obj_destroy(h)
    data = SceneData(tris, mats, lights)
    data.faces_skipped = ns.value
    return data


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().mrt_ctx_create(device, C.byref(h)), "mrt_ctx_create")
        self.h = h

    def close(self):
        # The library refuses while scenes or servers of the context live
        # (the handle is kept, and freed by the last close that succeeds).
        if self.h and lib().mrt_ctx_destroy(self.h) == abi.MRT_OK:
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_budget(self, nbytes: int) -> None:
        check(lib().mrt_ctx_set_budget(self.h, nbytes), "mrt_ctx_set_budget")

    def set_tuning(self, **knobs) -> None:
        """Launch-shape knobs (mrt_tuning): refill_lanes, fetch_chunk,
        launch_share_pct, max_slot_lanes, max_groups, overlap_loads,
        graph_loop, service_lanes; 0 (or omitted) keeps a default."""
        t = abi.Tuning()
        for k, v in knobs.items():
            setattr(t, k, int(v))
        check(lib().mrt_ctx_set_tuning(self.h, C.byref(t)), "mrt_ctx_set_tuning")

    def tuning(self) -> dict:
        t = abi.Tuning()
        check(lib().mrt_ctx_get_tuning(self.h, C.byref(t)), "mrt_ctx_get_tuning")
        return {k: getattr(t, k) for k, _ in abi.Tuning._fields_ if k != "reserved"}

    def camera_rays(self, cam: abi.Camera, uvdd: np.ndarray) -> np.ndarray:
        uvdd = np.ascontiguousarray(uvdd, np.float32)
        out = np.zeros(len(uvdd), abi.RAY_DTYPE)
        check(lib().mrt_camera_rays(self.h, C.byref(cam), uvdd.ctypes.data, len(uvdd), out.ctypes.data, MEM_HOST),
              "mrt_camera_rays")
        return out

    def intersect_pairs(self, category: int, prims: np.ndarray, rays: np.ndarray):
        n = len(rays)
        hits = np.zeros(n, abi.HIT_DTYPE)
        anyh = np.zeros(n, np.uint8)
        check(lib().mrt_intersect_pairs(self.h, category, prims.ctypes.data, rays.ctypes.data, n, hits.ctypes.data,
                                        anyh.ctypes.data, MEM_HOST), "mrt_intersect_pairs")
        return hits, anyh


class Scene:
    def __init__(self, ctx: Context, data: SceneData):
        self.ctx = ctx
        self.data = data
        h = C.c_void_p()
        desc = data.desc()
        check(lib().mrt_scene_create(ctx.h, C.byref(desc), C.byref(h)), "mrt_scene_create")
        self.h = h

    def close(self):
        # Refused while a render server holds the scene.
        if self.h and lib().mrt_scene_destroy(self.h) == abi.MRT_OK:
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, accelerator: int, camera_min=(0, 0, 0), camera_max=(0, 0, 0), leaf_size: int = 2,
              builder: int = abi.BUILDER_AUTO) -> None:
        p = abi.AccelParams()
        p.camera_min[:] = list(camera_min)
        p.camera_max[:] = list(camera_max)
        p.leaf_size = leaf_size
        p.builder = builder
        check(lib().mrt_scene_build(self.h, accelerator, C.byref(p)), "mrt_scene_build")
        self.accel = accelerator

    def footprint(self) -> int:
        v = C.c_uint64()
        check(lib().mrt_scene_footprint(self.h, C.byref(v)), "mrt_scene_footprint")
        return v.value

    def trace_closest(self, rays, mem: int = MEM_HOST, out=None):
        n = len(rays) if mem == MEM_HOST else rays.shape[0]
        if out is None:
            out = np.zeros(n, abi.HIT_DTYPE)
        check(lib().mrt_trace_closest(self.h, ptr(rays), n, ptr(out), mem), "mrt_trace_closest")
        return out

    def trace_any(self, rays, mem: int = MEM_HOST, out=None):
        n = len(rays) if mem == MEM_HOST else rays.shape[0]
        if out is None:
            out = np.zeros(n, np.uint8)
        check(lib().mrt_trace_any(self.h, ptr(rays), n, ptr(out), mem), "mrt_trace_any")
        return out

    def shade_rays(self, desc: abi.RenderDesc, rays: np.ndarray, depths=None, sample: int = 0):
        n = len(rays)
        rgb = np.zeros((n, 3), np.float32)
        flag = np.zeros(n, np.uint8)
        dp = None if depths is None else np.ascontiguousarray(depths, np.int32)
        check(lib().mrt_shade_rays(self.h, C.byref(desc), rays.ctypes.data, ptr(dp), n, sample, rgb.ctypes.data,
                                   flag.ctypes.data, MEM_HOST), "mrt_shade_rays")
        return rgb, flag

    def render(self, desc: abi.RenderDesc, bitmap=None, accum=None, stride_bytes: int | None = None,
               mem: int = MEM_HOST):
        """Renderer::renderFrame.  Host numpy buffers by default; with
        mem=MEM_DEVICE `bitmap`/`accum` are device tensors."""
        if bitmap is None:
            bitmap = np.zeros((desc.height, desc.width), np.uint32)
        if stride_bytes is None:
            stride_bytes = desc.width * 4
        stats = abi.RenderStats()
        check(lib().mrt_render(self.h, C.byref(desc), ptr(bitmap), stride_bytes, ptr(accum), mem, C.byref(stats)),
              "mrt_render")
        return bitmap, stats

    def render_with_aovs(self, desc: abi.RenderDesc, bitmap=None, accum=None, aov=None,
                         stride_bytes: int | None = None, mem: int = MEM_HOST):
        """render() and aovs() of the same frame in one call; returns
        (bitmap, aov, stats)."""
        if bitmap is None:
            bitmap = np.zeros((desc.height, desc.width), np.uint32)
        if aov is None:
            aov = np.zeros((desc.height, desc.width, 7), np.float32)
        if stride_bytes is None:
            stride_bytes = desc.width * 4
        stats = abi.RenderStats()
        check(lib().mrt_render_with_aovs(self.h, C.byref(desc), ptr(bitmap), stride_bytes, ptr(accum), ptr(aov), mem,
                                         C.byref(stats)), "mrt_render_with_aovs")
        return bitmap, aov, stats

    def aovs(self, desc: abi.RenderDesc, out=None, mem: int = MEM_HOST):
        """First-hit albedo, normal and depth per pixel (height x width x 7)."""
        if out is None:
            out = np.zeros((desc.height, desc.width, 7), np.float32)
        check(lib().mrt_render_aovs(self.h, C.byref(desc), ptr(out), mem), "mrt_render_aovs")
        return out

    def stop(self) -> None:
        check(lib().mrt_render_stop(self.h), "mrt_render_stop")

    def progress(self) -> int:
        v = C.c_uint32()
        check(lib().mrt_render_progress(self.h, C.byref(v)), "mrt_render_progress")
        return v.value


class RenderServer:
    """Continuous-batching render server over resident scenes (mrt_server_*).

    submit() queues a render and returns a job id; wait() blocks until its
    bitmap (and accumulation) are in the caller's buffers.  idle_direct=True
    lets an idle server render a lone small job through mrt_render's
    one-launch megakernel instead of its pipelines."""

    def __init__(self, ctx: Context, max_slots: int = 0, max_jobs: int = 0, max_scenes: int = 0,
                 pipelines: int = 0, job_memory: int = 0, scene_memory: int = 0, sample_memory: int = 0,
                 idle_direct: bool = False):
        self.ctx = ctx
        d = abi.ServerDesc()
        d.max_slots, d.max_jobs, d.max_scenes, d.pipelines = max_slots, max_jobs, max_scenes, pipelines
        d.job_memory, d.scene_memory, d.sample_memory = job_memory, scene_memory, sample_memory
        d.idle_direct = 1 if idle_direct else 0
        h = C.c_void_p()
        check(lib().mrt_server_create(ctx.h, C.byref(d), C.byref(h)), "mrt_server_create")
        self.h = h
        self.scenes = []
        self._buffers = {}

    def close(self):
        if self.h:
            lib().mrt_server_destroy(self.h)
            self.h = None
            self._buffers.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: Scene) -> int:
        v = C.c_uint32()
        check(lib().mrt_server_add_scene(self.h, scene.h, C.byref(v)), "mrt_server_add_scene")
        self.scenes.append(scene)
        return v.value

    def submit(self, scene_id: int, desc: abi.RenderDesc, bitmap=None, accum=None, stride_bytes: int | None = None,
               mem: int = MEM_HOST) -> int:
        if bitmap is None:
            bitmap = np.zeros((desc.height, desc.width), np.uint32)
        if stride_bytes is None:
            stride_bytes = desc.width * 4
        jid = C.c_uint64()
        check(lib().mrt_server_submit(self.h, scene_id, C.byref(desc), ptr(bitmap), stride_bytes, ptr(accum), mem,
                                      C.byref(jid)), "mrt_server_submit")
        self._buffers[jid.value] = (bitmap, accum)
        return jid.value

    def wait(self, job_id: int, timeout_us: int = -1):
        """Returns (bitmap, accum, JobInfo); raises MrtError on a failed job."""
        info = abi.JobInfo()
        rc = lib().mrt_server_wait(self.h, job_id, timeout_us, C.byref(info))
        if rc == abi.MRT_E_TIMEOUT:
            return None, None, info
        bufs = self._buffers.pop(job_id, (None, None))
        check(rc, "mrt_server_wait")
        return bufs[0], bufs[1], info

    def poll(self, job_id: int) -> abi.JobInfo:
        info = abi.JobInfo()
        check(lib().mrt_server_poll(self.h, job_id, C.byref(info)), "mrt_server_poll")
        return info

    def cancel(self, job_id: int) -> None:
        check(lib().mrt_server_cancel(self.h, job_id), "mrt_server_cancel")

    def stats(self) -> abi.ServingStats:
        st = abi.ServingStats()
        check(lib().mrt_server_stats(self.h, C.byref(st)), "mrt_server_stats")
        return st


# ---------------------------------------------------------------- denoiser --
def denoiser_desc(levels: int = 5, channels=(32, 64, 96, 128, 128), depth_scale: float = 1.0) -> abi.DenoiserDesc:
    d = abi.DenoiserDesc()
    d.levels = levels
    ch = list(channels)[:levels] + [0] * (5 - min(levels, len(channels)))
    d.channels[:] = ch[:5]
    d.depth_scale = depth_scale
    return d


def denoiser_weight_count(desc: abi.DenoiserDesc) -> int:
    v = C.c_uint64()
    check(lib().mrt_denoiser_weight_count(C.byref(desc), C.byref(v)), "mrt_denoiser_weight_count")
    return v.value


class Denoiser:
    """The U-Net denoiser of mrt_denoiser_* (BF16 on MFMA).  `weights` is the
    flat FP32 vector in the library's layer order ([cout][cin][3][3] then
    bias, per layer)."""

    def __init__(self, ctx: Context, weights: np.ndarray, desc: abi.DenoiserDesc | None = None):
        self.ctx = ctx
        self.desc = desc if desc is not None else denoiser_desc()
        w = np.ascontiguousarray(weights, np.float32).ravel()
        h = C.c_void_p()
        check(lib().mrt_denoiser_create(ctx.h, C.byref(self.desc), w.ctypes.data, w.size, C.byref(h)),
              "mrt_denoiser_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mrt_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, radiance, aov, radiance_scale: float = 1.0, out=None, mem: int = MEM_HOST, width=None,
            height=None):
        """Denoised radiance (height x width x 3) of a frame's radiance
        (height x width x 3, times radiance_scale) and AOVs (height x width x 7)."""
        if mem == MEM_HOST:
            radiance = np.ascontiguousarray(radiance, np.float32)
            aov = np.ascontiguousarray(aov, np.float32)
            height, width = radiance.shape[:2]
            if aov.shape[:2] != (height, width) or aov.shape[2] != 7 or radiance.shape[2] != 3:
                raise ValueError("radiance must be (h, w, 3) and aov (h, w, 7)")
            if out is None:
                out = np.zeros((height, width, 3), np.float32)
        check(lib().mrt_denoiser_run(self.h, ptr(radiance), radiance_scale, ptr(aov), width, height, ptr(out), mem),
              "mrt_denoiser_run")
        return out

    def last_ms(self) -> float:
        v = C.c_float()
        check(lib().mrt_denoiser_last_ms(self.h, C.byref(v)), "mrt_denoiser_last_ms")
        return v.value


def denoiser_train_desc(batch: int = 8, crop: int = 256, seed: int = 0, beta1: float = 0.9, beta2: float = 0.999,
                        eps: float = 1e-8) -> abi.DenoiserTrainDesc:
    d = abi.DenoiserTrainDesc()
    d.batch, d.crop, d.seed = batch, crop, seed
    d.beta1, d.beta2, d.eps = beta1, beta2, eps
    return d


def splitmix64(z: int) -> int:
    """The splitmix64 finaliser of z + 0x9E3779B97F4A7C15 (the trainer's crop
    schedule)."""
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


class DenoiserTrainer:
    """mrt_denoiser_trainer_*: trains the U-Net on the GPU (BF16 MFMA forward,
    backward-data and weight-gradient kernels, FP32 master weights, Adam).
    Pairs: packed inputs (n, h, w, 10) and log-space targets (n, h, w, 3),
    FP32."""

    def __init__(self, ctx: Context, weights: np.ndarray, desc: abi.DenoiserDesc | None = None,
                 train: abi.DenoiserTrainDesc | None = None):
        self.ctx = ctx
        self.desc = desc if desc is not None else denoiser_desc()
        self.train = train if train is not None else denoiser_train_desc()
        w = np.ascontiguousarray(weights, np.float32).ravel()
        self.count = w.size
        h = C.c_void_p()
        check(lib().mrt_denoiser_trainer_create(ctx.h, C.byref(self.desc), w.ctypes.data, w.size,
                                                C.byref(self.train), C.byref(h)), "mrt_denoiser_trainer_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mrt_denoiser_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pairs(self, inputs: np.ndarray, targets: np.ndarray) -> None:
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        if x.ndim != 4 or x.shape[3] != 10 or t.shape != x.shape[:3] + (3,):
            raise ValueError("inputs must be (n, h, w, 10) and targets (n, h, w, 3)")
        n, h, w = x.shape[:3]
        check(lib().mrt_denoiser_trainer_set_pairs(self.h, x.ctypes.data, t.ctypes.data, n, w, h, MEM_HOST),
              "mrt_denoiser_trainer_set_pairs")

    def crops(self, step: int) -> np.ndarray:
        out = np.zeros((self.train.batch, 5), np.uint32)
        check(lib().mrt_denoiser_trainer_crops(self.h, step, out.ctypes.data), "mrt_denoiser_trainer_crops")
        return out

    def step(self, steps: int, lr: float) -> float:
        loss = C.c_double()
        check(lib().mrt_denoiser_trainer_step(self.h, steps, lr, C.byref(loss)), "mrt_denoiser_trainer_step")
        return loss.value

    def gradient(self, inputs: np.ndarray, targets: np.ndarray, lr: float = 0.0) -> tuple[float, np.ndarray]:
        """(loss, flat gradient) of one batch (b, crop, crop, 10) / (b, crop,
        crop, 3); an Adam step follows when lr > 0."""
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        b = x.shape[0]
        g = np.zeros(self.count, np.float32)
        loss = C.c_double()
        check(lib().mrt_denoiser_trainer_gradient(self.h, x.ctypes.data, t.ctypes.data, b, lr, C.byref(loss),
                                                  g.ctypes.data, g.size, MEM_HOST), "mrt_denoiser_trainer_gradient")
        return loss.value, g

    def weights(self) -> np.ndarray:
        w = np.zeros(self.count, np.float32)
        check(lib().mrt_denoiser_trainer_weights(self.h, w.ctypes.data, w.size), "mrt_denoiser_trainer_weights")
        return w

    def stats(self) -> abi.DenoiserTrainStats:
        st = abi.DenoiserTrainStats()
        check(lib().mrt_denoiser_trainer_stats(self.h, C.byref(st)), "mrt_denoiser_trainer_stats")
        return st


def conv3x3_bf16(ctx: Context, x: np.ndarray, w: np.ndarray, bias: np.ndarray | None, relu: bool) -> np.ndarray:
    """One 3x3 convolution layer on the GPU: x (h, w, cin) and w (3, 3, cout,
    cin) as BF16 bit patterns (uint16), bias FP32; returns (h, w, cout) BF16
    bits."""
    x = np.ascontiguousarray(x, np.uint16)
    w = np.ascontiguousarray(w, np.uint16)
    h, wd, cin = x.shape
    cout = w.shape[2]
    if w.shape != (3, 3, cout, cin):
        raise ValueError("weights must be (3, 3, cout, cin)")
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    y = np.zeros((h, wd, cout), np.uint16)
    check(lib().mrt_conv3x3_bf16(ctx.h, x.ctypes.data, h, wd, cin, w.ctypes.data, ptr(b), cout, int(relu),
                                 y.ctypes.data, MEM_HOST), "mrt_conv3x3_bf16")
    return y
