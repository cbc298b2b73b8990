"""Shared test helpers: scene builders, ray batches and comparisons."""
from __future__ import annotations

import ctypes as C

import numpy as np

import mobileraytracer_amd as mrt
from mobileraytracer_amd import abi


def f32(v) -> np.float32:
    return np.float32(float(v)) if not isinstance(v, str) else np.float32(v)


def vec(v) -> np.ndarray:
    return np.array([f32(x) for x in v], np.float32)


def product_triangle(a, b, c, mat=-1, normals=None, colors=None) -> np.ndarray:
    out = np.zeros(1, abi.TRIANGLE_DTYPE)
    a, b, c = vec(a), vec(b), vec(c)
    n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(9)
    col = None if colors is None else np.ascontiguousarray(colors, np.float32).reshape(9)
    abi.check(abi.lib().mrt_triangle_make(a.ctypes.data, b.ctypes.data, c.ctypes.data, mat,
                                          None if n is None else n.ctypes.data,
                                          None if col is None else col.ctypes.data, out.ctypes.data),
              "mrt_triangle_make")
    return out


def make_rays(orig, dirs, tmax=1.0e30, exclude=abi.ID_NONE) -> np.ndarray:
    orig = np.asarray(orig, np.float32).reshape(-1, 3)
    dirs = np.asarray(dirs, np.float32).reshape(-1, 3)
    n = max(len(orig), len(dirs))
    r = np.zeros(n, abi.RAY_DTYPE)
    r["org"] = np.broadcast_to(orig, (n, 3))
    r["dir"] = np.broadcast_to(dirs, (n, 3))
    r["tmax"] = tmax
    r["exclude"] = exclude
    return r


def accel_params(cam_min=(0, 0, 0), cam_max=(0, 0, 0), leaf=2) -> abi.AccelParams:
    p = abi.AccelParams()
    p.camera_min[:] = [float(x) for x in cam_min]
    p.camera_max[:] = [float(x) for x in cam_max]
    p.leaf_size = leaf
    return p


def cornell_camera(w=256, h=256):
    ratio = np.float32(w) / np.float32(h)
    hf = max(float(ratio), 1.0)
    vf = max(float(np.float32(1) / ratio), 1.0)
    return mrt.make_camera(abi.CAMERA_PERSPECTIVE, (0, 0, -3.4), (0, 0, 1), (0, 1, 0),
                           float(np.float32(45.0) * np.float32(hf)), float(np.float32(45.0) * np.float32(vf)))


def soup_camera(w=1920, h=1080):
    ratio = np.float32(w) / np.float32(h)
    hf = max(float(ratio), 1.0)
    vf = max(float(np.float32(1) / ratio), 1.0)
    return mrt.make_camera(abi.CAMERA_PERSPECTIVE, (0.5, 0.5, -1.0), (0.5, 0.5, 0.5), (0, 1, 0),
                           float(np.float32(45.0) * np.float32(hf)), float(np.float32(45.0) * np.float32(vf)))


def mirror_scene() -> mrt.SceneData:
    """Box of axis-aligned planes with a mirror sphere, a glass sphere, a
    diffuse triangle and two area lights: exercises planes, spheres,
    reflection and transmission."""
    mats = np.zeros(6, abi.MATERIAL_DTYPE)
    mats["ior"] = 1.0
    mats[0]["kd"] = (0.7, 0.7, 0.7)
    mats[1]["kd"] = (0.8, 0.15, 0.1)
    mats[2]["kd"] = (0.1, 0.2, 0.8)
    mats[3]["ks"] = (0.9, 0.9, 0.9)                      # mirror
    mats[4]["kt"] = (0.9, 0.9, 0.9); mats[4]["ior"] = 1.9  # glass
    mats[5]["kd"] = (0.9, 0.9, 0.1); mats[5]["ks"] = (0.3, 0.3, 0.3)
    planes = np.zeros(6, abi.PLANE_DTYPE)
    spec = [((0, 0, 1), (0, 0, -1), 0), ((0, 0, -4), (0, 0, 1), 0), ((0, -1, 0), (0, 1, 0), 0),
            ((0, 1, 0), (0, -1, 0), 0), ((-1, 0, 0), (1, 0, 0), 1), ((1, 0, 0), (-1, 0, 0), 2)]
    for i, (p, n, m) in enumerate(spec):
        planes[i]["point"] = p
        planes[i]["normal"] = n
        planes[i]["material_id"] = m
    spheres = np.zeros(2, abi.SPHERE_DTYPE)
    spheres[0]["center"] = (0.45, -0.65, 0.4); spheres[0]["sq_radius"] = np.float32(0.35) * np.float32(0.35)
    spheres[0]["material_id"] = 3
    spheres[1]["center"] = (-0.4, -0.3, 0.0); spheres[1]["sq_radius"] = np.float32(0.35) * np.float32(0.35)
    spheres[1]["material_id"] = 4
    tris = np.concatenate([product_triangle((0.5, -0.5, 0.99), (0.5, 0.5, 0.99), (-0.5, -0.5, 0.99), 5),
                           product_triangle((-0.6, 0.2, 0.7), (-0.2, 0.6, 0.7), (-0.9, 0.7, 0.7), -1,
                                            colors=[[0.2, 0.8, 0.3]] * 3)])
    lights = np.zeros(2, abi.LIGHT_DTYPE)
    for i, pts in enumerate([((-0.25, 0.99, -0.25), (0.25, 0.99, -0.25), (0.25, 0.99, 0.25)),
                             ((0.25, 0.99, 0.25), (-0.25, 0.99, 0.25), (-0.25, 0.99, -0.25))]):
        L = abi.Light()
        L.kind = abi.LIGHT_AREA
        L.radiance.le[:] = [0.9, 0.9, 0.9]
        L.radiance.ior = 1.0
        t = product_triangle(*pts, -1)
        C.memmove(C.addressof(L.triangle), t.ctypes.data, 112)
        L.sampler = mrt.sampler(abi.SAMPLER_RANDOM, 0.5, 11 + i)
        C.memmove(lights[i:i + 1].ctypes.data, C.addressof(L), C.sizeof(abi.Light))
    return mrt.SceneData(tris, mats, lights, spheres, planes)


def random_rays(n, seed, lo=-1.0, hi=1.0, tmax=1.0e30) -> np.ndarray:
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o, d.astype(np.float32), tmax)


def pixel_uvdd(w, h, jitter=None):
    """(u, v, du, dv) per pixel as Renderer::renderScene computes them."""
    ys, xs = np.mgrid[0:h, 0:w]
    inv_w = np.float32(1.0) / np.float32(w)
    inv_h = np.float32(1.0) / np.float32(h)
    u = xs.astype(np.float32) * inv_w
    v = ys.astype(np.float32) * inv_h
    out = np.zeros((h * w, 4), np.float32)
    out[:, 0] = u.reshape(-1)
    out[:, 1] = v.reshape(-1)
    return out


def oracle_scene(data: mrt.SceneData, accel: int, cam_min=(0, 0, 0), cam_max=(0, 0, 0), leaf=2):
    import oracle

    desc = data.desc()
    sc = oracle.OracleScene(desc)
    sc._keep = (data, desc)
    sc.build(accel, accel_params(cam_min, cam_max, leaf))
    return sc


class DeviceArray:
    """Device memory from the HIP runtime the library itself loaded (no
    torch: its wheel ships another HIP runtime, and two in one process do not
    share devices)."""

    _hip = None

    def __init__(self, nbytes: int):
        if DeviceArray._hip is None:
            abi.lib()  # loads the runtime the library links
            DeviceArray._hip = C.CDLL("libamdhip64.so.7")
        p = C.c_void_p()
        rc = DeviceArray._hip.hipMalloc(C.byref(p), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        self.ptr = p.value
        self.nbytes = nbytes

    def data_ptr(self) -> int:
        return self.ptr

    def upload(self, a: np.ndarray) -> None:
        assert a.nbytes == self.nbytes
        rc = DeviceArray._hip.hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.nbytes == self.nbytes
        rc = DeviceArray._hip.hipMemcpy(C.c_void_p(a.ctypes.data), C.c_void_p(self.ptr), C.c_size_t(a.nbytes), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
        return a

    def free(self) -> None:
        if self.ptr:
            DeviceArray._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None
