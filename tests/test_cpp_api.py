"""The reference-shaped C++ API (include/mobilert/*.hpp, libmobilert.so)
through its test program tests/cpp/api_check (built by build()).

Host checks run without a GPU: sampler counters (HaltonSeq.cpp:14-22,
Stratified.cpp:14-23, Sampler.cpp:9-13), the tile scheduler
(Camera.cpp:28-37), camera basis, construction, boxes and material
equality.  The GPU checks render the Cornell box through
Renderer::renderFrame and through RayTrace and compare the bitmaps with the
oracle bit for bit, and compare the per-ray methods (accelerator trace,
Triangle::intersect, shadowTrace, traceLights) with each other.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import mobileraytracer_amd as mrt
from mobileraytracer_amd import abi
from helpers import cornell_camera, oracle_scene

BIN = Path(__file__).resolve().parent / "cpp" / "api_check"
RANDOM = abi.SAMPLER_RANDOM
# Default light-choice and hemisphere streams of MobileRT::Shader.
LIGHT_SEED, HEMI_SEED = 0x6C696768, 0x68656D69


def run(*args: str) -> subprocess.CompletedProcess:
    if not BIN.exists():
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build()")
    print("api_check", *args, flush=True)
    return subprocess.run([str(BIN), *args], capture_output=True, text=True, timeout=300)


def test_cpp_api_host_checks():
    r = run("host")
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr


def cpp_desc():
    """The render descriptor the C++ Renderer builds for its Cornell frame."""
    cam = cornell_camera(64, 64)
    return mrt.render_desc(abi.SHADER_WHITTED, 64, 64, 1, cam, depth_min=4, depth_max=8,
                           light=mrt.sampler(RANDOM, 0.0, LIGHT_SEED), hemi=mrt.sampler(RANDOM, 0.0, HEMI_SEED))


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [abi.BVH, abi.REGULAR_GRID, abi.NAIVE])
def test_cpp_renderer_matches_oracle(tmp_path, accel):
    out = tmp_path / "bm.bin"
    r = run("render", str(out), str(accel))
    assert r.returncode == 0, r.stderr
    bm = np.fromfile(out, np.uint32).reshape(64, 64)
    desc = cpp_desc()
    cam_min, cam_max = mrt.camera_aabb(desc.camera)
    orc = oracle_scene(mrt.cornell_scene(1), accel, cam_min, cam_max)
    bo, cnt, _ = orc.render(desc, threads=8)
    assert int(cnt[0]) == 64 * 64
    assert np.array_equal(bm, bo)


@pytest.mark.gpu
def test_cpp_raytrace_entry_matches_renderer(tmp_path):
    a, b = tmp_path / "raytrace.bin", tmp_path / "renderer.bin"
    r1 = run("raytrace", str(a))
    assert r1.returncode == 0, r1.stderr
    r2 = run("render", str(b), str(abi.BVH))
    assert r2.returncode == 0, r2.stderr
    assert np.array_equal(np.fromfile(a, np.uint32), np.fromfile(b, np.uint32))


@pytest.mark.gpu
def test_cpp_per_ray_methods():
    r = run("rays")
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("shader", [abi.SHADER_NOSHADOWS, abi.SHADER_DEPTHMAP, abi.SHADER_DIFFUSE])
def test_cpp_one_hit_shaders(tmp_path, shader):
    """Components::NoShadows / DepthMap / DiffuseMaterial through the C++
    Renderer and through RayTrace (its shader ids 0, 3, 4) equal the oracle's
    frame."""
    a, b = tmp_path / "renderer.bin", tmp_path / "raytrace.bin"
    r1 = run("shader", str(shader), str(a))
    assert r1.returncode == 0, r1.stderr
    r2 = run("shader", str(shader), str(b), "rt")
    assert r2.returncode == 0, r2.stderr
    bm = np.fromfile(a, np.uint32).reshape(64, 64)
    assert np.array_equal(bm, np.fromfile(b, np.uint32).reshape(64, 64))
    desc = cpp_desc()
    desc.shader = shader
    desc.depth_max_point[:] = [1.0, 1.0, 1.0]
    if shader != abi.SHADER_NOSHADOWS:
        desc.samples_light = 0
    cam_min, cam_max = mrt.camera_aabb(desc.camera)
    orc = oracle_scene(mrt.cornell_scene(1), abi.BVH, cam_min, cam_max)
    bo, cnt, _ = orc.render(desc, threads=8)
    assert np.array_equal(bm, bo)


OBJ_DIR = Path(__file__).resolve().parent / "golden" / "obj"


def test_cpp_obj_loader():
    """Components::OBJLoader (OBJLoader.cpp:23-250): process() counts, the
    lambda called once per AreaLight, materials deduplicated against the
    scene's own (the scene's "mirror" is reused), a missing file fails."""
    r = run("obj", str(OBJ_DIR / "scene.obj"), str(OBJ_DIR / "scene.mtl"))
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr


@pytest.mark.gpu
def test_cpp_raytrace_obj(tmp_path):
    """RayTrace with an OBJ path renders the loaded scene (default camera)."""
    out = tmp_path / "obj.bin"
    r = run("objrt", str(OBJ_DIR / "scene.obj"), str(OBJ_DIR / "scene.mtl"), str(out))
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
