"""GPU: presentation interop of the ray-cast pass: the overlay_web (sRGB) epilogue against the
oracle, the exported device pointers against the read-back, and linon_render captured by a
caller into a HIP graph (through PyTorch's graph API on a PyTorch stream) and replayed."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import linon_amd as L
from linon_amd import scenes
from oracle import oracle as O

from gpu_common import run_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def view():
    v = L.MainView(16, 16)
    yield v
    v.close()


def _startup(view, W, H, overlay_mode=1):
    v, f, c = scenes.reference_startup_scene()
    view.resize_texture(W, H)
    st = L.Settings.make(overlay_mode=overlay_mode, central_difference_delta=1, lyapunov_scaling=50.0)
    return v, f, L.Camera.arcball_default(c), st


def test_overlay_web_srgb_epilogue(view):
    """overlay_web (src/overlay.wgsl:156-165) against the oracle's, applied to the GPU's own
    ray-cast colour and mapping: the rgb channels sRGB-encoded, alpha unchanged."""
    v, f, cam, st = _startup(view, 80, 60)
    print("[case] overlay_web 80x60", flush=True)
    view.set_output_srgb(True)
    try:
        gpu = run_gpu(view, cam, st, L.Integrator.make(), "mirage_spherical_sigmoid", v, f)
    finally:
        view.set_output_srgb(False)
    ref = O.overlay(O.as_oracle_struct(st, O.SettingsIn), gpu.ray_casting, gpu.mapping, web=True)
    d = np.abs(gpu.final.astype(int) - ref["final"].astype(int))
    assert d.max() <= 1 and (d == 0).mean() >= 0.999, (d.max(), (d == 0).mean())
    desk = O.overlay(O.as_oracle_struct(st, O.SettingsIn), gpu.ray_casting, gpu.mapping)
    # the encoding brightens every channel strictly inside (0, 1)
    mid = (desk["final"][..., :3] > 0) & (desk["final"][..., :3] < 255)
    assert mid.any() and (gpu.final[..., :3][mid] >= desk["final"][..., :3][mid]).all()
    assert np.array_equal(gpu.final[..., 3], desk["final"][..., 3])
    assert view._lib.linon_set_output_srgb(view._h, 2) == 1  # LINON_ERR_ARGUMENT


def test_device_buffers_are_the_outputs(view):
    """linon_get_device_buffers: the exported pointers hold exactly what linon_read_* returns."""
    v, f, cam, st = _startup(view, 40, 24)
    print("[case] device buffers 40x24", flush=True)
    fr = run_gpu(view, cam, st, L.Integrator.make(), "mirage_spherical_sigmoid", v, f)
    d = view.device_buffers()
    assert (d.width, d.height) == (40, 24)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []

    def fetch(ptr, like):
        out = np.empty_like(like)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    assert np.array_equal(fetch(d.final_rgba8, fr.final), fr.final)
    assert np.array_equal(fetch(d.ray_casting_rgba8, fr.ray_casting), fr.ray_casting)
    assert np.array_equal(fetch(d.mapping_rgba32f, fr.mapping).view(np.uint32), fr.mapping.view(np.uint32))
    assert np.array_equal(fetch(d.hit_face_i32, fr.hit_faces), fr.hit_faces)
    assert np.array_equal(fetch(d.steps_i32, fr.steps), fr.steps)


def test_render_captured_into_a_graph(view):
    """A caller captures linon_render on its own stream into a HIP graph and replays it; with
    eager renders of another field between the replays, every replay leaves the frame the
    eager render of the captured field gives, bit for bit."""
    v, f, cam, st = _startup(view, 64, 48)
    print("[case] graph capture 64x48", flush=True)
    want = run_gpu(view, cam, st, L.Integrator.make(), "mirage_spherical_sigmoid", v, f)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        view.render(stream=s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        view.set_field("swirl")             # eager work between the replays
        view.render()
        other = view.read_frame()
        assert not np.array_equal(other.mapping, want.mapping)
        view.set_field("mirage_spherical_sigmoid")
        g.replay()
        torch.cuda.synchronize()
        got = view.read_frame()
        assert np.array_equal(got.mapping.view(np.uint32), want.mapping.view(np.uint32)), rep
        assert np.array_equal(got.final, want.final), rep
        assert np.array_equal(got.hit_faces, want.hit_faces), rep
    del g
