"""GPU: the device build of the uniform grid (count / scan / fill / per-cell sort / distance
field kernels) against the host build, array for array, and the frames each grid renders."""
from __future__ import annotations

import numpy as np
import pytest

import linon_amd as L
from linon_amd import configs, scenes

pytestmark = pytest.mark.gpu


def _grid(view, builder):
    view.set_grid_builder(builder)
    return view.read_grid()


# ico8 is the 4K stress scene: the Cornell box plus the level-8 icosphere (1,310,756 faces) at
# the resolution the context picks for it
@pytest.mark.parametrize("mesh,res", [("cornell", 0), ("ico3", 0), ("ico6", 128), ("ico6", 256), ("ico8", 0),
                                      ("empty", 0)])
def test_device_grid_equals_host_grid(mesh, res):
    if mesh == "cornell":
        v, f = scenes.cornell_box_centered()
    elif mesh == "empty":
        v, f = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    else:
        v, f = scenes.benchmark_scene(level=int(mesh[3:]))
    print(f"[case] grid build {mesh} ({len(f)} faces) res={res}", flush=True)
    view = L.MainView(64, 48)
    try:
        view.set_accel(L.ACCEL_GRID, res)
        view.set_grid_builder(L.GRID_BUILD_HOST)
        view.update_model(v, f)
        h = view.read_grid()
        d = _grid(view, L.GRID_BUILD_DEVICE)
        assert np.array_equal(h["dims"], d["dims"])
        assert np.array_equal(h["cell_start"], d["cell_start"])
        assert np.array_equal(h["cell_faces"], d["cell_faces"])
        assert np.array_equal(h["dist"], d["dist"])
        assert np.array_equal(h["tris"], d["tris"])
    finally:
        view.close()


def test_device_grid_renders_the_same_frame():
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=4)
    view = L.MainView(96, 64)
    try:
        view.update_camera(wl.camera())
        view.update_settings(wl.settings())
        view.set_integrator(wl.integrator())
        view.set_field("lens")
        frames = []
        for builder in (L.GRID_BUILD_HOST, L.GRID_BUILD_DEVICE):
            print(f"[case] grid render builder={builder}", flush=True)
            view.set_grid_builder(builder)
            view.update_model(v, f)
            view.render()
            frames.append(view.read_frame())
        a, b = frames
        assert (a.hit_faces >= 0).mean() > 0.05
        assert np.array_equal(a.hit_faces, b.hit_faces)
        assert np.array_equal(a.mapping, b.mapping)
        assert np.array_equal(a.final, b.final)
    finally:
        view.close()
