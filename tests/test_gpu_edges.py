"""GPU: edge cases of the ray-cast pass against the FP32 oracle: ragged and minimal image sizes
(partial 8x8 tiles), degenerate and duplicated triangles (the determinant rejection and the
lowest-index tie-break of src/main_view.wgsl:150-220), an empty scene, a ray-sample count that
does not divide the sample steps."""
from __future__ import annotations

import numpy as np
import pytest

import linon_amd as L
from linon_amd import configs, scenes
from oracle import oracle as O

from gpu_common import oracle_structs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def view():
    v = L.MainView(16, 16)
    yield v
    v.close()


def _render_both(view, W, H, cam, st, ig, field, verts, faces):
    view.resize_texture(W, H)
    view.update_model(verts, faces)
    view.update_camera(cam)
    view.update_settings(st)
    view.set_integrator(ig)
    view.set_field(field)
    view.render()
    g = view.read_frame()
    c, s, i = oracle_structs(cam, st, ig)
    r = O.render(c, s, i, L.FIELDS[field], verts, faces, W, H, threads=8)
    return g, r


@pytest.mark.parametrize("W,H", [(2, 2), (9, 7), (67, 33), (8, 70), (16384, 2), (2, 16384)])
def test_ragged_sizes_identity_field(view, W, H):
    """The identity field is a straight line: no field arithmetic, so hit faces must agree
    exactly and end points to rounding, on images whose edge tiles are partial, up to the
    largest width and height the context accepts (16384)."""
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    ig = L.Integrator.make(h_initial=0.01, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0)
    print(f"[case] ragged {W}x{H}", flush=True)
    view.set_math(L.MATH_IEEE)   # correctly rounded division and sqrt, as the oracle
    try:
        g, r = _render_both(view, W, H, wl.camera(), wl.settings(), ig, "identity", v, f)
    finally:
        view.set_math(L.MATH_AUTO)
    assert np.array_equal(g.hit_faces, r["hit"])
    assert np.array_equal(g.steps, r["steps"])
    d = np.abs(g.mapping[..., :3].astype(np.float64) - r["mapping"][..., :3]) / np.maximum(1.0, np.abs(r["mapping"][..., :3]))
    assert d.max() <= 1e-5, d.max()
    assert np.abs(g.ray_casting.astype(int) - r["color8"].astype(int)).max() <= 1


def test_degenerate_and_duplicate_triangles(view):
    """A zero-area triangle (rejected by |det| < 1e-7), a collinear one, and an exact duplicate
    of a valid triangle listed after it (a tie: the lower index wins) in front of the camera."""
    verts = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0],          # a valid triangle in z = 0
                      [0.2, 0.2, 0.5], [0.2, 0.2, 0.5], [0.2, 0.2, 0.5],   # zero area
                      [-0.5, 0, 0.3], [0, 0, 0.3], [0.5, 0, 0.3],   # collinear
                      ], np.float32)
    faces = np.array([[3, 4, 5], [6, 7, 8], [0, 1, 2], [0, 1, 2], [2, 1, 0]], np.uint32)
    cam = L.Camera.look_at((0.0, 0.0, 3.0), (0, 0, 0), fov_deg=60.0)
    ig = L.Integrator.make(h_initial=0.01, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0)
    print("[case] degenerate/duplicate triangles", flush=True)
    view.set_math(L.MATH_IEEE)
    try:
        g, r = _render_both(view, 48, 40, cam, L.Settings.make(), ig, "identity", verts, faces)
    finally:
        view.set_math(L.MATH_AUTO)
    assert np.array_equal(g.hit_faces, r["hit"])
    hit = r["hit"][r["hit"] >= 0]
    # never the degenerate faces 0 and 1, never 3 (an exact duplicate of 2, so an exact tie);
    # 4 lists the same corners in the other order and may round its t below face 2's
    assert hit.size > 100 and 2 in hit and not np.isin(hit, [0, 1, 3]).any()


def test_empty_scene(view):
    """No faces: every ray misses (hit -1, black with alpha 1) and ends where the integration
    ends; swirl bends the rays."""
    v, f = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    ig = L.Integrator.make(h_initial=0.01, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0)
    wl = configs.PLAIN_256_SWIRL
    print("[case] empty scene", flush=True)
    g, r = _render_both(view, 24, 16, wl.camera(), wl.settings(), ig, "swirl", v, f)
    assert (g.hit_faces == -1).all() and (r["hit"] == -1).all()
    assert (g.ray_casting[..., :3] == 0).all() and (g.ray_casting[..., 3] == 255).all()
    d = np.abs(g.mapping[..., :3].astype(np.float64) - r["mapping"][..., :3]) / np.maximum(1.0, np.abs(r["mapping"][..., :3]))
    assert np.quantile(d, 0.99) <= 1e-4


def test_sample_count_not_dividing_steps(view):
    """64 records over 5000 steps (stride 78, the stress configuration's count): records at
    i = 0, 78, ..., 4914, matching the oracle."""
    v, f, c = scenes.reference_startup_scene()
    cam = L.Camera.arcball_default(c)
    st = L.Settings.make(mouse_pos=(0.45, 0.5))
    ig = L.Integrator.make(samples_per_ray=64, sample_steps=5000)
    view.resize_texture(16, 16)
    view.update_model(v, f)
    view.update_camera(cam)
    view.update_settings(st)
    view.set_integrator(ig)
    view.set_field("mirage_spherical_sigmoid")
    print("[case] 64 sample records", flush=True)
    view.render()
    got = view.read_ray_samples()
    c2, s2, i2 = oracle_structs(cam, st, ig)
    ref = O.sample_rays(c2, s2, i2, 3, 16, 16)
    assert got.shape == ref.shape == (8, 64, 2, 4)
    np.testing.assert_array_equal(got[:, :, 1], ref[:, :, 1])
    err = np.abs(got[:, :, 0, :3] - ref[:, :, 0, :3]) / np.maximum(1.0, np.abs(ref[:, :, 0, :3]))
    assert err.max() <= 1e-3, err.max()
