"""GPU: the persistent multi-view batch kernel (work queue + wave-level ray compaction),
its hipGraph replay and the launch-shape autotuner."""
from __future__ import annotations

import numpy as np
import pytest

import linon_amd as L
from linon_amd import configs, scenes

from gpu_common import compare, run_oracles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    v, f = scenes.benchmark_scene(level=3)
    view = L.MainView(64, 64)
    view.update_model(v, f)
    view.set_integrator(configs.batch_integrator(max_steps=3000))
    views = configs.batch_views(num_views=9, seed=3, scale=16)
    yield view, views
    view.close()


def _plain(view, bv, max_steps=3000):
    view.resize_texture(bv.width, bv.height)
    view.update_camera(bv.camera)
    view.update_settings(L.Settings.make(field_weight=bv.field_weight))
    ig = configs.batch_integrator(max_steps=max_steps)
    ig.h_initial = bv.h_initial
    view.set_integrator(ig)
    view.set_field(bv.field)
    view.render()
    return view.read_frame()


def test_batch_views_equal_the_plain_kernel(setup):
    """Every view of a batch (mixed fields, step sizes and resolutions, escape mode) is bit for
    bit the frame the plain kernel renders for that view alone."""
    view, views = setup
    print(f"[case] batch of {len(views)} views", flush=True)
    view.set_batch(views)
    view.render_batch()
    c = view.counters()
    got = [view.read_batch_view(i) for i in range(len(views))]
    total_steps = 0
    for i, bv in enumerate(views):
        ref = _plain(view, bv)
        g = got[i]
        assert np.array_equal(g.hit_faces, ref.hit_faces), i
        assert np.array_equal(g.steps, ref.steps), i
        assert np.array_equal(g.mapping, ref.mapping), i
        assert np.array_equal(g.ray_casting, ref.ray_casting), i
        total_steps += int(g.steps.sum())
    assert c["rays"] == sum(v.width * v.height for v in views)
    assert c["rk4_steps"] == total_steps
    # escape mode: uneven path lengths
    st = np.concatenate([g.steps.ravel() for g in got])
    assert st.min() < st.max() / 4


def test_graph_replays_match_eager(setup):
    view, views = setup
    view.set_batch(views)
    view.render_batch()
    first = [view.read_batch_view(i) for i in range(len(views))]
    for rep in range(3):
        _plain(view, views[rep])          # eager work between replays
        view.render_batch()
        for i in range(len(views)):
            g = view.read_batch_view(i)
            assert np.array_equal(g.mapping, first[i].mapping), (rep, i)
            assert np.array_equal(g.hit_faces, first[i].hit_faces), (rep, i)


@pytest.mark.parametrize("tuning", [(4, 1, 1, 1), (8, 4, 256, 64), (32, 2, 16, 48)])
def test_launch_shapes_do_not_change_results(setup, tuning):
    view, views = setup
    view.set_batch(views)
    view.set_batch_tuning(*tuning)
    view.render_batch()
    a = [view.read_batch_view(i) for i in range(len(views))]
    view.set_batch_tuning(16, 1, 32, 16)
    view.render_batch()
    for i in range(len(views)):
        b = view.read_batch_view(i)
        assert np.array_equal(a[i].mapping, b.mapping) and np.array_equal(a[i].hit_faces, b.hit_faces)


def test_autotuner_picks_a_valid_shape(setup):
    view, views = setup
    view.set_batch(views)
    t, ms = view.autotune_batch(downscale=2)
    assert ms > 0
    assert t.waves_per_block in (1, 2, 4) and 1 <= t.refill_min <= 64 and t.compact_every >= 1
    assert view.batch_tuning().as_dict() == t.as_dict()
    view.render_batch()
    view.read_batch_view(0)


def test_batch_argument_errors(setup):
    view, views = setup
    bad = configs.batch_views(num_views=1, seed=0, scale=16)
    bad[0].field = 0
    with pytest.raises(L.LinonError):
        view.set_batch(bad)
    with pytest.raises(L.LinonError):
        view.set_batch_tuning(0, 1, 1, 1)


def test_full_resolution_batch_equals_plain():
    """Size-independent property at the benchmark's own sizes: four 1920x1080 views, one per
    field, different eyes and step sizes, in one batch equal the plain kernel's frames bit for
    bit."""
    v, f = scenes.benchmark_scene(level=5)
    view = L.MainView(1920, 1080)
    try:
        view.update_model(v, f)
        view.set_integrator(configs.batch_integrator(max_steps=4000))
        eyes = [(0.0, 0.0, 3.0), (2.0, 1.0, 2.2), (-2.5, 0.5, 1.5), (0.3, -2.8, 1.0)]
        steps = [0.005, 0.01, 0.0025, 0.005]
        views = [L.View.make(L.Camera.look_at(e, (0, 0, 0), fov_deg=60.0), 1920, 1080, name, h)
                 for e, name, h in zip(eyes, ("swirl", "lens", "hot_air", "warp"), steps)]
        print("[case] full-resolution batch of 4 views", flush=True)
        view.set_batch(views)
        view.render_batch()
        got = [view.read_batch_view(i) for i in range(len(views))]
        for i, bv in enumerate(views):
            ref = _plain(view, bv, max_steps=4000)
            assert np.array_equal(got[i].hit_faces, ref.hit_faces), i
            assert np.array_equal(got[i].mapping, ref.mapping), i
            assert np.array_equal(got[i].steps, ref.steps), i
    finally:
        view.close()


def test_mesh_change_recaptures_the_batch():
    """The captured batch graph holds the mesh and grid buffers by value, and linon_set_mesh
    after linon_set_batch frees them.  The batch is captured again on the new mesh: its frames
    equal those of a batch set up after the mesh change, bit for bit, and differ from the first
    mesh's."""
    v1, f1 = scenes.benchmark_scene(level=2)
    v2, f2 = scenes.benchmark_scene(level=3)
    views = configs.batch_views(num_views=4, seed=5, scale=16)
    print("[case] batch across a mesh change", flush=True)
    with L.MainView(32, 32) as view:
        view.set_integrator(configs.batch_integrator(max_steps=2000))
        view.update_model(v1, f1)
        view.set_batch(views)
        view.render_batch()
        first = [view.read_batch_view(i) for i in range(len(views))]
        view.update_model(v2, f2)
        view.render_batch()
        got = [view.read_batch_view(i) for i in range(len(views))]
        view.set_batch(views)
        view.render_batch()
        ref = [view.read_batch_view(i) for i in range(len(views))]
    for g, r in zip(got, ref):
        assert np.array_equal(g.hit_faces, r.hit_faces)
        assert np.array_equal(g.mapping, r.mapping)
    assert any(not np.array_equal(g.hit_faces, a.hit_faces) for g, a in zip(got, first))


def test_batch_follows_integrator_and_accel_changes(setup):
    """linon_set_integrator and linon_set_accel after linon_set_batch reach the captured batch:
    a smaller max_steps caps every view's iterations, and the brute-force loop renders the same
    frames as the grid."""
    view, views = setup
    view.set_integrator(configs.batch_integrator(max_steps=3000))
    view.set_batch(views)
    view.render_batch()
    full = [view.read_batch_view(i) for i in range(len(views))]
    assert max(int(g.steps.max()) for g in full) > 40
    print("[case] batch after set_integrator(max_steps=40)", flush=True)
    view.set_integrator(configs.batch_integrator(max_steps=40))
    view.render_batch()
    capped = [view.read_batch_view(i) for i in range(len(views))]
    for g, f in zip(capped, full):
        assert g.steps.max() <= 40
        assert np.array_equal(g.steps, np.minimum(f.steps, 40))
    print("[case] batch after set_accel(brute)", flush=True)
    view.set_accel(L.ACCEL_BRUTE)
    view.render_batch()
    c_brute = view.counters()
    brute = [view.read_batch_view(i) for i in range(len(views))]
    view.set_accel(L.ACCEL_AUTO)
    view.render_batch()
    c_grid = view.counters()
    for b, g in zip(brute, capped):
        assert np.array_equal(b.hit_faces, g.hit_faces) and np.array_equal(b.mapping, g.mapping)
    assert c_brute["triangle_tests"] > c_grid["triangle_tests"]
    view.set_integrator(configs.batch_integrator(max_steps=3000))


def test_batch_views_against_the_oracle():
    """Batch views (mixed fields, step sizes and resolutions; escape mode, first-hit stop)
    compared with the oracle's escape mode directly, not only with the plain kernel: hit face,
    end point and iteration count per view."""
    v, f = scenes.benchmark_scene(level=2)
    views = configs.batch_views(num_views=5, seed=19, scale=32, enhance=False)   # 32x32 and 60x33, 4 fields
    with L.MainView(32, 32) as view:
        view.update_model(v, f)
        view.set_integrator(configs.batch_integrator(max_steps=3000))
        print(f"[case] batch of {len(views)} views against the oracle", flush=True)
        view.set_batch(views)
        view.render_batch()
        got = [view.read_batch_view(i) for i in range(len(views))]
    fields = set()
    for i, bv in enumerate(views):
        ig = configs.batch_integrator(max_steps=3000)
        ig.h_initial = bv.h_initial
        ref, ref64 = run_oracles(bv.camera, L.Settings.make(field_weight=bv.field_weight), ig, bv.field, v, f,
                                 bv.width, bv.height)
        info = compare(got[i], ref, ref64, f"batch-view-{i}-field{bv.field}-h{bv.h_initial}-{bv.width}x{bv.height}")
        assert info["hit_ok"] and info["exit_pass"] and info["q99_ok"], info
        assert (got[i].steps == ref["steps"]).mean() >= 0.99
        fields.add(bv.field)
    assert len(fields) == 4


def test_the_65_view_benchmark_batch():
    """The benchmark's own batch (64 views drawn from seed 0 plus the Enhance view: resolutions
    up to 2048^2, the four fields, h in {0.01, 0.005, 0.0025} and h/4; escape mode, at most 8000
    RK4 steps; the 20,516-face scene) in one graph replay.  Its ray count is every view's
    pixels, and a sample of its views -- view 0 and its Enhance copy, the first view of each
    field and the largest view -- equals the plain kernel's frames bit for bit."""
    wl = configs.OVERLAY_1080P
    v, f = wl.mesh()
    views = configs.batch_views(num_views=64, seed=0)
    names = {val: k for k, val in L.FIELDS.items()}
    pick = {0, len(views) - 1, max(range(len(views)), key=lambda i: views[i].width * views[i].height)}
    for name in ("swirl", "lens", "hot_air", "warp"):
        pick.add(next(i for i, bv in enumerate(views) if names[bv.field] == name))
    with L.MainView(64, 64) as view:
        view.update_model(v, f)
        view.set_integrator(configs.batch_integrator(max_steps=8000))
        print(f"[case] 65-view benchmark batch, views {sorted(pick)} checked", flush=True)
        view.set_batch(views)
        view.render_batch()
        c = view.counters()
        assert c["rays"] == sum(bv.width * bv.height for bv in views)
        got = {i: view.read_batch_view(i) for i in sorted(pick)}
        for i in sorted(pick):
            ref = _plain(view, views[i], max_steps=8000)
            assert np.array_equal(got[i].hit_faces, ref.hit_faces), i
            assert np.array_equal(got[i].mapping.view(np.uint32), ref.mapping.view(np.uint32)), i
            assert np.array_equal(got[i].steps, ref.steps), i
