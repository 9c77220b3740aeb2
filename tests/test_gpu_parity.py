"""GPU parity: the HIP ray-cast and overlay kernels (through the C ABI) against the FP32 CPU
oracle on the same camera, field, step size and mesh."""
from __future__ import annotations

import numpy as np
import pytest

import linon_amd as L
from linon_amd import configs, scenes
from oracle import oracle as O

from gpu_common import EXIT_RTOL, HIT_MIN_AGREE, compare, oracle_pixels, rel_err, run_gpu, run_oracle, run_oracles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def view():
    v = L.MainView(64, 48)
    yield v
    v.close()


def _case(view, W, H, cam, st, ig, field, verts, faces, label, accel=L.ACCEL_AUTO):
    print(f"[case] {label} {W}x{H} field={field} accel={accel}", flush=True)
    view.resize_texture(W, H)
    gpu = run_gpu(view, cam, st, ig, field, verts, faces, accel)
    fid = L.FIELDS[field] if isinstance(field, str) else field
    ref, ref64 = run_oracles(cam, st, ig, fid, verts, faces, W, H)
    info = compare(gpu, ref, ref64, label)
    assert info["low_lambda"] > 0, info
    assert info["hit_ok"], info
    assert info["exit_pass"], info
    assert info["q99_ok"], info
    assert info["color_maxdiff"] <= 1, info
    # rays whose adaptive steps were chosen by rounding noise still find the same faces
    assert info["noisy_hit_agree"] >= 0.95, info
    return gpu, ref


def test_reference_default_scene(view):
    """linon's start-up state: Cornell box (x, z negated, normalised), arcball default camera,
    MirageSphericalSigmoid, h 0.1 adaptive, t <= 5 (src/application.rs:105-176)."""
    v, f, c = scenes.reference_startup_scene()
    gpu, ref = _case(view, 64, 48, L.Camera.arcball_default(c), L.Settings.make(), L.Integrator.make(),
                     "mirage_spherical_sigmoid", v, f, "reference-default")
    assert (ref["hit"] >= 0).mean() > 0.5
    quiet = ref["noise_flag"] == 0
    assert np.array_equal(gpu.steps[quiet], ref["steps"][quiet])


@pytest.mark.parametrize("field", ["identity", "swirl", "lens", "hot_air", "warp"])
def test_config_camera_cornell(view, field):
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    _case(view, 64, 64, wl.camera(), wl.settings(), wl.integrator(), field, v, f, f"cornell-{field}")


@pytest.mark.parametrize("weight", [0.5, 0.0])
@pytest.mark.parametrize("field", ["swirl", "hot_air"])
def test_field_weight_blend(view, field, weight):
    """The blend cur_dir = (1 - w) cur_dir + w V (src/main_view.wgsl:255) for a field weight
    other than 1, which runs the step loop's general form (the measured configurations all have
    w = 1); w = 0 ignores the field and draws straight rays."""
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    st = L.Settings.make(field_weight=weight)
    gpu, ref = _case(view, 64, 64, wl.camera(), st, wl.integrator(), field, v, f, f"weight{weight}-{field}")
    if weight == 0.0:  # straight rays: the identity field's frame, bit for bit, in one build
        view.set_math(L.MATH_IEEE)  # (the default runs swirl and the identity in the WGSL build)
        try:
            a = run_gpu(view, wl.camera(), st, wl.integrator(), field, v, f)
            b = run_gpu(view, wl.camera(), st, wl.integrator(), "identity", v, f)
        finally:
            view.set_math(L.MATH_AUTO)
        assert np.array_equal(a.hit_faces, b.hit_faces)
        assert np.array_equal(a.mapping, b.mapping)


@pytest.mark.parametrize("math", ["ieee", "wgsl"])
@pytest.mark.parametrize("field", ["swirl", "lens", "warp"])
def test_math_builds(view, field, math):
    """Both kernel builds against the oracle on the 64x64 Cornell view: IEEE (correctly rounded
    division and sqrt, full-precision transcendentals) and WGSL accuracy (hardware division,
    sqrt, sin/cos/exp within WGSL's bounds; AUTO's choice for swirl and warp)."""
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    view.set_math({"ieee": L.MATH_IEEE, "wgsl": L.MATH_WGSL}[math])
    try:
        _case(view, 64, 64, wl.camera(), wl.settings(), wl.integrator(), field, v, f, f"cornell-{field}-{math}")
    finally:
        view.set_math(L.MATH_AUTO)


@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_benchmark_scene_small(view, field):
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=2)
    _case(view, 48, 32, wl.camera(), wl.settings(), wl.integrator(), field, v, f, f"bench-scene-{field}")


@pytest.mark.parametrize("field", ["lens", "hot_air"])
def test_benchmark_scene_small_exact_build(view, field):
    """The exact functors MATH_IEEE runs for lens and hot air (every quotient, root and
    exponential correctly rounded; the default math choice runs variants) against the oracle
    on the benchmark scene."""
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=2)
    view.set_math(L.MATH_IEEE)
    try:
        _case(view, 48, 32, wl.camera(), wl.settings(), wl.integrator(), field, v, f, f"bench-scene-ieee-{field}")
    finally:
        view.set_math(L.MATH_AUTO)


@pytest.mark.parametrize("clip_exit", [0, 1])
@pytest.mark.parametrize("need_mapping", [0, 1])
@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_escape_mode(view, field, need_mapping, clip_exit):
    """The escape mode that batches and the surrogate's RK4 legs integrate with (a ray stops when
    it leaves the radius-2 sphere moving outward, at max_steps, or at its first hit unless
    need_mapping; clip_exit ends it on the sphere) against the oracle's escape mode, on the
    benchmark scene: hit face, end point, and the iteration count."""
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=2)
    ig = L.Integrator.make(h_initial=0.005, t_max=1e30, adaptive=False, max_steps=4000, mode=L.MODE_ESCAPE,
                           escape_radius=2.0, samples_per_ray=0, sample_steps=0, need_mapping=bool(need_mapping),
                           clip_exit=bool(clip_exit))
    gpu, ref = _case(view, 48, 32, wl.camera(), wl.settings(), ig, field, v, f,
                     f"escape-{field}-map{need_mapping}-clip{clip_exit}")
    # the iteration count is decided by the escape test |p| > 2 on the f32 path; rays that end
    # within rounding of the sphere may take one step more or less
    assert (gpu.steps == ref["steps"]).mean() >= 0.99
    hit = ref["hit"] >= 0
    if not need_mapping:   # a hit ends the ray: its iteration count is the hit's step
        assert hit.any()
    if clip_exit:  # rays that crossed the sphere from inside (no hit, or mapping kept) end on it
        left = (ref["steps"] < 4000) & (~hit | bool(need_mapping))
        on = np.abs(np.linalg.norm(gpu.mapping[..., :3].astype(np.float64), axis=-1) - 2.0) <= 1e-5
        on_ref = np.abs(np.linalg.norm(ref["mapping"][..., :3].astype(np.float64), axis=-1) - 2.0) <= 1e-5
        # rays that never entered the sphere (swirl turns most camera rays away from it) end
        # outside it; lens captures many rays near its singularity
        if left.any():
            assert (on == on_ref)[left].mean() >= 0.99
        assert on.any() or field == "lens"


@pytest.mark.parametrize("fid", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_predefined_fields(view, fid):
    """Every body of PredefinedFunction (src/functions.rs:33-128) on the start-up scene."""
    v, f, c = scenes.reference_startup_scene()
    ig = L.Integrator.make(h_initial=0.05, samples_per_ray=0, sample_steps=0)
    _case(view, 64, 48, L.Camera.arcball_default(c), L.Settings.make(), ig, fid, v, f, f"predefined-{fid}")


def test_grid_equals_brute_force(view):
    """The grid path (distance-field skip + cell walk) returns the same face, colour and end
    point as the reference's loop over all faces, bit for bit."""
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=3)
    view.resize_texture(96, 64)
    outs = []
    for accel in (L.ACCEL_BRUTE, L.ACCEL_GRID):
        print(f"[case] grid-vs-brute accel={accel}", flush=True)
        fr = run_gpu(view, wl.camera(), wl.settings(), wl.integrator(), "warp", v, f, accel)
        outs.append((fr, view.counters()))
    (b, cb), (g, cg) = outs
    assert (b.hit_faces >= 0).mean() > 0.03
    assert np.array_equal(b.hit_faces, g.hit_faces)
    assert np.array_equal(b.ray_casting, g.ray_casting)
    assert np.array_equal(b.mapping, g.mapping)
    assert cg["triangle_tests"] < cb["triangle_tests"]


def test_overlay_kernel(view):
    """Overlay kernel against the oracle's overlay applied to the GPU's own ray-cast outputs."""
    v, f, c = scenes.reference_startup_scene()
    W, H = 80, 60
    view.resize_texture(W, H)
    for delta in (1, 3):
        st = L.Settings.make(overlay_mode=1, central_difference_delta=delta, lyapunov_scaling=50.0)
        gpu = run_gpu(view, L.Camera.arcball_default(c), st, L.Integrator.make(), "mirage_spherical_sigmoid", v, f)
        ov = O.overlay(O.as_oracle_struct(st, O.SettingsIn), gpu.ray_casting, gpu.mapping)
        assert np.abs(gpu.final.astype(int) - ov["final"].astype(int)).max() <= 1
        inner = ov["lyapunov"] >= 0
        assert np.array_equal(inner, gpu.lyapunov >= 0)
        # lambda via tr^2 - 4 det cancels when the two singular values are close; hold the GPU
        # to the FP32 restatement's own error against FP64 on the same inputs
        ov64 = O.overlay(O.as_oracle_struct(st, O.SettingsIn), gpu.ray_casting, gpu.mapping.astype(np.float64),
                         precision="f64")
        l64 = ov64["lyapunov"][inner].astype(np.float64)
        e32 = np.abs(ov["lyapunov"][inner] - l64)
        eg = np.abs(gpu.lyapunov[inner] - l64)
        assert (eg <= np.maximum(4 * e32, 2e-5 * np.abs(l64) + 1e-7)).mean() >= 0.999
        # exponents buffer: row stride H, compare the slots only one texel writes
        idx = np.arange(W * H).reshape(H, W)
        ey, ex = np.nonzero(inner)
        slots = H * ey + ex
        ok = slots < W * H
        uniq, counts = np.unique(slots[ok], return_counts=True)
        single = np.isin(slots, uniq[counts == 1]) & ok
        e_gpu = gpu.exponents[slots[single]]
        e_ref = ov["exponents"][slots[single]]
        np.testing.assert_allclose(e_gpu, e_ref, rtol=1e-3, atol=1e-6)
        del idx


def test_overlay_disabled_copies_colour(view):
    v, f, c = scenes.reference_startup_scene()
    view.resize_texture(40, 30)
    gpu = run_gpu(view, L.Camera.arcball_default(c), L.Settings.make(overlay_mode=0), L.Integrator.make(),
                  "mirage_spherical_sigmoid", v, f)
    assert np.array_equal(gpu.final, gpu.ray_casting)


def test_sample_rays(view):
    """The 8 sample rays (src/main_view.wgsl:281-314,470-479) against the FP32 oracle."""
    v, f, c = scenes.reference_startup_scene()
    cam = L.Camera.arcball_default(c)
    st = L.Settings.make(mouse_pos=(0.4, 0.55))
    ig = L.Integrator.make()
    view.resize_texture(64, 48)
    run_gpu(view, cam, st, ig, "mirage_spherical_sigmoid", v, f)
    got = view.read_ray_samples()
    c2, s2, i2 = (O.as_oracle_struct(cam, O.CameraIn), O.as_oracle_struct(st, O.SettingsIn),
                  O.as_oracle_struct(ig, O.IntegratorIn))
    ref = O.sample_rays(c2, s2, i2, 3, 64, 48)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, :, 1], ref[:, :, 1])   # colours
    pos_err = np.abs(got[:, :, 0, :3] - ref[:, :, 0, :3]) / np.maximum(1.0, np.abs(ref[:, :, 0, :3]))
    assert pos_err.max() <= 1e-3, pos_err.max()


@pytest.mark.parametrize("W,H", [(96, 64), (40, 72)])
def test_outline_rays(view, W, H):
    """Outline-ray selection (src/main_view.wgsl:381-468): the search runs on the exponents of
    the previous render's overlay and must find exactly the oracle's pixels; the sample rays
    then start there.  W > H and H > W cover the row-stride-H aliasing and index range."""
    v, f, c = scenes.reference_startup_scene()
    cam = L.Camera.arcball_default(c)
    st = L.Settings.make(overlay_mode=1)
    view.resize_texture(W, H)
    print(f"[case] outline rays {W}x{H}", flush=True)
    run_gpu(view, cam, st, L.Integrator.make(), "mirage_spherical_sigmoid", v, f)
    exps = view.read_exponents()
    assert (exps >= 0.8).any()
    ig = L.Integrator.make(outline_rays=True)
    view.set_integrator(ig)
    view.render()
    got = view.read_outline()
    want = O.outline_search(exps, W, H)
    np.testing.assert_array_equal(got, want)
    samples = view.read_ray_samples()
    c2, s2, i2 = (O.as_oracle_struct(cam, O.CameraIn), O.as_oracle_struct(st, O.SettingsIn),
                  O.as_oracle_struct(ig, O.IntegratorIn))
    ref = O.sample_rays(c2, s2, i2, 3, W, H, outline=want)
    ref64 = O.sample_rays(c2, s2, i2, 3, W, H, precision="f64", outline=want)
    np.testing.assert_array_equal(samples[:, :, 1], ref[:, :, 1])
    # outline rays start where the exponent is high, i.e. where neighbouring rays diverge: over
    # 5000 steps FP32 rounding alone moves them, so the GPU is held to 1e-3 or to 4x the FP32
    # oracle's own distance from FP64, whichever is larger
    rel = lambda a, b: (np.abs(a[..., 0, :3] - b[..., 0, :3]) / np.maximum(1.0, np.abs(b[..., 0, :3]))).max(-1)
    e_gpu, e_32 = rel(samples.astype(np.float64), ref64), rel(ref.astype(np.float64), ref64)
    assert (e_gpu <= np.maximum(1e-3, 4.0 * e_32)).all(), (e_gpu.max(), e_32.max())
    # the mouse-position rays differ from the outline rays
    view.set_integrator(L.Integrator.make())
    view.render()
    assert not np.array_equal(view.read_ray_samples()[:, 1:, 0], samples[:, 1:, 0])
    # MainView::render_outline_rays: one outline render, the integrator unchanged
    exps = view.read_exponents()
    view.render_outline_rays()
    np.testing.assert_array_equal(view.read_outline(), O.outline_search(exps, W, H))
    assert view.integrator().outline_rays == 0


def test_high_accuracy_render(view):
    """Enhance (src/main_view.rs:778-802): one render at h_initial = 0.001."""
    v, f, c = scenes.reference_startup_scene()
    cam = L.Camera.arcball_default(c)
    W, H = 24, 16
    view.resize_texture(W, H)
    run_gpu(view, cam, L.Settings.make(), L.Integrator.make(samples_per_ray=0, sample_steps=0),
            "mirage_spherical_sigmoid", v, f)
    view.render_high_accuracy()
    gpu = view.read_frame()
    assert view.integrator().h_initial == pytest.approx(0.1)
    ref, ref64 = run_oracles(cam, L.Settings.make(),
                             L.Integrator.make(h_initial=0.001, samples_per_ray=0, sample_steps=0), 3, v, f, W, H)
    info = compare(gpu, ref, ref64, "enhance")
    assert info["gated"] > 0 and info["hit_ok"] and info["exit_pass"] and info["q99_ok"]
    assert int(np.median(gpu.steps)) >= 5000


def test_counters_and_errors(view):
    v, f = scenes.cornell_box_centered()
    wl = configs.PLAIN_256_SWIRL
    view.resize_texture(32, 32)
    run_gpu(view, wl.camera(), wl.settings(), wl.integrator(), "identity", v, f)
    c = view.counters()
    assert c["rays"] == 32 * 32
    assert c["rk4_steps"] == int(view.read_steps().sum())
    assert c["last_raycast_ms"] > 0
    with pytest.raises(L.LinonError):
        view.set_field(0)          # Custom has no predefined body
    with pytest.raises(L.LinonError):
        view.update_model(v, np.array([[0, 1, 999]], np.uint32))
    with pytest.raises(L.LinonError):
        view.update_settings(L.Settings.make(overlay_mode=7))


@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_1080p_full_resolution_sample(field):
    """The 1920x1080 benchmark frame (Cornell box + 20,480-face icosphere, h 0.005): the whole
    frame on the GPU, 1024 random pixels through the FP32 and FP64 oracles; lambda from the
    GPU's own overlay at full resolution."""
    wl = configs.OVERLAY_1080P
    v, f = wl.mesh()
    with L.MainView(wl.width, wl.height) as big:
        big.update_model(v, f)
        configs.setup_view(big, wl, field)
        big.render()
        fr = big.read_frame()
    rng = np.random.default_rng(7)
    gy = rng.integers(0, wl.height, 1024)
    gx = rng.integers(0, wl.width, 1024)
    px = np.stack([gx, gy], 1)
    ty = wl.height - 1 - gy  # texel row of invocation row gy
    r32, ref64 = oracle_pixels(wl, L.FIELDS[field], v, f, px)
    class G:
        pass
    g = G()
    g.hit_faces = fr.hit_faces[ty, gx][None]
    g.mapping = fr.mapping[ty, gx][None]
    g.ray_casting = fr.ray_casting[ty, gx][None]
    ref = {"mapping": r32["mapping"][None], "hit": r32["hit"][None], "noise_flag": np.zeros((1, len(px)), np.uint8),
           "color8": g.ray_casting}
    lam = fr.lyapunov[ty, gx][None]
    info = compare(g, ref, ref64, f"1080p-{field}", lam_gpu=lam)
    assert info["gated"] > 100, info
    assert info["hit_ok"], info
    assert info["exit_pass"], info
    assert info["q99_ok"], info


@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_wgsl_body_matches_predefined_functor(view, field):
    """MainView::reload_shader path: the WGSL body of a benchmark field, translated and compiled
    with hiprtc, renders the same frame as the statically compiled functor."""
    wl = configs.OVERLAY_1080P
    v, f = scenes.benchmark_scene(level=2)
    view.resize_texture(64, 40)
    try:
        for math in (L.MATH_IEEE, L.MATH_WGSL):  # the body is compiled for the same build as the functor
            view.set_math(math)
            print(f"[case] wgsl-vs-functor {field} math={math}", flush=True)
            a = run_gpu(view, wl.camera(), wl.settings(), wl.integrator(), field, v, f)
            b = run_gpu(view, wl.camera(), wl.settings(), wl.integrator(), configs.FIELD_WGSL[field], v, f)
            assert np.array_equal(a.hit_faces, b.hit_faces)
            assert np.array_equal(a.steps, b.steps)
            d = np.abs(a.mapping.astype(np.float64) - b.mapping) / np.maximum(1.0, np.abs(a.mapping))
            assert d.max() <= 1e-5, d.max()
    finally:
        view.set_math(L.MATH_AUTO)


def test_wgsl_compile_error_keeps_previous_field(view):
    v, f, c = scenes.reference_startup_scene()
    view.resize_texture(32, 24)
    a = run_gpu(view, L.Camera.arcball_default(c), L.Settings.make(), L.Integrator.make(), "identity", v, f)
    with pytest.raises(L.LinonError) as e:
        view.reload_shader("return v +;")
    assert e.value.status == 2 and "field_function:1" in e.value.message
    view.render()
    b = view.read_frame()
    assert np.array_equal(a.mapping, b.mapping)


def test_shared_denominator_division_is_ieee():
    """The IEEE build divides vec3 by a scalar through one shared, correctly rounded reciprocal
    and Markstein's correction, and the predefined fields' in-range quotients (sigmoid's
    reciprocal, the refraction indices and their ratio) the same way without range tests; all
    must equal IEEE division bit for bit: every mantissa of the reciprocal at exponents
    -126..125, every numerator mantissa against 16 structured denominators, 2^28 random pairs
    (zeros, out-of-range and signed operands included), sigmoid_floor at every float and
    refraction_in_range on 2^28 random temperature pairs and directions.  On the same draws, the
    first RK4 stage of hot air and the four mirages, which skips the index ratio n(T)/n(T) = 1,
    must equal the field evaluated at p_prev = p.  The bounded divisions of normalize and of
    the lens field (whose relaxed range admits zeros and numerators down to 2^-102) are also
    checked at every numerator mantissa of 18 exponents from the subnormals up to 2^0, beside
    signed zeros, against 19 denominators in and out of their ranges; the lens quotient at
    its floor s = 0.001 (the captured rays' path) at every numerator of magnitude below 2^13; and
    the RK4 average's division by six at every float (apart from its documented ties of
    subnormal quotients)."""
    print("[case] division self-test", flush=True)
    res = L.selftest_division(random_pairs=1 << 28)
    print(res, flush=True)
    assert res["comparisons"] > 20_000_000_000
    assert res["recip_mismatches"] == 0 and res["quotient_mismatches"] == 0, res


def test_square_root_is_ieee():
    """The IEEE build's square root (the hardware root plus a one-ulp correction, inputs below
    2^-96 rescaled) equals the compiler's correctly rounded square root at every float, and so
    does its form without the rescaling (refraction's root) at every float whose magnitude is
    not in (0, 2^-96)."""
    print("[case] sqrt self-test", flush=True)
    res = L.selftest_sqrt()
    print(res, flush=True)
    assert res["comparisons"] == 1 << 32
    assert res["mismatches"] == 0, res


def test_exponential_is_within_the_wgsl_bound():
    """Hot air's default variant takes the sigmoid's exponential as the hardware 2^t of
    RN(y log2 e): at every float y in [-87, 87] within WGSL's bound for exp, 3 + 2 |y| ulp of
    the FP64 exponential (the product's rounding alone can cost about 1.2 |y| ulp)."""
    print("[case] exp self-test", flush=True)
    res = L.selftest_exp()
    print(res, flush=True)
    assert res["comparisons"] == 1 << 32
    assert res["over_bound"] == 0, res


def test_wgsl_integer_division_is_total(view):
    """A WGSL body's integer / and % follow WGSL, not C++: x / 0 = x, x % 0 = 0,
    INT_MIN / -1 = INT_MIN, INT_MIN % -1 = 0 (i32 and u32).  The body returns a constant
    velocity built from those results, so each ray moves in a straight line at that velocity and
    the end point reveals it: V = (1, 0.5, 2)."""
    body = ("let z = i32(p.x * 0.0);\n"
            "let m = -2147483647 - 1 + z;\n"
            "let n = z - 1;\n"
            "let u = u32(z);\n"
            "return vec3<f32>(f32(7 / z) / 7.0, f32(7 % z) + f32(m % n) + f32(9u % u) + 0.5,\n"
            "                 f32(m / n) / -2147483648.0 + f32(9u / u) / 9.0);")
    view.resize_texture(16, 8)
    ig = L.Integrator.make(h_initial=0.1, adaptive=False, samples_per_ray=0, sample_steps=0)
    cam = L.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov_deg=60.0)
    fr = run_gpu(view, cam, L.Settings.make(), ig, body, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    v = (fr.mapping[..., :3].astype(np.float64) - [0.0, 0.0, 3.0]) / (fr.steps[..., None] * 0.1)
    np.testing.assert_allclose(v, np.broadcast_to([1.0, 0.5, 2.0], v.shape), atol=1e-4)


@pytest.mark.parametrize("field", ["lens", "hot_air"])
def test_outline_search_at_4k(field):
    """The outline search on the GPU's own 3840x2160 exponents buffer (row stride H, rows
    aliasing since W > H) finds exactly the oracle search's pixels.  Lens's buffer holds
    highlighted texels (exponent >= 0.8) for the search to find; hot air's, at this pixel pitch,
    holds none (its largest exponent is printed), so every search comes back empty and the
    sample rays start at texel (0, 0), as in the reference.  Cornell box, h 0.005 (the search
    reads only the exponents)."""
    wl = configs.STRESS_4K
    v, f = scenes.cornell_box_centered()
    with L.MainView(wl.width, wl.height) as big:
        big.update_model(v, f)
        big.update_camera(wl.camera())
        big.update_settings(wl.settings())
        ig = wl.integrator()
        ig.h_initial = 0.005
        big.set_integrator(ig)
        big.set_field(field)
        print(f"[case] 4K outline search {field}", flush=True)
        big.render()          # exponents of this frame ...
        exps = big.read_exponents()
        big.render()          # ... are what the next frame's outline search reads
        got = big.read_outline()
    n_hi = int((exps >= 0.8).sum())
    print(f"{field}: {n_hi} exponents >= 0.8 (max {exps.max():.3f}), found {got[:, 2].tolist()}", flush=True)
    np.testing.assert_array_equal(got, O.outline_search(exps, wl.width, wl.height))
    assert (got[:, 2].sum() > 0) == (n_hi > 0)
    if field == "lens":
        assert n_hi > 0


@pytest.mark.parametrize("field", ["lens", "hot_air"])
def test_4k_stress_frame_sample(field):
    """The 3840x2160 stress frame (Cornell box + the 1,310,720-face level-8 icosphere, h 0.001,
    5001 RK4 steps per ray, overlay, outline rays): the whole frame on the GPU, 256 random pixels
    through the FP32 and FP64 oracles (brute force over all faces, with the exact block
    prefilter); lambda from the GPU's own overlay, scaled to the 1080-row pitch."""
    wl = configs.STRESS_4K
    v, f = wl.mesh()
    with L.MainView(wl.width, wl.height) as big:
        big.update_model(v, f)
        configs.setup_view(big, wl, field)
        print(f"[case] 4K stress frame {field}", flush=True)
        big.render()
        fr = big.read_frame()
    rng = np.random.default_rng(11)
    gy = rng.integers(0, wl.height, 256)
    gx = rng.integers(0, wl.width, 256)
    px = np.stack([gx, gy], 1)
    ty = wl.height - 1 - gy
    r32, ref64 = oracle_pixels(wl, L.FIELDS[field], v, f, px)

    class G:
        pass
    g = G()
    g.hit_faces = fr.hit_faces[ty, gx][None]
    g.mapping = fr.mapping[ty, gx][None]
    g.ray_casting = fr.ray_casting[ty, gx][None]
    ref = {"mapping": r32["mapping"][None], "hit": r32["hit"][None], "noise_flag": np.zeros((1, len(px)), np.uint8),
           "color8": g.ray_casting}
    lam = fr.lyapunov[ty, gx][None] * (wl.height / 1080)
    info = compare(g, ref, ref64, f"4k-{field}", lam_gpu=lam)
    assert info["gated"] > 30, info
    assert info["hit_ok"], info
    assert info["exit_pass"], info
    assert info["q99_ok"], info
    assert (r32["hit"] >= 0).any()
