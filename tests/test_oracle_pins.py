"""Pins of the CPU oracle against what the mathematics fixes (no GPU needed).

Each test is chosen so that a plausible mistake in the restatement (a dropped term, a wrong
sign or index, a swapped operand, a wrong stage weight) fails it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from linon_amd import FIELDS, Camera, scenes
from oracle import oracle as O


# ---------------------------------------------------------------------------------------
# Moeller-Trumbore (src/main_view.wgsl:150-175)
# ---------------------------------------------------------------------------------------
TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float64)


@pytest.mark.parametrize("o,d,t", [
    ((0.25, 0.25, 1.0), (0, 0, -1), 1.0),     # interior
    ((0.25, 0.25, -2.0), (0, 0, 1), 2.0),     # two-sided: from behind
    ((0.5, 0.0, 1.0), (0, 0, -1), 1.0),       # on edge v = 0 (inclusive)
    ((0.0, 0.5, 1.0), (0, 0, -1), 1.0),       # on edge u = 0
    ((0.5, 0.5, 1.0), (0, 0, -1), 1.0),       # on edge u + v = 1
    ((1.0, 0.0, 1.0), (0, 0, -1), 1.0),       # vertex
    ((0.6, 0.6, 1.0), (0, 0, -1), -1.0),      # outside u + v > 1
    ((-0.1, 0.2, 1.0), (0, 0, -1), -1.0),     # outside u < 0
    ((0.25, 0.25, 1.0), (1, 0, 0), -1.0),     # parallel
    ((0.25, 0.25, 1.0), (0, 0, 1), -1.0),     # behind the origin
    ((0.25, 0.25, 0.0), (0, 0, -1), -1.0),    # t = 0 is not > eps
])
def test_mt_known_answers(o, d, t):
    assert O.hit_triangle(TRI, o, d) == pytest.approx(t, abs=1e-12)
    assert O.hit_triangle(TRI, o, d, precision="f32") == pytest.approx(t, abs=1e-6)


def test_mt_degenerate_determinant():
    tiny = TRI * 1e-4  # edge lengths 1e-4: a = 1e-8 < 1e-7 is rejected
    assert O.hit_triangle(tiny, (2e-5, 2e-5, 1.0), (0, 0, -1)) == -1.0
    small = TRI * 1e-3  # a = 1e-6 is accepted
    assert O.hit_triangle(small, (2e-4, 2e-4, 1.0), (0, 0, -1)) == pytest.approx(1.0)


def test_mt_random_against_plane_barycentric():
    """FP64 oracle against an independent plane-intersection + barycentric solve."""
    rng = np.random.default_rng(0)
    n_checked = 0
    for _ in range(2000):
        v = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 2
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        t = O.hit_triangle(v.ravel(), o, d)
        # solve o + t d = v0 + u e1 + w e2
        A = np.stack([-d, v[1] - v[0], v[2] - v[0]], 1)
        if abs(np.linalg.det(A)) < 1e-6:
            continue
        tt, u, w = np.linalg.solve(A, o - v[0])
        inside = u >= 0 and w >= 0 and u + w <= 1 and tt > 1e-7
        margin = min(abs(u), abs(w), abs(1 - u - w), abs(tt))
        if margin < 1e-9:
            continue
        n_checked += 1
        if inside:
            assert t == pytest.approx(tt, rel=1e-9, abs=1e-9)
        else:
            assert t == -1.0
    assert n_checked > 1500


def test_ray_color_closest_and_ties():
    # two coplanar copies of one triangle (faces 0 and 2) and a nearer one (face 1)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5]], np.float32)
    f = np.array([[0, 1, 2], [3, 4, 5], [0, 1, 2]], np.uint32)
    face, rgba = O.ray_color(v, f, (0.2, 0.2, 1.0), (0, 0, -1), 10.0)
    assert face == 1 and rgba[3] == pytest.approx(0.5)
    # lowest index wins an exact tie
    face, _ = O.ray_color(v, np.array([[3, 4, 5], [0, 1, 2], [3, 4, 5]], np.uint32), (0.2, 0.2, 1.0), (0, 0, -1), 10)
    assert face == 0
    f2 = np.array([[0, 1, 2], [0, 1, 2]], np.uint32)
    face, _ = O.ray_color(v, f2, (0.2, 0.2, 1.0), (0, 0, -1), 10.0)
    assert face == 0
    # max_dist is exclusive: t must be < max_dist
    face, _ = O.ray_color(v, f2, (0.2, 0.2, 1.0), (0, 0, -1), 1.0)
    assert face == -1
    face, _ = O.ray_color(v, f2, (0.2, 0.2, 1.0), (0, 0, -1), 1.0001)
    assert face == 0


def test_ray_color_shading_closed_form():
    """Camera-as-light Blinn-Phong (src/main_view.wgsl:199-218): normal (0,0,1) seen head-on
    gives intensity 1: colour = (0.01 + 1 + 1) * |n| = (0, 0, 2.01)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    face, rgba = O.ray_color(v, f, (0.2, 0.2, 1.0), (0, 0, -1), 10.0)
    np.testing.assert_allclose(rgba[:3], [0, 0, 2.01], rtol=1e-6)
    # 60 degrees off the normal, seen from behind: intensity 0.5, specular 0.5^64
    d = np.array([0.0, math.sin(math.pi / 3), math.cos(math.pi / 3)], np.float32)
    face, rgba = O.ray_color(v, f, np.array([0.2, 0.2, 0.0]) - d, d, 10.0)
    assert face == 0
    np.testing.assert_allclose(rgba[2], 0.01 + 0.5 + 0.5 ** 64, rtol=1e-5)


# ---------------------------------------------------------------------------------------
# Camera (src/main_view.wgsl:347-366)
# ---------------------------------------------------------------------------------------
def test_camera_centre_and_corner_rays():
    cam = O.camera_in((0, 0, 3), (0, 0, -1), (0, 1, 0))
    W, H = 101, 51
    np.testing.assert_allclose(O.camera_ray(cam, W, H, 50, 25), [0, 0, -1], atol=1e-15)
    # corner (0, 0): u = -vw/2, v = -vh/2 along horizontal = cross(up, w) = +x, vertical = +y
    vh = 0.025
    vw = W / H * vh
    expect = np.array([-vw / 2, -vh / 2, -0.035])
    np.testing.assert_allclose(O.camera_ray(cam, W, H, 0, 0), expect / np.linalg.norm(expect), rtol=1e-12)
    # field-of-view extension: the top-centre ray is fov/2 above the axis
    cam60 = O.camera_in((0, 0, 3), (0, 0, -1), (0, 1, 0), 60.0)
    d = O.camera_ray(cam60, W, H, 50, H - 1)
    assert math.degrees(math.atan2(d[1], -d[2])) == pytest.approx(30.0, abs=1e-9)


# ---------------------------------------------------------------------------------------
# refraction helpers (src/main_view.wgsl:111-136)
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("T,n", [(15, 0.9993990), (20, 0.9823091), (30, 0.9498201), (100, 0.7711449),
                                 (200, 0.6075901)])
def test_refraction_index_values(T, n):
    assert O.refraction_index(T) == pytest.approx(n, abs=5e-7)


def test_refraction_snell_and_total_reflection():
    n = np.array([0.0, 1.0, 0.0])
    for theta in (0.0, 0.3, 0.7, 1.2):
        v = np.array([math.sin(theta), -math.cos(theta), 0.0])  # travelling against n
        r = O.refraction(200.0, 15.0, v, n)                      # hot -> cold: eta = n200/n15 < 1
        eta = O.refraction_index(200.0) / O.refraction_index(15.0)
        assert np.linalg.norm(r) == pytest.approx(1.0, abs=1e-12)
        assert r[0] == pytest.approx(eta * math.sin(theta), abs=1e-12)   # Snell: tangential part * eta
        assert r[1] < 0
    # cold -> hot at grazing incidence: eta > 1 -> total internal reflection = mirror image
    theta = 1.4
    v = np.array([math.sin(theta), -math.cos(theta), 0.0])
    r = O.refraction(15.0, 200.0, v, n)
    np.testing.assert_allclose(r, [v[0], -v[1], 0.0], atol=1e-12)


# ---------------------------------------------------------------------------------------
# field bodies (src/functions.rs:33-128 and the benchmark fields)
# ---------------------------------------------------------------------------------------
def test_lorenz_and_roessler_fixed_points():
    beta, rho = 8.0 / 3.0, 28.0
    c = math.sqrt(beta * (rho - 1))
    z = np.zeros(3)
    for p in ([c, c, rho - 1], [-c, -c, rho - 1], [0, 0, 0]):
        np.testing.assert_allclose(O.field(FIELDS["lorenz"], z, p, z, z, 0.0), 0, atol=1e-12)
    np.testing.assert_allclose(O.field(FIELDS["lorenz"], z, [1, 2, 3], z, z, 0.0), [10, 23, 2 - 8], atol=1e-12)
    a, b, cc = 0.1, 0.1, 14.0
    disc = math.sqrt(cc * cc - 4 * a * b)
    for s in (+1, -1):
        x = (cc + s * disc) / 2
        p = [x, -x / a, x / a]
        np.testing.assert_allclose(O.field(FIELDS["roessler"], z, p, z, z, 0.0), 0, atol=1e-9)


def test_translation_rotation_fields():
    z = np.zeros(3)
    v0 = np.array([0.3, -0.2, 0.9])
    np.testing.assert_allclose(O.field(FIELDS["translation_x"], z, z, v0, z, 2.0), v0 + [1.0, 0, 0])
    np.testing.assert_allclose(O.field(FIELDS["translation_z"], z, z, v0, z, 2.0), v0 + [0, 0, 1.0])
    # rotation: rotateZ drops z (zero third row), so at t = 0: normalize(v0 + (v0.x, v0.y, 0))
    r0 = O.field(FIELDS["rotation"], z, z, v0, z, 0.0)
    e = v0 + [v0[0], v0[1], 0.0]
    np.testing.assert_allclose(r0, e / np.linalg.norm(e), atol=1e-14)
    # at t = 1 (phi = pi): rotateZ(v0) = (-x, -y, 0) -> normalize((0, 0, z))
    np.testing.assert_allclose(O.field(FIELDS["rotation"], z, z, v0, z, 1.0), [0, 0, 1], atol=1e-12)


def test_swirl_lens_warp_fields():
    z = np.zeros(3)
    v0 = np.array([0.0, 0.1, -1.0])
    p = np.array([0.0, 0.0, 2.0])          # |p| = 2 -> rotation by 1 rad about +y
    f = O.field(FIELDS["swirl"], z, p, v0, z, 0.0)
    np.testing.assert_allclose(f, [math.sin(1.0) * -1.0, 0.1, math.cos(1.0) * -1.0], atol=1e-14)
    v = np.array([0.0, 0.0, -1.0])
    np.testing.assert_allclose(O.field(FIELDS["lens"], z, [2.0, 0, 0], z, v, 0.0), v - [0.3 * 2 / 8, 0, 0])
    np.testing.assert_allclose(O.field(FIELDS["lens"], z, [0.01, 0, 0], z, v, 0.0), v - [0.3 * 0.01 / 1e-3, 0, 0])
    q = np.array([math.pi / 8, math.pi / 8, 0.0])   # sin(4 * pi/8) = 1
    np.testing.assert_allclose(O.field(FIELDS["warp"], z, q, v0, z, 0.0), v0 + [0.3, 0, 0.3], atol=1e-14)


def test_mirage_and_hot_air_temperature_regimes():
    v = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    far = np.array([3.0, 3.0, 3.0])
    for fid in (1, 3, 19):   # far from the source: equal temperatures -> eta = 1 -> v
        np.testing.assert_allclose(O.field(fid, far, far, v, v, 0.0), v, atol=1e-12)
    # plane mirages: beyond max_dist from the plane y = 0.1 the medium is uniform
    for fid in (2, 4):
        p = np.array([0.0, 0.5, 0.0])
        np.testing.assert_allclose(O.field(fid, p, p, v, v, 0.0), v, atol=1e-12)
    # crossing from hot (at the plane) to cold: the ray bends away from the normal
    pin, pout = np.array([0, 0.1, 0]), np.array([0, 0.12, 0])
    r = O.field(2, pin, pout, v, v, 0.0)
    assert abs(r[0]) < abs(v[0]) or abs(r[1]) != abs(v[1])


# Temperature laws of the refracting fields, pinned at points where the law has a closed-form
# value.  Each field evaluation across two such points must be Snell refraction at the index
# ratio n(T(p_prev)) / n(T(p)) (n is pinned by its table above) about the field's normal.  The
# refracted direction below is built from the angle law itself (tangential part scaled by eta,
# unit length, same side as v), not from refraction()'s formula.
def _sig(x):
    return 1.0 / (1.0 + math.exp(-x))


def _snell(v, n, eta):
    v = np.asarray(v, np.float64)
    n = np.asarray(n, np.float64) / np.linalg.norm(n)
    vn = float(np.dot(v, n))
    tang = eta * (v - vn * n)
    s2 = float(np.dot(tang, tang))
    assert s2 < 1.0, "pin cases avoid total internal reflection"
    return tang + math.copysign(math.sqrt(1.0 - s2), vn) * n


def _unit(*x):
    x = np.asarray(x, np.float64)
    return x / np.linalg.norm(x)


HOT_AIR_C = np.array([0.0, -0.5, 0.0])          # functions of the benchmark's hot-air field
MIRAGE_SPH_C = np.array([-0.5, 0.5, -0.5])      # src/functions.rs:39,70
PLANE_Y = 0.1                                   # src/functions.rs:56,87


# (field id, p_prev, p, T(p_prev), T(p), normal) with the temperatures in closed form:
#  hot air  T = 20 + 80 sigmoid(8 (0.5 - |p - c|)): 60 at |p - c| = 0.5, 20 + 80 s(4) at c,
#           20 + 80 s(2) at 0.25, 20 + 80 s(-4) at 1.0
#  spherical mirage (linear) part = min(1, d / 0.25): 200 at the centre, 107.5 at d = 0.125,
#           153.75 at 0.0625, 15 from 0.25 on
#  spherical sigmoid part = min(1, s(12 d / 0.25 - 6)): 107.5 at d = 0.125, 15 s(-6) + 200 (1 - s(-6))
#           at the centre, 15 s(3) + 200 (1 - s(3)) at 0.1875
#  plane mirages (y = 0.1, max_dist 0.01, 30 -> 15): linear 30 on the plane, 22.5 at 0.005 either
#           side, 15 from 0.01 on; sigmoid 22.5 at 0.005, 15 s(-6) + 30 (1 - s(-6)) on the plane
_u = _unit(0.3, 0.8, -0.52)
_w = _unit(-0.6, 0.2, 0.77)
TEMPERATURE_CASES = [
    ("hot-air 60 -> 20+80s(2)", 19, HOT_AIR_C + 0.5 * _u, HOT_AIR_C + 0.25 * _w, 60.0, 20 + 80 * _sig(2.0),
     _w),
    ("hot-air centre -> 60", 19, HOT_AIR_C, HOT_AIR_C + 0.5 * _w, 20 + 80 * _sig(4.0), 60.0, _w),
    ("hot-air 20+80s(-4) -> 60", 19, HOT_AIR_C + 1.0 * _u, HOT_AIR_C + 0.5 * _u, 20 + 80 * _sig(-4.0), 60.0, _u),
    ("mirage-sph 200 -> 107.5", 1, MIRAGE_SPH_C, MIRAGE_SPH_C + 0.125 * _u, 200.0, 107.5, _u),
    ("mirage-sph 153.75 -> 15", 1, MIRAGE_SPH_C + 0.0625 * _w, MIRAGE_SPH_C + 0.3 * _u, 153.75, 15.0, _u),
    ("mirage-sph-sig centre -> 107.5", 3, MIRAGE_SPH_C, MIRAGE_SPH_C + 0.125 * _w,
     15 * _sig(-6.0) + 200 * (1 - _sig(-6.0)), 107.5, _w),
    ("mirage-sph-sig 0.1875 -> 107.5", 3, MIRAGE_SPH_C + 0.1875 * _u, MIRAGE_SPH_C + 0.125 * _u,
     15 * _sig(3.0) + 200 * (1 - _sig(3.0)), 107.5, _u),
    ("mirage-plane 30 -> 22.5", 2, np.array([0.2, PLANE_Y, -0.1]), np.array([0.2, PLANE_Y + 0.005, -0.1]), 30.0,
     22.5, np.array([0.0, 1.0, 0.0])),
    ("mirage-plane 22.5 (below) -> 15", 2, np.array([0.0, PLANE_Y - 0.005, 0.3]), np.array([0.0, PLANE_Y + 0.02, 0.3]),
     22.5, 15.0, np.array([0.0, 1.0, 0.0])),
    ("mirage-plane-sig plane -> 22.5", 4, np.array([0.1, PLANE_Y, 0.0]), np.array([0.1, PLANE_Y - 0.005, 0.0]),
     15 * _sig(-6.0) + 30 * (1 - _sig(-6.0)), 22.5, np.array([0.0, 1.0, 0.0])),
]


@pytest.mark.parametrize("label,fid,p_prev,p,t_in,t_out,normal", TEMPERATURE_CASES,
                         ids=[c[0] for c in TEMPERATURE_CASES])
def test_refracting_fields_temperature_laws(label, fid, p_prev, p, t_in, t_out, normal):
    eta = O.refraction_index(t_in) / O.refraction_index(t_out)
    assert abs(eta - 1.0) > 1e-3, "the pin needs a real index step"
    # directions at 0.2 and 0.5 rad from the normal, travelling along it and against it, each
    # below total internal reflection (eta <= 1.4 here, so sin(0.5) eta < 0.7)
    n = np.asarray(normal, np.float64)
    tan = np.cross(n, [0.3, 0.5, 0.81])
    tan /= np.linalg.norm(tan)
    for theta in (0.2, 0.5):
        for side in (1.0, -1.0):
            v = side * math.cos(theta) * n + math.sin(theta) * tan
            got = O.field(fid, p_prev, p, v, v, 0.0)
            np.testing.assert_allclose(got, _snell(v, n, eta), atol=1e-12, err_msg=f"{label} theta={theta}")


def test_temperature_pins_separate_plausible_mistakes():
    """The temperature pairs above differ enough that a mistaken law (sigmoid argument negated,
    slope 8 -> 4, centre moved by 0.1, hot and cold swapped) gives an index ratio whose
    refraction differs by more than the pins' tolerance."""
    v = _unit(0.1, -0.3, 1.0)
    n = _w
    good = _snell(v, n, O.refraction_index(20 + 80 * _sig(4.0)) / O.refraction_index(60.0))
    for wrong_t_in in (20 + 80 * _sig(-4.0), 20 + 80 * _sig(2.0), 20 + 80 * _sig(8 * (0.5 - 0.1))):
        bad = _snell(v, n, O.refraction_index(wrong_t_in) / O.refraction_index(60.0))
        assert np.abs(bad - good).max() > 1e-4
    swapped = _snell(v, n, O.refraction_index(60.0) / O.refraction_index(20 + 80 * _sig(4.0)))
    assert np.abs(swapped - good).max() > 1e-4


# ---------------------------------------------------------------------------------------
# RK4 loop (src/main_view.wgsl:229-279)
# ---------------------------------------------------------------------------------------
def _ig(**k):
    base = dict(h_initial=0.01, t_max=5.0, adaptive=False, max_steps=100000, samples_per_ray=0, sample_steps=0)
    base.update(k)
    return O.integrator_in(**base)


def test_f32_step_count_is_the_f32_running_sum():
    for h in (0.1, 0.01, 0.005, 0.003):
        t = np.float32(0)
        n = 0
        while t <= np.float32(5.0):
            t = np.float32(t + np.float32(h))
            n += 1
        steps, _, _ = O.trace(FIELDS["identity"], _ig(h_initial=h), (0, 0, 0), (0, 0, -1), precision="f32")
        assert steps == n, (h, steps, n)


def test_identity_field_is_a_straight_line():
    d = np.array([0.2, -0.1, -1.0])
    d /= np.linalg.norm(d)
    steps, end, face = O.trace(FIELDS["identity"], _ig(h_initial=0.01), (0.1, 0.2, 3.0), d)
    t_final = steps * 0.01
    np.testing.assert_allclose(end, np.array([0.1, 0.2, 3.0]) + d * t_final, atol=1e-12)


def test_identity_field_hits_like_a_linear_ray_cast():
    v, f, c = scenes.reference_startup_scene()
    cam = O.as_oracle_struct(Camera.arcball_default(c), O.CameraIn)
    W, H = 16, 12
    r = O.render(cam, O.settings_in(), _ig(h_initial=0.01), FIELDS["identity"], v, f, W, H, threads=4)
    for gy in range(H):
        for gx in range(W):
            d = O.camera_ray(cam, W, H, gx, gy).astype(np.float32)
            face, _ = O.ray_color(v, f, np.array(cam.origin[:3], np.float32), d, 100.0)
            assert r["hit"][H - 1 - gy, gx] == face


def test_translation_field_rk4_is_exact():
    """f = v0 + (0.5 t, 0, 0) is polynomial in t, so RK4 reproduces p(t) = p0 + v0 t + (t^2/4, 0, 0)."""
    d = np.array([0.0, 0.6, -0.8])
    steps, end, _ = O.trace(FIELDS["translation_x"], _ig(h_initial=0.0625, t_max=1e9, max_steps=32), (0, 0, 0), d)
    t = 32 * 0.0625
    np.testing.assert_allclose(end, d * t + [t * t / 4, 0, 0], atol=1e-12)


@pytest.mark.parametrize("field", ["swirl", "warp"])
def test_rk4_fourth_order_convergence(field):
    """Fields that depend on p and v0 only give p' = f(p): halving h must cut the error ~16x."""
    d = np.array([0.1, 0.05, -1.0])
    d /= np.linalg.norm(d)
    o = np.array([0.2, 0.1, 1.5])
    T = 2.0
    ends = {}
    for k in (3, 4, 5, 8):
        h = 2.0 ** -k
        _, ends[k], _ = O.trace(FIELDS[field], _ig(h_initial=h, t_max=1e9, max_steps=int(T / h)), o, d)
    e3 = np.linalg.norm(ends[3] - ends[8])
    e4 = np.linalg.norm(ends[4] - ends[8])
    e5 = np.linalg.norm(ends[5] - ends[8])
    assert 12 < e3 / e4 < 20 and 12 < e4 / e5 < 20, (e3, e4, e5)


def test_adaptive_refinement_triggers_on_a_jump():
    """A ray crossing the plane mirage's 0.01-thick hot layer (y = 0.1) sees |V - last_V| jump
    more than 10x: with adaptive sampling the loop redoes that step at h = 0.001."""
    o, d = (0.0, 0.5, 0.0), (0.0, -1.0, 0.05)
    n_fixed, end_fixed, _ = O.trace(FIELDS["mirage_plane"], _ig(h_initial=0.1, adaptive=False), o, d)
    n_adapt, end_adapt, _ = O.trace(FIELDS["mirage_plane"], _ig(h_initial=0.1, adaptive=True), o, d)
    assert n_fixed == 51
    assert n_adapt > n_fixed
    assert not np.allclose(end_fixed, end_adapt)


def test_escape_mode_stops_outside_the_sphere():
    steps, end, _ = O.trace(FIELDS["identity"], _ig(h_initial=0.01, mode=1, escape_radius=2.0), (0, 0, 3.0),
                            (0, 1.0, 0.0))
    assert steps == 1 and end[1] == pytest.approx(0.01)
    steps, end, _ = O.trace(FIELDS["identity"], _ig(h_initial=0.01, mode=1, escape_radius=2.0), (0, 0, 3.0),
                            (0, 0.0, -1.0))
    # enters at z = 2, leaves at z = -2 moving outward: first step with |p| > 2 and p.v > 0
    assert end[2] < -2.0 and end[2] > -2.0 - 0.0101


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_clip_exit_is_the_sphere_crossing(precision):
    """With clip_exit, a straight ray (identity field) ends where the line leaves the sphere:
    o + t d with |o + t d| = R, t the larger root of t^2 + 2 (o.d) t + |o|^2 - R^2 = 0."""
    rng = np.random.default_rng(11)
    R = 2.0
    tol = 1e-12 if precision == "f64" else 2e-5   # f32: p += d h rounds once per step
    for _ in range(200):
        o = rng.normal(size=3)
        o *= rng.uniform(0.0, 1.9) / np.linalg.norm(o)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        h = float(rng.choice([0.01, 0.037, 0.1]))
        ig = _ig(h_initial=h, t_max=1e30, mode=1, escape_radius=R, clip_exit=True)
        _, end, _ = O.trace(FIELDS["identity"], ig, o, d, precision=precision)
        b = float(np.dot(o, d))
        t = -b + math.sqrt(b * b - (float(np.dot(o, o)) - R * R))
        np.testing.assert_allclose(end, o + t * d, atol=tol)
        # without the clip the ray ends at the first step point past the sphere
        _, end_raw, _ = O.trace(FIELDS["identity"], _ig(h_initial=h, t_max=1e30, mode=1, escape_radius=R), o, d,
                                precision=precision)
        assert np.linalg.norm(end_raw) > R and np.linalg.norm(end_raw - end) <= h * (1 + 1e-5)


def test_clip_exit_on_a_curved_path():
    """Swirl bends the path; the clipped end lies on the sphere, on the last step's segment."""
    ig = _ig(h_initial=0.05, t_max=1e30, mode=1, escape_radius=2.0, clip_exit=True)
    ig_raw = _ig(h_initial=0.05, t_max=1e30, mode=1, escape_radius=2.0)
    for o, d in (((0.3, 0.2, 1.9), (-0.1, 0.05, -1.0)), ((1.5, -0.4, 0.2), (-1.0, 0.3, 0.1))):
        dd = np.asarray(d) / np.linalg.norm(d)
        n, end, _ = O.trace(FIELDS["swirl"], ig, o, dd)
        n_raw, end_raw, _ = O.trace(FIELDS["swirl"], ig_raw, o, dd)
        _, prev, _ = O.trace(FIELDS["swirl"], _ig(h_initial=0.05, t_max=1e30, max_steps=n_raw - 1), o, dd)
        assert n == n_raw
        assert np.linalg.norm(end) == pytest.approx(2.0, abs=1e-12)
        # collinear with the last step and between its ends
        seg = end_raw - prev
        u = float(np.dot(end - prev, seg) / np.dot(seg, seg))
        assert 0.0 <= u <= 1.0
        np.testing.assert_allclose(prev + u * seg, end, atol=1e-12)


# ---------------------------------------------------------------------------------------
# Lyapunov overlay (src/overlay.wgsl:4-146)
# ---------------------------------------------------------------------------------------
def test_lyapunov_equals_largest_singular_value():
    rng = np.random.default_rng(3)
    for _ in range(200):
        A = rng.normal(size=(3, 2)) * rng.uniform(0.01, 3)
        b = rng.normal(size=3)
        m = lambda x, y: A @ np.array([x, y]) + b
        for delta in (1, 3):
            lam = O.lyapunov(m(delta, 0), m(-delta, 0), m(0, delta), m(0, -delta))
            s = np.linalg.svd(A, compute_uv=False)
            assert lam == pytest.approx(delta * s[0], rel=1e-6, abs=1e-9)


def test_overlay_blend_and_exponent_layout():
    W, H, delta = 12, 9, 1
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    mapping = np.zeros((H, W, 4), np.float32)
    lam = 0.098
    mapping[..., 0] = lam * xx      # lambda = 0.098 everywhere inside
    mapping[..., 3] = 1
    color = np.full((H, W, 4), 100, np.uint8)
    color[..., 3] = 255
    st = O.settings_in(overlay_mode=1, delta=delta, scaling=50.0)
    ov = O.overlay(st, color, mapping)
    s = math.exp(50 * lam - 5)        # 0.905 >= 0.8: blend with alpha = s / 2
    inside = ov["lyapunov"] >= 0
    assert inside.sum() == (W - 4) * (H - 4)
    np.testing.assert_allclose(ov["lyapunov"][inside], lam, rtol=1e-6)
    alpha = 0.5 * s
    exp_rgb = round((alpha + (1 - alpha) * 100 / 255) * 255)
    assert np.all(ov["final"][inside][:, :3] == exp_rgb)
    assert np.all(ov["final"][inside][:, 3] == 255)
    assert np.all(ov["final"][~inside] == color[~inside])
    # row stride of the exponents buffer is the height (src/overlay.wgsl:128)
    y, x = 3, 5
    assert ov["exponents"][H * y + x] == pytest.approx(s, rel=1e-5)
    # below the threshold: colour copied
    mapping[..., 0] = 0.05 * xx
    ov = O.overlay(st, color, mapping)
    assert np.array_equal(ov["final"], color)
    # overlay disabled: copy
    ov = O.overlay(O.settings_in(overlay_mode=0), color, mapping)
    assert np.array_equal(ov["final"], color)


def test_sample_rays_identity_are_straight():
    cam = O.camera_in((0, 0, 2), (0, 0, -1), (0, 1, 0))
    st = O.settings_in(mouse_pos=(0.5, 0.5))
    ig = O.integrator_in(samples_per_ray=100, sample_steps=5000)
    out = O.sample_rays(cam, st, ig, FIELDS["identity"], 64, 48, precision="f64")
    W, H = 64, 48
    vh, vw = 0.025, 64 / 48 * 0.025
    offs = [(-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0)]
    for r, (ox, oy) in enumerate(offs):
        px, py = 0.5 + 0.01 * ox, 0.5 + 0.01 * oy
        d = np.array([px * vw - vw / 2, py * vh - vh / 2, -0.035])
        d /= np.linalg.norm(d)
        for k in (0, 1, 57, 99):
            np.testing.assert_allclose(out[r, k, 0, :3], np.array([0, 0, 2.0]) + d * (k * 50 * 0.001), atol=1e-9)
        assert out[r, 0, 1, 3] == 0.5
    np.testing.assert_allclose(out[0, 0, 1, :3], [1, 0, 0])
    np.testing.assert_allclose(out[7, 0, 1, :3], [0.5, 0, 0.5])


def test_prefilter_does_not_change_results():
    v, f = scenes.benchmark_scene(level=1)
    cam = O.camera_in((0.3, 0.2, 3.0), (-0.1, -0.06, -1.0), (0, 1, 0), 60.0)
    res = []
    for on in (False, True):
        O.set_prefilter(on)
        res.append(O.render(cam, O.settings_in(), _ig(h_initial=0.01), FIELDS["warp"], v, f, 24, 16, threads=4))
    O.set_prefilter(True)
    for k in ("hit", "mapping", "color8", "steps"):
        assert np.array_equal(res[0][k], res[1][k]), k
    assert res[1]["tests"] < res[0]["tests"]


def test_rounding_probe_variants():
    """The ulp-nudged FP32 variants move swirl rays (sin/cos/sqrt on every step) by rounding
    amounts only, leave the FP64 instantiation alone and reset after each call."""
    v, f = scenes.benchmark_scene(level=1)
    cam = O.camera_in((0.3, 0.2, 3.0), (-0.1, -0.06, -1.0), (0, 1, 0), 60.0)
    args = (cam, O.settings_in(), _ig(h_initial=0.01), FIELDS["swirl"], v, f, 24, 16)
    base = O.render(*args, threads=4)
    b64 = O.render(*args, threads=4, precision="f64")
    for var in ("f32up", "f32dn"):
        r = O.render(*args, threads=4, precision=var)
        d = np.abs(r["mapping"][..., :3].astype(np.float64) - base["mapping"][..., :3]).max(-1)
        assert (d > 0).mean() > 0.5, var           # the nudge reaches the arithmetic
        assert np.median(d) < 1e-4, var            # and stays at rounding level
        assert np.array_equal(O.render(*args, threads=4)["mapping"], base["mapping"])  # reset
        assert np.array_equal(O.render(*args, threads=4, precision="f64")["mapping"], b64["mapping"])
    up = O.render(*args, threads=4, precision="f32up")["mapping"]
    dn = O.render(*args, threads=4, precision="f32dn")["mapping"]
    assert not np.array_equal(up, dn)
    with pytest.raises(ValueError):
        O.render(*args, precision="f16")


# ---------------------------------------------------------------------------------------
# Outline-ray search (src/main_view.wgsl:381-468), expectations worked out by hand
# ---------------------------------------------------------------------------------------
def _exps(W, H, pixels, value=1.0):
    e = np.zeros(W * H, np.float32)
    for x, y in pixels:
        e[y * H + x] = value          # the overlay's row stride H
    return e


def test_outline_search_sides_and_corners():
    W = H = 10  # square: every index y * H + x names one texel
    A, B, Cc, D = (1, 7), (8, 1), (2, 2), (7, 6)
    got = O.outline_search(_exps(W, H, [A, B, Cc, D]), W, H)
    # corner keys: 0 (W-x)+y -> A 16; 2 x+y -> D 13; 4 x+(H-y) -> B 17; 6 (W-x)+(H-y) -> C 16
    # sides: 1 largest y -> A; 3 largest x -> B; 5 smallest y -> B; 7 smallest x -> A
    want = [A, A, D, B, B, B, Cc, A]
    assert [tuple(int(v) for v in r[:2]) for r in got] == want
    assert (got[:, 2] == 1).all()


def test_outline_search_row_stride_aliasing():
    # W > H: the stride-H index of texel (10, 1) is 19 = 2 * 9 + 1, so the scans also meet it
    # at (1, 2); each search keeps whichever alias its own order prefers
    W, H = 12, 9
    got = O.outline_search(_exps(W, H, [(10, 1)]), W, H)
    a, b = (10, 1), (1, 2)
    # keys: 0 -> b (13 vs 3); 2 -> a (11 vs 3); 4 -> a (18 vs 8); 6 -> b (18 vs 10);
    # 1 largest y -> b; 3 largest x -> a; 5 smallest y -> a; 7 smallest x -> b
    assert [tuple(int(v) for v in r[:2]) for r in got] == [b, b, a, a, a, a, b, b]


def test_outline_search_ties_threshold_and_empty():
    W = H = 12
    # equal key 14 for search 0: (1, 3) and (3, 5); `key >= sum` keeps the later one in x-major order
    got = O.outline_search(_exps(W, H, [(1, 3), (3, 5)]), W, H)
    assert tuple(got[0, :2]) == (3, 5) and tuple(got[7, :2]) == (1, 3)
    # 0.8 is inclusive, just below is not
    assert O.outline_search(_exps(W, H, [(4, 4)], 0.8), W, H)[:, 2].all()
    below = np.nextafter(np.float32(0.8), np.float32(0))
    assert not O.outline_search(_exps(W, H, [(4, 4)], below), W, H)[:, 2].any()
    none = O.outline_search(np.zeros(W * H, np.float32), W, H)
    assert (none == 0).all()


def test_outline_search_row_range_quirks():
    # search 1 runs y from W - 1 (not H - 1) down: with W > H it reads rows y >= H through the
    # stride-H index; entry 16 = 5 * 3 + 1 is only reachable that way
    W, H = 6, 3
    e = np.zeros(W * H, np.float32)
    e[16] = 1.0
    got = O.outline_search(e, W, H)
    assert tuple(got[1]) == (1, 5, 1)
    assert not np.delete(got[:, 2], 1).any()
    # H > W: indices y * H + x past W * H read as 0, so with every entry set, search 0's
    # maximum of (W - x) + y is at x = 0 on the last row whose index stays inside: y = 2
    W, H = 3, 6
    got = O.outline_search(np.ones(W * H, np.float32), W, H)
    assert tuple(got[0]) == (0, 2, 1)


# ---- overlay_web's sRGB encoding (overlay.wgsl:148-165) ----------------------------------------

def _srgb_decode(c):
    """The sRGB decoding function of IEC 61966-2-1 (the inverse of the encoding)."""
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_srgb_encoding_known_values_and_inverse():
    """overlay.wgsl:148-154.  Above the 0.0031308 cutoff it is the sRGB encoding of IEC 61966-2-1:
    its published values (linear 0.5 -> 0.7353570, 0.18 -> 0.4613561, 1 -> 1) and a round trip
    through the standard's decoding, which a wrong exponent or branch would break.  Below the
    cutoff the shader divides by 12.92 where the standard multiplies: kept as the reference
    writes it (0.001 -> 0.001 / 12.92), so the function jumps at the cutoff."""
    assert O.srgb_from_linear(0.0) == 0.0
    assert abs(O.srgb_from_linear(1.0) - 1.0) <= 1e-6
    assert abs(O.srgb_from_linear(0.5) - 0.7353569830524495) <= 2e-6
    assert abs(O.srgb_from_linear(0.18) - 0.46135612950044164) <= 2e-6
    assert O.srgb_from_linear(0.001) == float(np.float32(0.001) / np.float32(12.92))
    below = O.srgb_from_linear(np.nextafter(np.float32(0.0031308), np.float32(0)))
    assert abs(below - 0.0031308 / 12.92) <= 1e-8
    assert abs(O.srgb_from_linear(0.0031308) - 0.040448) <= 1e-5
    xs = np.linspace(0.0031308, 1.0, 1001)
    enc = np.array([O.srgb_from_linear(x) for x in xs])
    np.testing.assert_allclose(_srgb_decode(enc), xs, rtol=2e-5, atol=1e-7)
    assert (np.diff(enc) >= 0).all()


def test_overlay_web_encodes_the_copied_colour():
    """overlay_mode 0 with overlay_web: every texel's rgb is the sRGB encoding of its colour,
    alpha untouched (against the published decoding: within one unorm8 step)."""
    W, H = 16, 16
    rc = np.zeros((H, W, 4), np.uint8)
    rc[..., 0] = np.arange(256, dtype=np.uint8).reshape(H, W)
    rc[..., 1] = 255 - rc[..., 0]
    rc[..., 2] = 7
    rc[..., 3] = np.arange(256, dtype=np.uint8).reshape(H, W)[::-1]
    mapping = np.zeros((H, W, 4), np.float32)
    st = O.settings_in(overlay_mode=0)
    web = O.overlay(st, rc, mapping, web=True)["final"]
    desk = O.overlay(st, rc, mapping)["final"]
    assert np.array_equal(desk, rc)
    assert np.array_equal(web[..., 3], rc[..., 3])
    # decoding the encoded byte lands within one (encoded) step of the original linear value
    enc = web[..., :3].astype(int)
    hi = _srgb_decode(np.minimum(enc + 1, 255) / 255.0)
    lo = _srgb_decode(np.maximum(enc - 1, 0) / 255.0)
    x = rc[..., :3] / 255.0
    assert ((lo <= x + 1e-9) & (x <= hi + 1e-9)).all()
    assert (enc >= rc[..., :3]).all()
