"""ctypes wrapper of the CPU oracle (liblinon_oracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check of __graft_entry__.py and the CPU-baseline leg of bench.py use
this module, always as the checker; the product library never loads it.

What pins each oracle function (tests/test_oracle_pins.py unless noted):
  hit_triangle       textbook known answers (centroid, edges/vertices inclusive, parallel,
                     behind, |a| < 1e-7) and an FP64 brute-force comparison with an
                     independent numpy plane/barycentric intersection
  ray_color          brute force over tiny meshes, duplicated faces -> lowest index
  camera ray         closed form: the centre pixel of an odd-sized image looks along view_dir;
                     corner rays against the viewport geometry
  refraction_index   the values at 15/20/30/100/200 degC derived by hand from the formula;
                     refraction: Snell's law (sin ratio), normal incidence, total internal
                     reflection equals a mirror reflection
  fields             each predefined body against its own mathematics at sample points
                     (Lorenz / Roessler derivatives, translations, rotation, swirl, lens,
                     warp, mirage temperatures)
  RK4 loop           identity field -> straight line (linear ray cast); translation fields are
                     polynomial in t so RK4 is exact (closed form); swirl/lens/warp: step
                     halving converges at 4th order towards the FP64 result; FP32 `t += h`
                     step counts against an exact float32 accumulation in numpy
  adaptive step      a hand-made field whose jump triggers the 0.001 redo
  lyapunov           affine mappings give a constant G whose sigma_max equals numpy's SVD
  overlay blend      closed-form alpha blend and unorm8 rounding, exponents row stride H
  sample_rays        the identity field gives straight sample rays at the documented offsets
  outline_search     hand-worked side/corner picks, ties, the inclusive 0.8 threshold, an
                     empty buffer, search 1's y range from W - 1 and out-of-range indices
  f32up / f32dn      rounding probes: swirl rays move by rounding amounts only, the FP64
                     instantiation is untouched, the nudge resets after every call
Parity unpinned: the exact f32 bits of the reference's naga/SPIR-V/driver build (its
transcendentals and FMA contraction), which no file of the reference pins.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# LINON_ORACLE_SANITIZE=1 selects the AddressSanitizer/UndefinedBehaviorSanitizer builds
_SAN = "_san" if os.environ.get("LINON_ORACLE_SANITIZE") == "1" else ""
LIB_PATH = HERE / f"liblinon_oracle{_SAN}.so"


class CameraIn(C.Structure):
    _fields_ = [("origin", C.c_float * 4), ("view_direction", C.c_float * 4), ("up", C.c_float * 4),
                ("view_matrix", C.c_float * 16)]


class SettingsIn(C.Structure):
    _fields_ = [("field_weight", C.c_float), ("mouse_pos", C.c_float * 2), ("overlay_mode", C.c_int32),
                ("central_difference_delta", C.c_int32), ("lyapunov_scaling", C.c_float)]


class IntegratorIn(C.Structure):
    _fields_ = [("h_initial", C.c_float), ("t_max", C.c_float), ("adaptive", C.c_int32),
                ("max_steps", C.c_int32), ("mode", C.c_int32), ("escape_radius", C.c_float),
                ("samples_per_ray", C.c_int32), ("sample_steps", C.c_int32), ("need_mapping", C.c_int32),
                ("outline_rays", C.c_int32), ("clip_exit", C.c_int32), ("pad", C.c_int32)]


_lib = None
_lib_fma = None
LIB_FMA_PATH = HERE / f"liblinon_oracle_fma{_SAN}.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)] + (["sanitize"] if _SAN else []), check=True)
    return LIB_PATH


def lib(fma: bool = False) -> C.CDLL:
    """The oracle library; fma=True gives the build with FMA contraction allowed (a second
    FP32 rounding path of the same restatement)."""
    global _lib, _lib_fma
    if fma:
        if _lib_fma is None:
            if not LIB_FMA_PATH.exists():
                build()
            _lib_fma = _bind(C.CDLL(str(LIB_FMA_PATH)))
        return _lib_fma
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def _bind(_lib: C.CDLL) -> C.CDLL:
    P = C.c_void_p
    _lib.oracle_render_f32.argtypes = [P, P, P, C.c_int, P, C.c_uint32, P, C.c_uint32, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P]
    _lib.oracle_render_f64.argtypes = _lib.oracle_render_f32.argtypes
    _lib.oracle_set_prefilter.argtypes = [C.c_int]
    _lib.oracle_set_prefilter.restype = None
    _lib.oracle_set_nudge.argtypes = [C.c_int]
    _lib.oracle_set_nudge.restype = None
    _lib.oracle_render_pixels_f32.argtypes = [P, P, P, C.c_int, P, C.c_uint32, P, C.c_uint32, C.c_int, C.c_int,
                                              P, C.c_int, C.c_int, P, P, P, P]
    _lib.oracle_render_pixels_f64.argtypes = _lib.oracle_render_pixels_f32.argtypes
    _lib.oracle_overlay_f32.argtypes = [P, C.c_int, C.c_int, P, P, P, P, P]
    _lib.oracle_overlay_f64.argtypes = _lib.oracle_overlay_f32.argtypes
    _lib.oracle_overlay_web_f32.argtypes = _lib.oracle_overlay_f32.argtypes
    _lib.oracle_srgb_from_linear_f32.argtypes = [C.c_float]
    _lib.oracle_srgb_from_linear_f32.restype = C.c_float
    _lib.oracle_sample_rays_f32.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, P]
    _lib.oracle_sample_rays_f64.argtypes = _lib.oracle_sample_rays_f32.argtypes
    _lib.oracle_outline_search.argtypes = [P, C.c_int, C.c_int, P]
    _lib.oracle_outline_search.restype = None
    _lib.oracle_sample_rays_outline_f32.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, P, P]
    _lib.oracle_sample_rays_outline_f64.argtypes = _lib.oracle_sample_rays_outline_f32.argtypes
    _lib.oracle_hit_triangle_f32.argtypes = [P, P, P]
    _lib.oracle_hit_triangle_f32.restype = C.c_float
    _lib.oracle_hit_triangle_f64.argtypes = [P, P, P]
    _lib.oracle_hit_triangle_f64.restype = C.c_double
    _lib.oracle_ray_color_f32.argtypes = [P, C.c_uint32, P, C.c_uint32, P, P, C.c_float, P]
    _lib.oracle_field_f64.argtypes = [C.c_int, P, P, P, P, C.c_double, P]
    _lib.oracle_field_f32.argtypes = [C.c_int, P, P, P, P, C.c_float, P]
    _lib.oracle_refraction_index_f64.argtypes = [C.c_double]
    _lib.oracle_refraction_index_f64.restype = C.c_double
    _lib.oracle_refraction_f64.argtypes = [C.c_double, C.c_double, P, P, P]
    _lib.oracle_lyapunov_f64.argtypes = [P, P, P, P]
    _lib.oracle_lyapunov_f64.restype = C.c_double
    _lib.oracle_camera_ray_f64.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, P]
    _lib.oracle_trace_f64.argtypes = [C.c_int, P, C.c_double, P, C.c_uint32, P, C.c_uint32, P, P, P, P]
    _lib.oracle_trace_f32.argtypes = [C.c_int, P, C.c_float, P, C.c_uint32, P, C.c_uint32, P, P, P, P]
    _lib.oracle_flowmap_f32.argtypes = [C.c_int, P, C.c_float, P, P, P, C.c_int, C.c_int]
    _lib.oracle_flowmap_f64.argtypes = [C.c_int, P, C.c_double, P, P, P, C.c_int, C.c_int]
    return _lib


def set_prefilter(on: bool) -> None:
    """Toggles the exact bounding-box prefilter of the all-faces loop (results unchanged)."""
    lib().oracle_set_prefilter(1 if on else 0)


# FP32 variants of one restatement: "f32" (one rounding per WGSL operation, the parity
# oracle), "f32fma" (FMA contraction allowed), "f32up" / "f32dn" (every exp, sin, cos, sqrt and
# division of the camera, field and step arithmetic moved one ulp up / down).  Their spread around the "f64"
# result measures how far FP32 rounding alone moves a ray.
FP32_VARIANTS = ("f32", "f32fma", "f32up", "f32dn")


class _variant:
    """Context manager: the oracle library of a precision name, with the ulp nudge set for
    the duration of one call."""

    def __init__(self, precision: str):
        if precision not in FP32_VARIANTS + ("f64",):
            raise ValueError(f"unknown precision {precision!r}")
        self.lib = lib(fma=precision == "f32fma")
        self.nudge = {"f32up": 1, "f32dn": -1}.get(precision, 0)

    def __enter__(self):
        self.lib.oracle_set_nudge(self.nudge)
        return self.lib

    def __exit__(self, *exc):
        self.lib.oracle_set_nudge(0)
        return False


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def camera_in(origin, view_direction, up, fov_deg=0.0) -> CameraIn:
    c = CameraIn()
    for i in range(3):
        c.origin[i], c.view_direction[i], c.up[i] = origin[i], view_direction[i], up[i]
    c.view_direction[3] = fov_deg
    return c


def settings_in(field_weight=1.0, mouse_pos=(0.5, 0.6), overlay_mode=0, delta=1, scaling=50.0) -> SettingsIn:
    s = SettingsIn()
    s.field_weight = field_weight
    s.mouse_pos[0], s.mouse_pos[1] = mouse_pos
    s.overlay_mode, s.central_difference_delta, s.lyapunov_scaling = overlay_mode, delta, scaling
    return s


def integrator_in(h_initial=0.1, t_max=5.0, adaptive=True, max_steps=100000, mode=0, escape_radius=2.0,
                  samples_per_ray=100, sample_steps=5000, need_mapping=False, clip_exit=False) -> IntegratorIn:
    i = IntegratorIn()
    i.h_initial, i.t_max, i.adaptive, i.max_steps, i.mode = h_initial, t_max, int(adaptive), max_steps, mode
    i.escape_radius, i.samples_per_ray, i.sample_steps = escape_radius, samples_per_ray, sample_steps
    i.need_mapping = int(need_mapping)
    i.clip_exit = int(clip_exit)
    return i


def as_oracle_struct(obj, cls):
    """Copies a layout-identical ctypes struct (e.g. linon_amd.Camera) into the oracle's type."""
    out = cls()
    C.memmove(C.byref(out), C.byref(obj), C.sizeof(cls))
    return out


def render(camera, settings, integ, field: int, verts, faces, W: int, H: int, rows=None, threads: int = 0,
           precision: str = "f32") -> dict:
    """Ray-cast pass; returns texel-space arrays (row 0 at the top).  ``rows`` = (gy0, gy1)
    restricts the invocation rows (gy counts from the image bottom)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    ft = np.float64 if precision == "f64" else np.float32
    color8 = np.zeros((H, W, 4), np.uint8)
    color = np.zeros((H, W, 4), ft)
    mapping = np.zeros((H, W, 4), ft)
    hit = np.full((H, W), -2, np.int32)
    steps = np.zeros((H, W), np.int32)
    counters = np.zeros(2, np.int64)
    flags = np.zeros((H, W), np.uint8)
    r0, r1 = rows if rows is not None else (0, H)
    with _variant(precision) as L:
        fn = L.oracle_render_f64 if precision == "f64" else L.oracle_render_f32
        rc = fn(C.byref(camera), C.byref(settings), C.byref(integ), field, _p(v), len(v), _p(f), len(f), W, H, r0,
                r1, threads, _p(color8), _p(color), _p(mapping), _p(hit), _p(steps), _p(counters), _p(flags))
    if rc != 0:
        raise ValueError(f"oracle_render failed: {rc}")
    return {"color8": color8, "color": color, "mapping": mapping, "hit": hit, "steps": steps,
            "noise_flag": flags, "rk4_steps": int(counters[0]), "tests": int(counters[1])}


def render_pixels(camera, settings, integ, field: int, verts, faces, W: int, H: int, pixels, threads: int = 0,
                  precision: str = "f32"):
    """Rays of an explicit (n, 2) list of invocation coordinates (gx, gy)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    px = np.ascontiguousarray(pixels, dtype=np.int32).reshape(-1, 2)
    n = len(px)
    mapping = np.zeros((n, 4), np.float64 if precision == "f64" else np.float32)
    hit = np.zeros(n, np.int32)
    steps = np.zeros(n, np.int32)
    counters = np.zeros(2, np.int64)
    with _variant(precision) as L:
        fn = L.oracle_render_pixels_f64 if precision == "f64" else L.oracle_render_pixels_f32
        rc = fn(C.byref(camera), C.byref(settings), C.byref(integ), field, _p(v), len(v), _p(f),
                len(f), W, H, _p(px), n, threads, _p(mapping), _p(hit), _p(steps), _p(counters))
    if rc != 0:
        raise ValueError(f"oracle_render_pixels failed: {rc}")
    return {"mapping": mapping, "hit": hit, "steps": steps, "rk4_steps": int(counters[0]), "tests": int(counters[1])}


def flowmap(fid, integ, states, field_weight=1.0, threads=0, precision="f32"):
    """Escape-mode flow map through the bounding sphere without a mesh: (n, 6) entry states
    (p, q) -> (n, 6) exit states (p_exit, dir_exit) and the RK4 iteration counts."""
    ft = np.float64 if precision == "f64" else np.float32
    x = np.ascontiguousarray(states, dtype=ft).reshape(-1, 6)
    out = np.zeros_like(x)
    steps = np.zeros(len(x), np.int32)
    with _variant(precision) as L:
        fn = L.oracle_flowmap_f64 if precision == "f64" else L.oracle_flowmap_f32
        rc = fn(fid, C.byref(integ), field_weight, _p(x), _p(out), _p(steps), len(x), threads)
    if rc != 0:
        raise ValueError(f"oracle_flowmap failed: {rc}")
    return out, steps


def threads_available() -> int:
    return int(lib().oracle_max_threads())


def srgb_from_linear(x: float) -> float:
    """overlay.wgsl:148-154 in f32 (one channel)."""
    return float(lib().oracle_srgb_from_linear_f32(float(x)))


def overlay(settings, ray_casting8, mapping, precision: str = "f32", web: bool = False) -> dict:
    """overlay_desktop, or with web=True overlay_web (sRGB-encoded rgb; f32 only)."""
    L = lib()
    if web and precision != "f32":
        raise ValueError("overlay_web is restated in f32 only")
    H, W = ray_casting8.shape[:2]
    rc8 = np.ascontiguousarray(ray_casting8, dtype=np.uint8)
    ft = np.float32 if precision == "f32" else np.float64
    m = np.ascontiguousarray(mapping, dtype=ft)
    final = np.zeros((H, W, 4), np.uint8)
    expo = np.zeros(W * H, np.float32)
    lyap = np.zeros((H, W), np.float32)
    fn = (L.oracle_overlay_web_f32 if web else L.oracle_overlay_f32) if precision == "f32" else L.oracle_overlay_f64
    rc = fn(C.byref(settings), W, H, _p(rc8), _p(m), _p(final), _p(expo), _p(lyap))
    if rc != 0:
        raise ValueError(f"oracle_overlay failed: {rc}")
    return {"final": final, "exponents": expo, "lyapunov": lyap}


def sample_rays(camera, settings, integ, field: int, W: int, H: int, precision: str = "f32",
                outline=None) -> np.ndarray:
    """The 8 sample rays; ``outline`` = (8, 3) int32 {x, y, found} from outline_search starts
    them at those pixels instead of around the mouse position."""
    L = lib()
    ft = np.float32 if precision == "f32" else np.float64
    out = np.zeros((8, integ.samples_per_ray, 2, 4), ft)
    if outline is None:
        fn = L.oracle_sample_rays_f32 if precision == "f32" else L.oracle_sample_rays_f64
        rc = fn(C.byref(camera), C.byref(settings), C.byref(integ), field, W, H, _p(out))
    else:
        xy = np.ascontiguousarray(outline, dtype=np.int32).reshape(8, 3)
        fn = L.oracle_sample_rays_outline_f32 if precision == "f32" else L.oracle_sample_rays_outline_f64
        rc = fn(C.byref(camera), C.byref(settings), C.byref(integ), field, W, H, _p(xy), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_sample_rays failed: {rc}")
    return out


def outline_search(exponents, W: int, H: int) -> np.ndarray:
    """The outline-ray searches over an exponents buffer (W * H floats, index y * H + x):
    (8, 3) int32 {x, y, found}."""
    e = np.ascontiguousarray(exponents, dtype=np.float32).reshape(-1)
    if e.size != W * H:
        raise ValueError("exponents must hold W * H values")
    out = np.zeros((8, 3), np.int32)
    lib().oracle_outline_search(_p(e), W, H, _p(out))
    return out


def hit_triangle(tri, origin, direction, precision: str = "f64") -> float:
    L = lib()
    ft = np.float64 if precision == "f64" else np.float32
    t = np.ascontiguousarray(tri, dtype=ft).reshape(9)
    o = np.ascontiguousarray(origin, dtype=ft)
    d = np.ascontiguousarray(direction, dtype=ft)
    fn = L.oracle_hit_triangle_f64 if precision == "f64" else L.oracle_hit_triangle_f32
    return float(fn(_p(t), _p(o), _p(d)))


def ray_color(verts, faces, origin, direction, max_dist):
    L = lib()
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    o = np.ascontiguousarray(origin, dtype=np.float32)
    d = np.ascontiguousarray(direction, dtype=np.float32)
    rgba = np.zeros(4, np.float32)
    face = L.oracle_ray_color_f32(_p(v), len(v), _p(f), len(f), _p(o), _p(d), C.c_float(max_dist), _p(rgba))
    return face, rgba


def field(fid: int, p_prev, p, v0, v, t, precision: str = "f64") -> np.ndarray:
    L = lib()
    ft = np.float64 if precision == "f64" else np.float32
    a = [np.ascontiguousarray(x, dtype=ft) for x in (p_prev, p, v0, v)]
    out = np.zeros(3, ft)
    fn = L.oracle_field_f64 if precision == "f64" else L.oracle_field_f32
    rc = fn(fid, *map(_p, a), t, _p(out))
    if rc != 0:
        raise ValueError(f"unknown field {fid}")
    return out


def refraction_index(t: float) -> float:
    return float(lib().oracle_refraction_index_f64(t))


def refraction(t_in, t_out, v, n) -> np.ndarray:
    out = np.zeros(3)
    vv = np.ascontiguousarray(v, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    lib().oracle_refraction_f64(t_in, t_out, _p(vv), _p(nn), _p(out))
    return out


def lyapunov(xn, xp, yn, yp) -> float:
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (xn, xp, yn, yp)]
    return float(lib().oracle_lyapunov_f64(*map(_p, a)))


def camera_ray(camera, W, H, gx, gy) -> np.ndarray:
    out = np.zeros(3)
    lib().oracle_camera_ray_f64(C.byref(camera), W, H, gx, gy, _p(out))
    return out


def trace(fid, integ, origin, direction, verts=None, faces=None, field_weight=1.0, precision="f64"):
    """One ray through the integrator: returns (iterations, end point, face)."""
    L = lib()
    v = np.zeros((0, 3), np.float32) if verts is None else np.ascontiguousarray(verts, dtype=np.float32)
    f = np.zeros((0, 3), np.uint32) if faces is None else np.ascontiguousarray(faces, dtype=np.uint32)
    ft = np.float64 if precision == "f64" else np.float32
    o = np.ascontiguousarray(origin, dtype=ft)
    d = np.ascontiguousarray(direction, dtype=ft)
    end = np.zeros(3, ft)
    face = C.c_int32(-2)
    fn = L.oracle_trace_f64 if precision == "f64" else L.oracle_trace_f32
    n = fn(fid, C.byref(integ), field_weight, _p(v), len(v), _p(f), len(f), _p(o), _p(d), _p(end), C.byref(face))
    return n, end, face.value
