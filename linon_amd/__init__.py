"""linon_amd — MI355X (gfx950) nonlinear ray casting behind linon's MainView interface.

The compute pass of linon (src/main_view.wgsl ray cast + src/overlay.wgsl Lyapunov overlay)
runs as hand-written HIP kernels in ``liblinon_hip.so``, which exports the C ABI declared in
``include/linon.h``.  This module binds that ABI with ctypes and mirrors the reference's
``MainView`` (src/main_view.rs:53-850) method for method, so that host code written against
the reference reads the same here:

    view = MainView(width, height)                       # MainView::new
    view.update_model(verts, faces)                       # MainView::update_model
    view.update_camera(camera)                            # MainView::update_camera
    view.update_settings(Settings(...))                   # MainView::update_settings
    view.reload_shader(field_body)                        # MainView::reload_shader (WGSL body)
    view.render()                                         # MainView::render
    view.render_high_accuracy()                           # MainView::render_high_accuracy

There is no CPU fallback: importing works without a GPU, but creating a view raises when the
library or a HIP device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field as dc_field
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "liblinon_hip.so"

LINON_OK = 0
STATUS_NAMES = {0: "OK", 1: "ARGUMENT", 2: "COMPILE", 3: "HIP", 4: "STATE", 5: "UNSUPPORTED"}

MODE_REFERENCE = 0
MODE_ESCAPE = 1
ACCEL_AUTO, ACCEL_BRUTE, ACCEL_GRID = 0, 1, 2
GRID_BUILD_AUTO, GRID_BUILD_HOST, GRID_BUILD_DEVICE = 0, 1, 2
MATH_AUTO, MATH_IEEE, MATH_WGSL = 0, 1, 2

# Field ids (include/linon.h): 1..9 = PredefinedFunction order of src/functions.rs:2-13.
FIELDS = {
    "custom": 0,
    "mirage_spherical": 1,
    "mirage_plane": 2,
    "mirage_spherical_sigmoid": 3,
    "mirage_plane_sigmoid": 4,
    "translation_x": 5,
    "translation_z": 6,
    "rotation": 7,
    "lorenz": 8,
    "roessler": 9,
    "identity": 16,
    "swirl": 17,
    "lens": 18,
    "hot_air": 19,
    "warp": 20,
}
PREDEFINED_FIELD_IDS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 16, 17, 18, 19, 20]
BENCH_FIELDS = ("swirl", "lens", "hot_air", "warp")


class LinonError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"linon error {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class Camera(C.Structure):
    """CameraUniform (src/main_view.rs:12-19), 112 bytes.  view_direction[3] > 0 is a
    vertical field of view in degrees; 0 keeps the reference's fixed viewport."""
    _fields_ = [("origin", C.c_float * 4), ("view_direction", C.c_float * 4), ("up", C.c_float * 4),
                ("view_matrix", C.c_float * 16)]

    @classmethod
    def make(cls, origin, view_direction, up, fov_deg: float = 0.0) -> "Camera":
        c = cls()
        for i in range(3):
            c.origin[i] = float(origin[i])
            c.view_direction[i] = float(view_direction[i])
            c.up[i] = float(up[i])
        c.view_direction[3] = float(fov_deg)
        for i in range(4):
            c.view_matrix[5 * i] = 1.0
        return c

    @classmethod
    def arcball_default(cls, center) -> "Camera":
        """ArcballCamera::new(center, 1, ..) followed by zoom(-1, 1) (src/main_view.rs:153-155,
        src/arcball.rs:33-46,91-98): eye at center + (0, 0, 2), looking down -z, y up.  The
        eye is computed in f32 as cgmath does (center + 2)."""
        c = np.asarray(center, dtype=np.float32)
        eye = np.array([c[0], c[1], c[2] + np.float32(2.0)], dtype=np.float32)
        return cls.make(eye, (0.0, 0.0, -1.0), (0.0, 1.0, 0.0))

    @classmethod
    def look_at(cls, eye, target, up_hint=(0.0, 1.0, 0.0), fov_deg: float = 60.0) -> "Camera":
        eye = np.asarray(eye, dtype=np.float64)
        d = np.asarray(target, dtype=np.float64) - eye
        d /= np.linalg.norm(d)
        up = np.asarray(up_hint, dtype=np.float64)
        if abs(float(np.dot(up, d))) > 0.999:
            up = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
        right = np.cross(d, up)
        right /= np.linalg.norm(right)
        up = np.cross(right, d)
        return cls.make(eye, d, up, fov_deg)


class Settings(C.Structure):
    """Settings (src/main_view.rs:43-51), 24 bytes."""
    _fields_ = [("field_weight", C.c_float), ("mouse_pos", C.c_float * 2), ("overlay_mode", C.c_int32),
                ("central_difference_delta", C.c_int32), ("lyapunov_scaling", C.c_float)]

    @classmethod
    def make(cls, field_weight=1.0, mouse_pos=(0.5, 0.6), overlay_mode=0, central_difference_delta=1,
             lyapunov_scaling=50.0) -> "Settings":
        s = cls()
        s.field_weight = field_weight
        s.mouse_pos[0], s.mouse_pos[1] = mouse_pos
        s.overlay_mode = overlay_mode
        s.central_difference_delta = central_difference_delta
        s.lyapunov_scaling = lyapunov_scaling
        return s


class Integrator(C.Structure):
    _fields_ = [("h_initial", C.c_float), ("t_max", C.c_float), ("adaptive", C.c_int32),
                ("max_steps", C.c_int32), ("mode", C.c_int32), ("escape_radius", C.c_float),
                ("samples_per_ray", C.c_int32), ("sample_steps", C.c_int32), ("need_mapping", C.c_int32),
                ("outline_rays", C.c_int32), ("clip_exit", C.c_int32), ("pad", C.c_int32)]

    @classmethod
    def make(cls, h_initial=0.1, t_max=5.0, adaptive=True, max_steps=100000, mode=MODE_REFERENCE,
             escape_radius=2.0, samples_per_ray=100, sample_steps=5000, need_mapping=False,
             outline_rays=False, clip_exit=False) -> "Integrator":
        i = cls()
        i.h_initial, i.t_max, i.adaptive, i.max_steps = h_initial, t_max, int(adaptive), max_steps
        i.mode, i.escape_radius = mode, escape_radius
        i.samples_per_ray, i.sample_steps, i.need_mapping = samples_per_ray, sample_steps, int(need_mapping)
        i.outline_rays = int(outline_rays)
        i.clip_exit = int(clip_exit)
        return i


class Counters(C.Structure):
    _fields_ = [("rk4_steps", C.c_uint64), ("triangle_tests", C.c_uint64), ("cell_visits", C.c_uint64),
                ("queries", C.c_uint64), ("skips", C.c_uint64), ("rays", C.c_uint64), ("last_render_ms", C.c_float),
                ("last_raycast_ms", C.c_float)]


class DeviceBuffers(C.Structure):
    _fields_ = [("final_rgba8", C.c_void_p), ("ray_casting_rgba8", C.c_void_p), ("mapping_rgba32f", C.c_void_p),
                ("exponents_f32", C.c_void_p), ("lyapunov_f32", C.c_void_p), ("hit_face_i32", C.c_void_p),
                ("steps_i32", C.c_void_p), ("ray_samples", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class View(C.Structure):
    """One view of a batch (linon_view)."""
    _fields_ = [("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32), ("field", C.c_int32),
                ("h_initial", C.c_float), ("field_weight", C.c_float), ("pad", C.c_int32 * 3)]

    @classmethod
    def make(cls, camera: Camera, width: int, height: int, field, h_initial: float, field_weight: float = 1.0):
        v = cls()
        v.camera = camera
        v.width, v.height = width, height
        v.field = FIELDS[field] if isinstance(field, str) else int(field)
        v.h_initial, v.field_weight = h_initial, field_weight
        return v


class BatchTuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int32), ("waves_per_block", C.c_int32), ("compact_every", C.c_int32),
                ("refill_min", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SurrogateConfig(C.Structure):
    """linon_surrogate_config: the learned flow-map surrogate of one field."""
    _fields_ = [("field", C.c_int32), ("train_rays", C.c_int32), ("batch", C.c_int32), ("steps", C.c_int32),
                ("lr", C.c_float), ("h", C.c_float), ("max_steps", C.c_int32), ("radius", C.c_float),
                ("conf_threshold", C.c_float), ("lyapunov_threshold", C.c_float), ("tolerance", C.c_float),
                ("rejudge_fraction", C.c_float), ("seed", C.c_uint32), ("coarse_h", C.c_float),
                ("pad", C.c_int32 * 2)]

    @classmethod
    def make(cls, field, train_rays=1 << 24, batch=65536, steps=2000, lr=1e-3, h=0.005, max_steps=4000,
             radius=2.0, conf_threshold=0.9, lyapunov_threshold=0.0955, tolerance=1e-3, rejudge_fraction=1 / 64,
             seed=0, coarse_h=0.0):
        c = cls()
        c.field = FIELDS[field] if isinstance(field, str) else int(field)
        c.train_rays, c.batch, c.steps, c.lr, c.h = train_rays, batch, steps, lr, h
        c.max_steps, c.radius, c.conf_threshold = max_steps, radius, conf_threshold
        c.lyapunov_threshold, c.tolerance, c.rejudge_fraction, c.seed = lyapunov_threshold, tolerance, \
            rejudge_fraction, seed
        c.coarse_h = coarse_h
        return c


class SurrogateStats(C.Structure):
    _fields_ = [("train_rays", C.c_uint64), ("train_loss", C.c_float), ("label_seconds", C.c_float),
                ("train_seconds", C.c_float), ("holdout_err_q50", C.c_float), ("holdout_err_q99", C.c_float),
                ("holdout_frac_within_tol", C.c_float), ("frame_rays", C.c_uint64), ("entered", C.c_uint64),
                ("accepted", C.c_uint64), ("conf_fallback", C.c_uint64), ("lyapunov_fallback", C.c_uint64),
                ("rejudged", C.c_uint64), ("rejudge_err_q50", C.c_float), ("rejudge_err_q99", C.c_float),
                ("rejudge_err_max", C.c_float), ("rejudge_frac_within_tol", C.c_float), ("frame_ms", C.c_float),
                ("approach_ms", C.c_float), ("mlp_ms", C.c_float), ("fallback_ms", C.c_float),
                ("overlay_ms", C.c_float), ("rejudge_ms", C.c_float), ("coarse_ms", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


assert C.sizeof(Camera) == 112 and C.sizeof(Settings) == 24 and C.sizeof(Integrator) == 48
assert C.sizeof(View) == 144 and C.sizeof(BatchTuning) == 16 and C.sizeof(SurrogateConfig) == 64

_lib = None

# Every symbol include/linon.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "linon_abi_version": (C.c_int, []),
    "linon_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint32, C.c_uint32]),
    "linon_destroy": (None, [_P]),
    "linon_last_error": (C.c_char_p, [_P]),
    "linon_resize": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "linon_set_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "linon_set_settings": (C.c_int, [_P, C.POINTER(Settings)]),
    "linon_set_mesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32]),
    "linon_set_accel": (C.c_int, [_P, C.c_int, C.c_int]),
    "linon_set_grid_builder": (C.c_int, [_P, C.c_int]),
    "linon_set_math": (C.c_int, [_P, C.c_int]),
    "linon_selftest_division": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64)]),
    "linon_read_grid": (C.c_int, [_P, _P, _P, C.c_size_t, _P, C.c_size_t, _P, C.c_size_t, _P, C.c_size_t, _P]),
    "linon_set_output_srgb": (C.c_int, [_P, C.c_int]),
    "linon_measure_fma_rate": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "linon_selftest_sqrt": (C.c_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "linon_selftest_exp": (C.c_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "linon_set_field_predefined": (C.c_int, [_P, C.c_int]),
    "linon_set_field_wgsl": (C.c_int, [_P, C.c_char_p]),
    "linon_wgsl_translate": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "linon_wgsl_compile_check": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
    "linon_set_integrator": (C.c_int, [_P, C.POINTER(Integrator)]),
    "linon_get_integrator": (C.c_int, [_P, C.POINTER(Integrator)]),
    "linon_render": (C.c_int, [_P, _P]),
    "linon_render_high_accuracy": (C.c_int, [_P, _P]),
    "linon_render_outline_rays": (C.c_int, [_P, _P]),
    "linon_synchronize": (C.c_int, [_P]),
    "linon_read_final": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_ray_casting": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_mapping": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_exponents": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_lyapunov": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_hit_faces": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_steps": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_ray_samples": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_read_outline": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_get_counters": (C.c_int, [_P, C.POINTER(Counters)]),
    "linon_get_device_buffers": (C.c_int, [_P, C.POINTER(DeviceBuffers)]),
    "linon_get_timings": (C.c_int, [_P, _P, _P, C.c_int, C.POINTER(C.c_int)]),
    "linon_set_batch": (C.c_int, [_P, _P, C.c_int]),
    "linon_render_batch": (C.c_int, [_P, _P]),
    "linon_read_batch_view": (C.c_int, [_P, C.c_int, _P, C.c_size_t, _P, C.c_size_t, _P, C.c_size_t, _P,
                                        C.c_size_t]),
    "linon_set_batch_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "linon_get_batch_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "linon_autotune_batch": (C.c_int, [_P, C.c_int, C.POINTER(BatchTuning), C.POINTER(C.c_float)]),
    "linon_surrogate_train": (C.c_int, [_P, C.POINTER(SurrogateConfig), C.POINTER(SurrogateStats)]),
    "linon_render_surrogate": (C.c_int, [_P, _P]),
    "linon_get_surrogate_stats": (C.c_int, [_P, C.POINTER(SurrogateStats)]),
    "linon_selftest_gemm": (C.c_int, [C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P,
                                      C.c_int, _P, _P]),
    "linon_reset_timings": (C.c_int, [_P]),
    "linon_surrogate_save": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "linon_surrogate_load": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_surrogate_infer": (C.c_int, [_P, _P, C.c_int, _P, C.c_int]),
    "linon_read_surrogate_accepted": (C.c_int, [_P, _P, C.c_size_t]),
    "linon_flowmap": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(Integrator), C.c_float, _P, C.c_int, _P, _P]),
    "linon_flowmap_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(Integrator), C.c_float, _P, C.c_int, _P,
                                        _P]),
}


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> C.CDLL:
    """Loads liblinon_hip.so (raises if it has not been built; there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(f"{_LIB_PATH} is missing: build it with `python -m linon_amd.build`")
    lib = C.CDLL(str(_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def wgsl_translate(body: str) -> str:
    """HIP C++ for a WGSL field body (no GPU needed); raises LinonError with the translator's
    line:col message for a body outside the supported WGSL subset."""
    lib = load_library()
    need = C.c_size_t(0)
    buf = C.create_string_buffer(1 << 16)
    st = lib.linon_wgsl_translate(body.encode(), buf, len(buf), C.byref(need))
    if need.value > len(buf):
        buf = C.create_string_buffer(need.value)
        st = lib.linon_wgsl_translate(body.encode(), buf, len(buf), C.byref(need))
    if st != LINON_OK:
        raise LinonError(st, buf.value.decode())
    return buf.value.decode()


def wgsl_compile_check(body: str) -> str:
    """Translates and compiles a field body into the gfx950 ray-cast kernel with hiprtc (no GPU
    needed).  Returns a short summary, raises LinonError with the log on failure."""
    lib = load_library()
    buf = C.create_string_buffer(1 << 20)
    st = lib.linon_wgsl_compile_check(body.encode(), buf, len(buf))
    if st != LINON_OK:
        raise LinonError(st, buf.value.decode(errors="replace"))
    return buf.value.decode()


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


@dataclass
class Frame:
    """Read-back of one render (texel row 0 at the top, as the reference's textures)."""
    final: np.ndarray        # (H, W, 4) uint8 — MainView::texture
    ray_casting: np.ndarray  # (H, W, 4) uint8
    mapping: np.ndarray      # (H, W, 4) float32
    hit_faces: np.ndarray    # (H, W) int32
    steps: np.ndarray        # (H, W) int32
    lyapunov: np.ndarray     # (H, W) float32
    exponents: np.ndarray    # (W*H,) float32, row stride H as src/overlay.wgsl:128
    counters: dict = dc_field(default_factory=dict)


class MainView:
    """Mirror of linon's MainView (src/main_view.rs:53-850) over the C ABI."""

    def __init__(self, width: int, height: int, device: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        st = self._lib.linon_create(C.byref(h), device, width, height)
        if st != LINON_OK:
            raise LinonError(st, f"linon_create({device}, {width}, {height}) failed (is a GPU visible?)")
        self._h = h
        self._num_faces = 0
        self.width, self.height = width, height

    # -- lifecycle ---------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.linon_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int) -> None:
        if st != LINON_OK:
            msg = self._lib.linon_last_error(self._h)
            raise LinonError(st, msg.decode() if msg else "")

    # -- MainView interface ------------------------------------------------------------
    def resize_texture(self, width: int, height: int) -> None:
        self._check(self._lib.linon_resize(self._h, width, height))
        self.width, self.height = width, height

    def update_camera(self, camera: Camera) -> None:
        self._check(self._lib.linon_set_camera(self._h, C.byref(camera)))

    def update_settings(self, settings: Settings) -> None:
        self._check(self._lib.linon_set_settings(self._h, C.byref(settings)))

    def update_model(self, verts, faces) -> None:
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        self._check(self._lib.linon_set_mesh(self._h, _ptr(v), len(v), _ptr(f), len(f)))
        self._num_faces = len(f)

    def set_accel(self, accel: int = ACCEL_AUTO, grid_resolution: int = 0) -> None:
        self._check(self._lib.linon_set_accel(self._h, accel, grid_resolution))

    def set_math(self, math: int = MATH_AUTO) -> None:
        """IEEE, WGSL-accuracy or per-field AUTO arithmetic for the ray-cast kernels."""
        self._check(self._lib.linon_set_math(self._h, math))

    def set_grid_builder(self, builder: int = GRID_BUILD_AUTO) -> None:
        """Host threads or device kernels build the grid (same arrays); AUTO: device above
        65536 faces.  Rebuilds the current mesh."""
        self._check(self._lib.linon_set_grid_builder(self._h, builder))

    def read_grid(self) -> dict:
        """The current grid: dims, cell_start, cell_faces, dist and the packed faces."""
        dims = np.zeros(3, np.int32)
        refs = C.c_uint64(0)
        self._check(self._lib.linon_read_grid(self._h, _ptr(dims), None, 0, None, 0, None, 0, None, 0, C.byref(refs)))
        nc = int(np.prod(dims.astype(np.int64)))
        out = {"dims": dims, "cell_start": np.zeros(nc + 1, np.uint32), "cell_faces": np.zeros(max(1, refs.value), np.uint32),
               "dist": np.zeros(nc, np.uint8), "tris": np.zeros((max(1, self._num_faces), 12), np.float32)}
        self._check(self._lib.linon_read_grid(
            self._h, None, _ptr(out["cell_start"]), out["cell_start"].nbytes, _ptr(out["cell_faces"]),
            out["cell_faces"].nbytes, _ptr(out["dist"]), out["dist"].nbytes, _ptr(out["tris"]), out["tris"].nbytes,
            None))
        out["cell_faces"] = out["cell_faces"][:refs.value]
        out["tris"] = out["tris"][:self._num_faces]
        return out

    def set_output_srgb(self, enable: bool) -> None:
        """overlay_web (sRGB-encoded final colour) instead of overlay_desktop."""
        self._check(self._lib.linon_set_output_srgb(self._h, 1 if enable else 0))

    def set_field(self, field) -> None:
        fid = FIELDS[field] if isinstance(field, str) else int(field)
        self._check(self._lib.linon_set_field_predefined(self._h, fid))

    def reload_shader(self, field_body: str) -> None:
        """MainView::reload_shader(None, body): compile a WGSL field body."""
        self._check(self._lib.linon_set_field_wgsl(self._h, field_body.encode()))

    def set_integrator(self, integ: Integrator) -> None:
        self._check(self._lib.linon_set_integrator(self._h, C.byref(integ)))

    def integrator(self) -> Integrator:
        i = Integrator()
        self._check(self._lib.linon_get_integrator(self._h, C.byref(i)))
        return i

    def render(self, stream: int | None = None) -> None:
        self._check(self._lib.linon_render(self._h, C.c_void_p(stream)))

    def render_high_accuracy(self, stream: int | None = None) -> None:
        self._check(self._lib.linon_render_high_accuracy(self._h, C.c_void_p(stream)))

    def render_outline_rays(self, stream: int | None = None) -> None:
        self._check(self._lib.linon_render_outline_rays(self._h, C.c_void_p(stream)))

    def synchronize(self) -> None:
        self._check(self._lib.linon_synchronize(self._h))

    # -- read-back ----------------------------------------------------------------------
    def _read(self, fn, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        self._check(fn(self._h, _ptr(out), out.nbytes))
        return out

    def read_final(self) -> np.ndarray:
        return self._read(self._lib.linon_read_final, (self.height, self.width, 4), np.uint8)

    def read_ray_casting(self) -> np.ndarray:
        return self._read(self._lib.linon_read_ray_casting, (self.height, self.width, 4), np.uint8)

    def read_mapping(self) -> np.ndarray:
        return self._read(self._lib.linon_read_mapping, (self.height, self.width, 4), np.float32)

    def read_hit_faces(self) -> np.ndarray:
        return self._read(self._lib.linon_read_hit_faces, (self.height, self.width), np.int32)

    def read_steps(self) -> np.ndarray:
        return self._read(self._lib.linon_read_steps, (self.height, self.width), np.int32)

    def read_lyapunov(self) -> np.ndarray:
        return self._read(self._lib.linon_read_lyapunov, (self.height, self.width), np.float32)

    def read_exponents(self) -> np.ndarray:
        return self._read(self._lib.linon_read_exponents, (self.width * self.height,), np.float32)

    def read_ray_samples(self) -> np.ndarray:
        n = self.integrator().samples_per_ray
        return self._read(self._lib.linon_read_ray_samples, (8, max(1, n), 2, 4), np.float32)

    def read_outline(self) -> np.ndarray:
        """(8, 3) int32: texel (x, y) and found flag of each outline-ray search of the last
        render with Integrator.outline_rays set."""
        return self._read(self._lib.linon_read_outline, (8, 3), np.int32)

    def counters(self) -> dict:
        c = Counters()
        self._check(self._lib.linon_get_counters(self._h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    # -- batches of views (one persistent launch per field, captured in a hipGraph) ----------
    def set_batch(self, views) -> None:
        arr = (View * len(views))(*views)
        self._batch = list(views)
        self._check(self._lib.linon_set_batch(self._h, C.cast(arr, C.c_void_p), len(views)))

    def render_batch(self, stream: int | None = None) -> None:
        self._check(self._lib.linon_render_batch(self._h, C.c_void_p(stream)))

    def read_batch_view(self, i: int) -> Frame:
        v = self._batch[i]
        H, W = v.height, v.width
        rgba = np.empty((H, W, 4), np.uint8)
        mapping = np.empty((H, W, 4), np.float32)
        hit = np.empty((H, W), np.int32)
        steps = np.empty((H, W), np.int32)
        self._check(self._lib.linon_read_batch_view(self._h, i, _ptr(rgba), rgba.nbytes, _ptr(mapping), mapping.nbytes,
                                                    _ptr(hit), hit.nbytes, _ptr(steps), steps.nbytes))
        empty = np.zeros(0, np.float32)
        return Frame(rgba, rgba, mapping, hit, steps, empty, empty, {})

    def batch_tuning(self) -> BatchTuning:
        t = BatchTuning()
        self._check(self._lib.linon_get_batch_tuning(self._h, C.byref(t)))
        return t

    def set_batch_tuning(self, waves_per_cu=16, waves_per_block=1, compact_every=32, refill_min=16) -> None:
        t = BatchTuning(waves_per_cu, waves_per_block, compact_every, refill_min)
        self._check(self._lib.linon_set_batch_tuning(self._h, C.byref(t)))

    def autotune_batch(self, downscale: int = 4):
        """Searches the launch shape on a 1/downscale proxy of the batch, confirms the five
        fastest on the batch itself; returns (tuning, the batch's time in ms with it)."""
        t = BatchTuning()
        ms = C.c_float(0)
        self._check(self._lib.linon_autotune_batch(self._h, downscale, C.byref(t), C.byref(ms)))
        return t, ms.value

    # -- learned flow-map surrogate ---------------------------------------------------------
    def train_surrogate(self, cfg: SurrogateConfig) -> dict:
        st = SurrogateStats()
        self._check(self._lib.linon_surrogate_train(self._h, C.byref(cfg), C.byref(st)))
        return st.as_dict()

    def render_surrogate(self, stream: int | None = None) -> None:
        self._check(self._lib.linon_render_surrogate(self._h, C.c_void_p(stream)))

    def surrogate_stats(self) -> dict:
        st = SurrogateStats()
        self._check(self._lib.linon_get_surrogate_stats(self._h, C.byref(st)))
        return st.as_dict()

    def save_surrogate(self, path=None) -> bytes:
        """The trained surrogate as a blob (linon_surrogate_save), also written to `path`."""
        need = C.c_size_t(0)
        self._check(self._lib.linon_surrogate_save(self._h, None, 0, C.byref(need)))
        buf = np.zeros(need.value, np.uint8)
        self._check(self._lib.linon_surrogate_save(self._h, _ptr(buf), buf.nbytes, C.byref(need)))
        data = buf.tobytes()
        if path is not None:
            Path(path).write_bytes(data)
        return data

    def read_surrogate_accepted(self) -> np.ndarray:
        """(H, W) bool: texels whose exit point the last surrogate frame took from the network."""
        m = np.zeros((self.height, self.width), np.uint8)
        self._check(self._lib.linon_read_surrogate_accepted(self._h, _ptr(m), m.nbytes))
        return m.astype(bool)

    def surrogate_infer(self, states, fused: bool = True) -> np.ndarray:
        """Raw network outputs (n, 7) for (n, 6) entry states (linon_surrogate_infer)."""
        x = np.ascontiguousarray(states, np.float32).reshape(-1, 6)
        out = np.zeros((len(x), 7), np.float32)
        self._check(self._lib.linon_surrogate_infer(self._h, _ptr(x), len(x), _ptr(out), 1 if fused else 0))
        return out

    def load_surrogate(self, src) -> None:
        """Installs a saved surrogate (bytes, or a path to them) (linon_surrogate_load)."""
        data = src if isinstance(src, (bytes, bytearray)) else Path(src).read_bytes()
        buf = np.frombuffer(bytes(data), np.uint8).copy()
        self._check(self._lib.linon_surrogate_load(self._h, _ptr(buf), buf.nbytes))

    def timings(self):
        """(raycast_ms, render_ms) arrays of the renders since reset_timings()."""
        a = np.zeros(256, np.float32)
        b = np.zeros(256, np.float32)
        n = C.c_int(0)
        self._check(self._lib.linon_get_timings(self._h, _ptr(a), _ptr(b), 256, C.byref(n)))
        return a[: n.value].copy(), b[: n.value].copy()

    def reset_timings(self) -> None:
        self._check(self._lib.linon_reset_timings(self._h))

    def device_buffers(self) -> DeviceBuffers:
        d = DeviceBuffers()
        self._check(self._lib.linon_get_device_buffers(self._h, C.byref(d)))
        return d

    def read_frame(self) -> Frame:
        return Frame(self.read_final(), self.read_ray_casting(), self.read_mapping(), self.read_hit_faces(),
                     self.read_steps(), self.read_lyapunov(), self.read_exponents(), self.counters())


FLOWMAP_NOALLOC_MAX = 65536   # LINON_FLOWMAP_NOALLOC_MAX


def flowmap(field, integrator: Integrator, states, field_weight: float = 1.0, math: int = 0, device: int = 0,
            noalloc: bool = False):
    """Flow map through the bounding sphere on the GPU (linon_flowmap, or with noalloc
    linon_flowmap_noalloc, which stages in the library's fixed device array and takes at most
    FLOWMAP_NOALLOC_MAX states): (n, 6) entry states (p, q) -> ((n, 6) end point and final
    direction, (n,) RK4 iterations)."""
    lib = load_library()
    fid = FIELDS[field] if isinstance(field, str) else int(field)
    x = np.ascontiguousarray(states, np.float32).reshape(-1, 6)
    out = np.zeros_like(x)
    steps = np.zeros(len(x), np.int32)
    fn = lib.linon_flowmap_noalloc if noalloc else lib.linon_flowmap
    st = fn(device, fid, math, C.byref(integrator), field_weight, _ptr(x), len(x), _ptr(out), _ptr(steps))
    if st != LINON_OK:
        msg = lib.linon_last_error(None) if noalloc else None
        raise LinonError(st, msg.decode() if msg else "linon_flowmap failed")
    return out, steps


def selftest_sqrt(device: int = 0) -> dict:
    """Bit-for-bit check of the IEEE build's square root against the compiler's correctly
    rounded one at every float (on the GPU): the mismatch count and the number of comparisons."""
    lib = load_library()
    m, n = C.c_uint64(0), C.c_uint64(0)
    st = lib.linon_selftest_sqrt(device, C.byref(m), C.byref(n))
    if st != LINON_OK:
        raise LinonError(st, "linon_selftest_sqrt failed")
    return {"mismatches": m.value, "comparisons": n.value}


def selftest_exp(device: int = 0) -> dict:
    """The exponential of hot air's default variant at every float y in [-87, 87] against the
    FP64 exponential (on the GPU): results off by more than WGSL's bound for exp, 3 + 2 |y| ulp,
    and the largest error in ulp."""
    lib = load_library()
    o, n, m = C.c_uint64(0), C.c_uint64(0), C.c_float(0.0)
    st = lib.linon_selftest_exp(device, C.byref(o), C.byref(n), C.byref(m))
    if st != LINON_OK:
        raise LinonError(st, "linon_selftest_exp failed")
    return {"over_bound": o.value, "comparisons": n.value, "max_ulp": m.value}


def measure_fma_rate(device: int = 0) -> float:
    """Measured non-packed FP32 FMA rate of the device in TFLOP/s (v_fma_f32 = 2 FLOP)."""
    lib = load_library()
    t = C.c_double(0.0)
    st = lib.linon_measure_fma_rate(device, C.byref(t))
    if st != LINON_OK:
        raise LinonError(st, "linon_measure_fma_rate failed")
    return float(t.value)


def selftest_division(random_pairs: int = 1 << 28, device: int = 0) -> dict:
    """Bit-for-bit check of the IEEE build's reciprocal-based divisions against IEEE division
    (on the GPU): the shared-denominator division, sigmoid_floor and refraction_in_range, and
    the refracting fields' first-stage form (first_stage(p, ...) against the field at p_prev = p).
    Returns the mismatch counts of the reciprocal and of the quotients, and the number of
    comparisons."""
    lib = load_library()
    r, q, n = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    st = lib.linon_selftest_division(device, random_pairs, C.byref(r), C.byref(q), C.byref(n))
    if st != LINON_OK:
        raise LinonError(st, "linon_selftest_division failed")
    return {"recip_mismatches": r.value, "quotient_mismatches": q.value, "comparisons": n.value}
