"""The benchmark workloads as host-side set-ups of a MainView.

* ``PLAIN_256_SWIRL``: 256x256, 1 ray/pixel, 60-degree pinhole at distance 3 looking at the origin,
  the Cornell box (centred), swirl field, h = 0.01, at most 1000 RK4 iterations, no overlay.
* ``OVERLAY_1080P``: 1920x1080, the Cornell box plus a seeded 20,480-face displaced icosphere,
  fields swirl / lens / hot-air / warp each rendered once, h = 0.005, at most 4000 RK4
  iterations, Lyapunov overlay over 3x3 neighbours (delta 1, scaling 50).
* ``STRESS_4K``: the stress case, 3840x2160 at h = 0.001 against 1.3 M faces (lens, hot-air).

Both keep the reference's integration loop `for (; t <= 5.0;)` with `t = t + h` in f32
(src/main_view.wgsl:244,265); the step cap never binds at these h.  They integrate with the
fixed step h: the reference's adaptive rule (src/main_view.wgsl:251-255,267-269) compares
|V - last_V| against 10x its previous value, and wherever the field is uniform both sides are
rounding noise, so which steps it refines is decided by the last bits of the arithmetic.
The adaptive path is exercised by its own tests (interactive default h = 0.1).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import FIELDS, MODE_REFERENCE, Camera, Integrator, Settings
from . import scenes


@dataclass
class Workload:
    name: str
    width: int
    height: int
    fields: tuple
    h: float
    max_steps: int
    overlay: bool
    level: int | None  # icosphere subdivision level (None: Cornell box only)
    t_max: float = 5.0
    adaptive: bool = False   # fixed-step RK4 ("step h = ..."); see the module docstring
    samples: bool = True
    samples_per_ray: int = 100
    outline_rays: bool = False  # sample rays start at the outermost high-lambda pixels

    def camera(self) -> Camera:
        return Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov_deg=60.0)

    def settings(self) -> Settings:
        return Settings.make(field_weight=1.0, mouse_pos=(0.5, 0.6), overlay_mode=1 if self.overlay else 0,
                             central_difference_delta=1, lyapunov_scaling=50.0)

    def integrator(self) -> Integrator:
        return Integrator.make(h_initial=self.h, t_max=self.t_max, adaptive=self.adaptive,
                               max_steps=self.max_steps, mode=MODE_REFERENCE,
                               samples_per_ray=self.samples_per_ray if self.samples else 0,
                               sample_steps=5000 if self.samples else 0, outline_rays=self.outline_rays)

    def mesh(self):
        if self.level is None:
            return scenes.cornell_box_centered()
        return scenes.benchmark_scene(self.level)

    def field_ids(self):
        return [FIELDS[f] for f in self.fields]


PLAIN_256_SWIRL = Workload("plain-rk4-256-swirl-cornell", 256, 256, ("swirl",), 0.01, 1000, False, None)
OVERLAY_1080P = Workload("plain-rk4-overlay-1080p-4fields-cornell+ico20k", 1920, 1080,
                   ("swirl", "lens", "hot_air", "warp"), 0.005, 4000, True, 5)
# Stress: 3840x2160, h = 0.001 (5001 RK4 iterations per ray under t <= 5), the Cornell box plus
# the level-8 displaced icosphere (1,310,720 faces), lens and hot-air, overlay, 8 outline rays
# (started at the outermost pixels the previous frame's overlay highlighted) recording 64
# positions each.
STRESS_4K = Workload("stress-4k-h0.001-lens+hotair-cornell+ico1.3M", 3840, 2160, ("lens", "hot_air"), 0.001,
                   20000, True, 8, samples_per_ray=64, outline_rays=True)


def setup_view(view, wl: Workload, field) -> None:
    view.update_camera(wl.camera())
    view.update_settings(wl.settings())
    view.set_integrator(wl.integrator())
    view.set_field(field)


# The four benchmark fields as WGSL field bodies (what a linon user would type into the
# field editor); linon_set_field_wgsl compiles them to the same arithmetic as the predefined
# functors LINON_FIELD_SWIRL / LENS / HOT_AIR / WARP.
FIELD_WGSL = {
    "swirl": "return rotateY(v0, 0.5 * length(p));",
    "lens": "let r = length(p);\nreturn v - 0.3 * p / max(r * r * r, 0.001);",
    "hot_air": (
        "let c = vec3<f32>(0.0, -0.5, 0.0);\n"
        "let temp_prev = 20.0 + 80.0 * sigmoid(8.0 * (0.5 - length(p_prev - c)));\n"
        "let temp = 20.0 + 80.0 * sigmoid(8.0 * (0.5 - length(p - c)));\n"
        "return refraction(temp_prev, temp, v, normalize(p - c));"),
    "warp": "return v0 + 0.3 * vec3<f32>(sin(4.0 * p.y), 0.0, sin(4.0 * p.x));",
}


# ---------------------------------------------------------------------------------------------
# Batched views: 64 views in one persistent launch per field (escape mode, uneven path lengths)
# ---------------------------------------------------------------------------------------------
BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080), (2048, 2048))
BATCH_STEPS = (0.01, 0.005, 0.0025)


def batch_views(num_views: int = 64, seed: int = 0, fov_deg: float = 60.0, enhance: bool = True,
                scale: int = 1):
    """Views drawn from numpy's PCG64(seed): resolution uniform over BATCH_RESOLUTIONS (divided
    by `scale`), eye uniform on a sphere of radius U(2.5, 4) looking at the origin, field
    uniform over the four benchmark fields, h uniform over BATCH_STEPS; plus one 'Enhance' view
    (a copy of view 0 at h/4) when `enhance`."""
    import numpy as np
    from . import View

    rng = np.random.default_rng(seed)
    views = []
    for _ in range(num_views):
        w, h = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = d * rng.uniform(2.5, 4.0)
        field = ("swirl", "lens", "hot_air", "warp")[int(rng.integers(4))]
        step = BATCH_STEPS[int(rng.integers(len(BATCH_STEPS)))]
        cam = Camera.look_at(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov_deg)
        views.append(View.make(cam, max(8, w // scale), max(8, h // scale), field, step))
    if enhance:
        v0 = views[0]
        views.append(View.make(v0.camera, v0.width, v0.height, v0.field, v0.h_initial / 4.0))
    return views


def batch_integrator(max_steps: int = 8000) -> Integrator:
    """Escape mode: a ray stops when it leaves the radius-2 sphere moving outward, at its first
    hit (no overlay in a batch), or after max_steps RK4 iterations."""
    return Integrator.make(h_initial=0.01, t_max=1e30, adaptive=False, max_steps=max_steps, mode=1,
                           escape_radius=2.0, samples_per_ray=0, sample_steps=0, need_mapping=False)


# ---------------------------------------------------------------------------------------------
# Learned flow-map surrogate (config 4): one per field at the workload's step h
# ---------------------------------------------------------------------------------------------
# coarse_h of the multi-fidelity form per field: the network learns the residual of the flow map
# at h over the flow map at coarse_h.  Chosen from the coarse map's own error against h = 0.005
# (max over x, y, z of |d| / max(1, |exit|), 2,000 uniform entry states, FP32 oracle): swirl at
# 0.1 is within 1e-3 on all of them (q99 5e-4; at 0.2 on 71%, and the network's correction left
# 6% of the accepted camera rays of the 1080p frame outside the tolerance), warp at 0.05 on 97%
# (q50 1e-4), hot air at 0.02 on 89%; lens rays are mostly captured by its singularity at any
# step, so it keeps the plain network (coarse_h 0), whose rays fall back to RK4.
SURROGATE_COARSE_H = {"swirl": 0.1, "warp": 0.05, "hot_air": 0.02, "lens": 0.0}
SURROGATE_STEPS = 4000
SURROGATE_DIR = __import__("pathlib").Path(__file__).resolve().parent / "surrogates"


def surrogate_config(field: str, wl: Workload = OVERLAY_1080P, steps: int = SURROGATE_STEPS, seed: int = 0):
    from . import SurrogateConfig
    return SurrogateConfig.make(field, train_rays=1 << 24, batch=65536, steps=steps, lr=1e-3, h=wl.h,
                                max_steps=wl.max_steps, tolerance=1e-3, conf_threshold=0.9,
                                coarse_h=SURROGATE_COARSE_H.get(field, 0.0), seed=seed)


def surrogate_blob_name(field: str, wl: Workload = OVERLAY_1080P) -> str:
    return f"{field}_h{wl.h:g}_max{wl.max_steps}_coarse{SURROGATE_COARSE_H.get(field, 0.0):g}.lnsg"


def saved_surrogate(field: str, wl: Workload = OVERLAY_1080P):
    """Path of the saved weights for this field and workload step, or None."""
    p = SURROGATE_DIR / surrogate_blob_name(field, wl)
    return p if p.exists() else None
