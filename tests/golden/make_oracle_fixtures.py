"""Writes oracle_fixtures.npz: small golden cases of the ray-cast pass, inputs and outputs of
the CPU oracle (the FP32 restatement, its rounding probes: the FMA-contracted build and the
builds with transcendentals and divisions one ulp up / down, and the FP64 instantiation).

Each case stores its inputs in full (camera / settings / integrator structs as raw bytes, the
mesh arrays, field id, image size), so the GPU tests replay exactly these inputs, and the
oracle's outputs: hit face, RK4 steps, colour (rgba8), end point (mapping), the FP32 overlay's
lambda and the adaptive-step noise flag.  Regenerate with

    python tests/golden/make_oracle_fixtures.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import linon_amd as L  # noqa: E402
from linon_amd import configs, scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).with_name("oracle_fixtures.npz")


def raw(struct) -> np.ndarray:
    return np.frombuffer(C.string_at(C.addressof(struct), C.sizeof(struct)), np.uint8).copy()


def cases():
    v, f, c = scenes.reference_startup_scene()
    # the start-up scene and field; a fixed step (the adaptive rule's decisions on this scene are
    # rounding noise, so those pixels would not be gated)
    yield ("startup_mirage", L.Camera.arcball_default(c), L.Settings.make(overlay_mode=1),
           L.Integrator.make(h_initial=0.02, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0),
           L.FIELDS["mirage_spherical_sigmoid"], v, f, 48, 36)
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    for name in ("swirl", "lens", "warp", "hot_air"):
        yield (f"cornell_{name}", wl.camera(), wl.settings(),
               L.Integrator.make(h_initial=0.01, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0),
               L.FIELDS[name], v, f, 32, 32)
    v, f = scenes.benchmark_scene(level=1)
    w2 = configs.OVERLAY_1080P
    yield ("bench1_rotation", w2.camera(), w2.settings(),
           L.Integrator.make(h_initial=0.02, adaptive=False, max_steps=1000, samples_per_ray=0, sample_steps=0),
           L.FIELDS["rotation"], v, f, 24, 16)


def main() -> None:
    data = {}
    names = []
    for name, cam, st, ig, fid, v, f, W, H in cases():
        c, s, i = (O.as_oracle_struct(cam, O.CameraIn), O.as_oracle_struct(st, O.SettingsIn),
                   O.as_oracle_struct(ig, O.IntegratorIn))
        r32 = O.render(c, s, i, fid, v, f, W, H, threads=8)
        r64 = O.render(c, s, i, fid, v, f, W, H, threads=8, precision="f64")
        probes = {var: O.render(c, s, i, fid, v, f, W, H, threads=8, precision=var)
                  for var in ("f32fma", "f32up", "f32dn")}
        ov = O.overlay(O.settings_in(overlay_mode=1, delta=1), r32["color8"], r32["mapping"])
        names.append(name)
        p = name + "/"
        data[p + "camera"], data[p + "settings"], data[p + "integrator"] = raw(cam), raw(st), raw(ig)
        data[p + "field"] = np.array([fid], np.int32)
        data[p + "size"] = np.array([W, H], np.int32)
        data[p + "verts"] = np.asarray(v, np.float32)
        data[p + "faces"] = np.asarray(f, np.uint32)
        data[p + "hit"] = r32["hit"]
        data[p + "steps"] = r32["steps"]
        data[p + "color8"] = r32["color8"]
        data[p + "mapping_f32"] = r32["mapping"]
        for var, rv in probes.items():
            data[p + "mapping_" + var] = rv["mapping"]
            data[p + "hit_" + var] = rv["hit"]
        data[p + "mapping_f64"] = r64["mapping"]
        data[p + "lyapunov_f32"] = ov["lyapunov"]
        data[p + "noise_flag"] = r32["noise_flag"]
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(OUT, OUT.stat().st_size, "bytes,", len(names), "cases")


if __name__ == "__main__":
    main()
