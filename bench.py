"""Benchmark: RK4 ray-steps/s and 1080p frame time of the HIP ray-cast pass on one MI355X.

One step = one 1920x1080 frame for each of the four fields (swirl, lens, hot-air, warp): the
ray-cast kernel (RK4 through the field, per-step segment queries against the Cornell box plus
a 20,480-face displaced icosphere) followed by the Lyapunov overlay, exactly the pass linon
dispatches in MainView::render (src/main_view.rs:831-849).  Inputs are resident on the
device before the timed region; nothing is read back inside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  `roofline` is for the dominant kernel (the ray-cast instantiation with
the most device time): FP32 FLOPs counted with the convention below from the kernel's own
step / triangle-test counters, divided by its mean launch time from HIP events recorded on
the launch stream over the timed region.  `cpu_baseline` times the FP32 CPU restatement
(brute force over all faces, as the reference shader does) on a strided pixel sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import linon_amd as L  # noqa: E402
from linon_amd import configs  # noqa: E402

# FP32 FLOP convention: +, -, * = 1, FMA = 2; div/sqrt/transcendentals are not counted.
# Integration 56 FLOP per RK4 step, +8 while the ray has not hit anything (normalise + length
# of the segment); the four field evaluations per step: swirl 48, lens 64, hot-air 217
# (temperature at p_prev once per step), warp 24; one Moeller-Trumbore test with stored
# edges 44.
FLOP_INTEGRATION = 56
FLOP_PRE_HIT = 8
FLOP_FIELD4 = {"swirl": 48, "lens": 64, "hot_air": 217, "warp": 24}
FLOP_TRIANGLE = 44
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _json_scalar(o):
    """numpy scalars (e.g. float32 timings) as plain numbers."""
    if hasattr(o, "item"):
        return o.item()
    raise TypeError(f"not JSON serialisable: {type(o).__name__}")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_flops(field: str, c: dict) -> float:
    pre = c["queries"] + c["skips"]
    return (c["rk4_steps"] * (FLOP_INTEGRATION + FLOP_FIELD4[field]) + pre * FLOP_PRE_HIT
            + c["triangle_tests"] * FLOP_TRIANGLE)


def _rel_err(a, b) -> np.ndarray:
    """Per-pixel exit-point error max_i |a_i - b_i| / max(1, |b_i|) over x, y, z."""
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    return (np.abs(a - b) / np.maximum(1.0, np.abs(b))).max(axis=1)


def _quantiles(e: np.ndarray) -> dict:
    return {"q50": float(np.quantile(e, 0.5)), "q99": float(np.quantile(e, 0.99)), "max": float(e.max()),
            "frac_within_1e-4": float((e <= 1e-4).mean())}


def frame_accuracy(wl, verts, faces, frames: dict, sample: np.ndarray, f32: dict, n64: int, threads: int) -> dict:
    """The GPU's full-size frames at the sampled pixels against the FP32 oracle, and both against
    the FP64 oracle on the first n64 of them.  `frames[field]` = (mapping, hit faces, lambda)
    read back from the GPU; `f32[field]` = the FP32 oracle's render of `sample`.  Pixels whose
    lambda is below 0.0955 (where the overlay highlights nothing) are reported on their own."""
    from oracle import oracle as O

    cam = O.as_oracle_struct(wl.camera(), O.CameraIn)
    st = O.as_oracle_struct(wl.settings(), O.SettingsIn)
    ig = O.as_oracle_struct(wl.integrator(), O.IntegratorIn)
    W, H = wl.width, wl.height
    rows, cols = H - 1 - sample[:, 1], sample[:, 0]   # invocation (gx, gy) -> texel (gx, H-1-gy)
    sub = sample[:n64]
    out = {"pixels": int(len(sample)), "pixels_f64": int(len(sub)),
           "error": "max over x, y, z of |exit - exit_ref| / max(1, |exit_ref|)"}
    for name in wl.fields:
        mapping, hit, lam = frames[name]
        g_map, g_hit, g_lam = mapping[rows, cols], hit[rows, cols], lam[rows, cols]
        ref = f32[name]
        low = (g_lam >= 0) & (g_lam < 0.0955)
        e32 = _rel_err(g_map, ref["mapping"])
        r64 = O.render_pixels(cam, st, ig, L.FIELDS[name], verts, faces, W, H, sub, threads, precision="f64")
        out[name] = {
            "gpu_vs_f32_oracle": {"hit_agree": float((g_hit == ref["hit"]).mean()),
                                  "hit_agree_low_lambda": float((g_hit[low] == ref["hit"][low]).mean()) if low.any() else None,
                                  "low_lambda_pixels": int(low.sum()),
                                  **_quantiles(e32),
                                  "low_lambda": _quantiles(e32[low]) if low.any() else None},
            "vs_f64_oracle": {"gpu": _quantiles(_rel_err(g_map[:n64], r64["mapping"])),
                              "f32_oracle": _quantiles(_rel_err(ref["mapping"][:n64], r64["mapping"]))},
        }
    return out


def sample_accuracy(camera, settings, integ, field: str, verts, faces, W: int, H: int, mapping, hits, n: int,
                    seed: int = 0) -> dict:
    """A GPU frame (texel-space mapping and hit faces) at n random pixels against the FP32
    oracle, and both against the FP64 oracle: hit agreement and the exit-point error (max over
    x, y, z of |exit - ref| / max(1, |ref|)).  Runs outside every timed region."""
    from oracle import oracle as O

    threads = max(1, min(16, O.threads_available()))
    c = O.as_oracle_struct(camera, O.CameraIn)
    s = O.as_oracle_struct(settings, O.SettingsIn)
    i = O.as_oracle_struct(integ, O.IntegratorIn)
    rng = np.random.default_rng(seed)
    px = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], 1).astype(np.int32)
    rows, cols = H - 1 - px[:, 1], px[:, 0]
    g_map, g_hit = mapping[rows, cols], hits[rows, cols]
    r32 = O.render_pixels(c, s, i, L.FIELDS[field], verts, faces, W, H, px, threads)
    r64 = O.render_pixels(c, s, i, L.FIELDS[field], verts, faces, W, H, px, threads, precision="f64")
    return {"pixels": n, "hit_agree_vs_f32_oracle": float((g_hit == r32["hit"]).mean()),
            "gpu_vs_f32_oracle": _quantiles(_rel_err(g_map, r32["mapping"])),
            "vs_f64_oracle": {"gpu": _quantiles(_rel_err(g_map, r64["mapping"])),
                              "f32_oracle": _quantiles(_rel_err(r32["mapping"], r64["mapping"]))}}


def bench_plain_256(reps: int = 5) -> dict:
    """The reference's CPU-runnable case: 256x256, Cornell box, swirl, h 0.01 (max 1000 RK4
    iterations), no overlay; device time of `reps` renders and the accuracy at 2048 pixels."""
    wl = configs.PLAIN_256_SWIRL
    v, f = wl.mesh()
    with L.MainView(wl.width, wl.height) as view:
        view.update_model(v, f)
        configs.setup_view(view, wl, "swirl")
        view.render()
        c = view.counters()
        view.reset_timings()
        for _ in range(reps):
            view.render()
        rms, pms = view.timings()
        fr = view.read_frame()
    r_ms = float(np.mean(rms))
    acc = sample_accuracy(wl.camera(), wl.settings(), wl.integrator(), "swirl", v, f, wl.width, wl.height,
                          fr.mapping, fr.hit_faces, 2048, seed=1)
    return {"workload": wl.name, "raycast_ms": r_ms, "frame_ms": float(np.mean(pms)), "rk4_steps": c["rk4_steps"],
            "gsteps_per_s": c["rk4_steps"] / (r_ms * 1e-3) / 1e9, "accuracy": acc}


def bench_adaptive(view, wl, verts, faces) -> dict:
    """The headline frames once more with the reference's adaptive step rule on (a redo at
    h = 0.001 when |V - last V| jumps tenfold, back to h after it calms; the shader's default,
    which the configurations' fixed step turns off): device time, RK4 steps (redos included)
    and the error of 256 random pixels against the FP32 and FP64 oracles, per field."""
    import dataclasses

    wla = dataclasses.replace(wl, adaptive=True)
    view.set_integrator(wla.integrator())
    out = {"workload": wl.name + " with adaptive steps", "per_field": {}}
    for name in wl.fields:
        view.set_field(name)
        view.render()
        c = view.counters()
        view.reset_timings()
        view.render()
        r_ms = float(view.timings()[0][0])
        fr = view.read_frame()
        out["per_field"][name] = {
            "raycast_ms": r_ms, "rk4_steps": c["rk4_steps"], "gsteps_per_s": c["rk4_steps"] / (r_ms * 1e-3) / 1e9,
            "accuracy": sample_accuracy(wla.camera(), wla.settings(), wla.integrator(), name, verts, faces,
                                        wla.width, wla.height, fr.mapping, fr.hit_faces, 256, seed=4)}
    view.set_integrator(wl.integrator())
    return out


def cpu_baseline(wl, verts, faces, budget_s: float, frames: dict | None = None, accuracy_pixels: int = 200_000) -> dict:
    """FP32 CPU restatement on a random sample of each field's pixels, timed with the reference's
    own scene loop: every face tested at every pre-hit step, in index order (the oracle's exact
    bounding-box prefilters switched off), as the reference shader does on any device.  With the
    GPU's frames given, a larger sample (the prefilters on: the same results, faster) yields the
    frame_accuracy() report, outside the timed region."""
    from oracle import oracle as O

    threads = max(1, min(16, O.threads_available()))
    cam = O.as_oracle_struct(wl.camera(), O.CameraIn)
    st = O.as_oracle_struct(wl.settings(), O.SettingsIn)
    ig = O.as_oracle_struct(wl.integrator(), O.IntegratorIn)
    W, H = wl.width, wl.height
    gx, gy = np.meshgrid(np.arange(W), np.arange(H))
    allpix = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.int32)
    rng = np.random.default_rng(0)
    order = rng.permutation(len(allpix))
    O.set_prefilter(False)
    try:
        # pilot: per-ray cost
        pilot = allpix[order[: 4 * threads]]
        t0 = time.perf_counter()
        for name in wl.fields:
            O.render_pixels(cam, st, ig, L.FIELDS[name], verts, faces, W, H, pilot, threads)
        per_ray = (time.perf_counter() - t0) / (len(pilot) * len(wl.fields))
        n = int(max(threads, min(len(allpix), budget_s / max(per_ray, 1e-9) / len(wl.fields))))
        sample = allpix[order[:n]]
        steps = 0
        t0 = time.perf_counter()
        for name in wl.fields:
            steps += O.render_pixels(cam, st, ig, L.FIELDS[name], verts, faces, W, H, sample, threads)["rk4_steps"]
        dt = time.perf_counter() - t0
    finally:
        O.set_prefilter(True)
    out = {"value": steps / dt, "unit": "RK4 ray-steps/s", "cores": threads, "kind": "port",
           "sample": f"{n} random pixels (seed 0) of each of the 4 fields at 1920x1080, FP32 CPU restatement "
                     f"of main_view.wgsl with its scene loop as the reference has it (all {len(faces)} faces "
                     f"tested at every pre-hit step), {steps} RK4 steps in {dt:.1f} s on {threads} threads",
           "frame_ms_extrapolated": dt / (n * len(wl.fields)) * W * H * 1e3}
    if frames:
        acc = allpix[order[:max(n, min(len(allpix), accuracy_pixels))]]
        f32 = {name: O.render_pixels(cam, st, ig, L.FIELDS[name], verts, faces, W, H, acc, threads)
               for name in wl.fields}
        out["accuracy"] = frame_accuracy(wl, verts, faces, frames, acc, f32, min(len(acc), 512), threads)
    return out


def copy_calibration() -> int:
    """1 GiB device-to-device copy by a plain copy kernel.  A 2 GiB fill in between evicts the
    source from the 256 MiB last-level cache, so the copy reads and writes HBM.  Profiled with
    the FETCH_SIZE and WRITE_SIZE counters, it gives the factor between each counter and bytes."""
    import torch
    n = 256 << 20  # floats: 1 GiB
    a = torch.ones(n, device="cuda")
    evict = torch.zeros(2 * n, device="cuda")
    evict.fill_(2.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {n * 4} bytes")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--batch-replays", type=int, default=2)
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=configs.SURROGATE_STEPS)
    ap.add_argument("--surrogate-train", action="store_true",
                    help="train the surrogates in this run even where saved weights match")
    ap.add_argument("--surrogate-fields", default="swirl,lens,hot_air,warp",
                    help="fields of the 1080p surrogate leg (empty: none; the stress leg keeps its own)")
    ap.add_argument("--skip-stress", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (the reference transfer that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        return copy_calibration()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the ray-cast pass has no multi-GPU path "
              "(rays are independent: it would shard by image tiles or by views, exchanging only "
              "the overlay's border rows between tiles)", file=sys.stderr)
        return 2

    wl = configs.OVERLAY_1080P
    verts, faces = wl.mesh()
    view = L.MainView(wl.width, wl.height)
    view.update_model(verts, faces)
    view.update_camera(wl.camera())
    view.update_settings(wl.settings())
    view.set_integrator(wl.integrator())

    # warm-up also records each field's (deterministic) work counters
    counters = {}
    frames = {}   # the last warm-up frame of each field, for the accuracy report
    for w in range(max(1, a.warmup)):
        last = w == max(1, a.warmup) - 1
        for name in wl.fields:
            view.set_field(name)
            view.render()
            counters[name] = view.counters()
            if last and not a.no_cpu_baseline:
                frames[name] = (view.read_mapping(), view.read_hit_faces(), view.read_lyapunov())
        log(f"warmup {w + 1}/{max(1, a.warmup)} done")

    view.synchronize()
    view.reset_timings()
    t0 = time.perf_counter()
    for k in range(a.steps):
        for name in wl.fields:
            view.set_field(name)
            view.render()
    view.synchronize()
    elapsed = time.perf_counter() - t0
    ray_ms, pass_ms = view.timings()
    log(f"timed {a.steps} steps in {elapsed:.3f} s")

    nf = len(wl.fields)
    assert len(ray_ms) == a.steps * nf
    per_field = {}
    for i, name in enumerate(wl.fields):
        rms = float(np.mean(ray_ms[i::nf]))
        pms = float(np.mean(pass_ms[i::nf]))
        c = counters[name]
        fl = kernel_flops(name, c)
        per_field[name] = {"raycast_ms": rms, "frame_ms": pms, "rk4_steps": c["rk4_steps"],
                           "triangle_tests": c["triangle_tests"], "pre_hit_steps": c["queries"] + c["skips"],
                           "flop": fl, "tflops": fl / (rms * 1e-3) / 1e12,
                           "gsteps_per_s": c["rk4_steps"] / (rms * 1e-3) / 1e9}
    steps_per_frame_set = sum(counters[n]["rk4_steps"] for n in wl.fields)
    value = steps_per_frame_set * a.steps / elapsed
    dom = max(per_field, key=lambda n: per_field[n]["raycast_ms"])
    d = per_field[dom]
    # the kernel each field runs under the default math choice, as rocprofv3 names it
    kname = {"swirl": "linon_wgslmath::raycast_kernel<lw::FieldSwirl>",
             "lens": "linon::raycast_kernel<lw::FieldLensT<true> >",
             "hot_air": "linon::raycast_kernel<lw::FieldHotAirT<true> >",
             "warp": "linon_wgslmath::raycast_kernel<lw::FieldWarp>"}[dom]
    roofline = {"bound": "valu", "kernel": f"void {kname}(RaycastParams)",
                "note": "compute-bound on FP32 vector ALUs (no matrix cores); peak = MI355X FP32 157.3 TFLOP/s",
                "achieved": d["tflops"],
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": d["tflops"] / PEAK_FP32_TFLOPS,
                "traffic": None,
                "flop_per_launch": d["flop"], "launch_ms": d["raycast_ms"],
                "all_raycast_kernels_tflops": sum(p["flop"] for p in per_field.values())
                / (sum(p["raycast_ms"] for p in per_field.values()) * 1e-3) / 1e12}
    # the attainable rate without packed FP32, measured on this device
    fma = L.measure_fma_rate()
    roofline["measured_fp32_fma_tflops"] = fma
    roofline["frac_of_measured_fma"] = d["tflops"] / fma if fma > 0 else None
    traffic_file = ROOT / "profiles" / "traffic.json"
    if traffic_file.exists():
        try:
            t = json.loads(traffic_file.read_text()).get(dom)
            if t:
                roofline["traffic"] = t
        except (ValueError, OSError):
            pass

    out = {
        "metric": "RK4 ray-steps/s + 1080p frame ms; surrogate vs plain RK4 at \u22641e-3 exit error",
        "value": value,
        "unit": "RK4 ray-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "frame_ms_1080p": float(np.mean([p["frame_ms"] for p in per_field.values()])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "width": wl.width, "height": wl.height, "fields": list(wl.fields),
                   "math": "auto: WGSL accuracy for swirl/warp, IEEE for lens/hot-air",
                   "h": wl.h, "t_max": wl.t_max, "max_steps": wl.max_steps, "adaptive": wl.adaptive,
                   "overlay": wl.overlay, "faces": int(len(faces)), "sample_rays": wl.samples,
                   "step": "one 1080p frame per field (4 frames)"},
        "per_field": per_field,
        "roofline": roofline,
    }
    extra = {}
    log("adaptive 1080p ...")
    extra["adaptive_1080p"] = bench_adaptive(view, wl, verts, faces)
    view.close()
    log("plain 256x256 ...")
    extra["plain_256_swirl"] = bench_plain_256()
    if not a.skip_batch:
        log("batched views ...")
        extra["batched_views"] = bench_batch(verts, faces, a.batch_replays)
    if not a.skip_surrogate:
        log("surrogate ...")
        extra["surrogate"] = bench_surrogate(a.surrogate_steps, [f for f in a.surrogate_fields.split(",") if f],
                                             train=a.surrogate_train)
    if not a.skip_stress:
        log("4K stress ...")
        extra["stress_4k"] = bench_stress(0 if a.skip_surrogate else a.surrogate_steps, a.surrogate_train)
    out["other_workloads"] = extra
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(wl, verts, faces, a.cpu_seconds, frames)
        # exit-point error of the GPU frames and of the FP32 oracle, both against FP64
        out["accuracy"] = out["cpu_baseline"].pop("accuracy", None)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out, default=_json_scalar), flush=True)
    return 0


def one_graph_ms(frames) -> float:
    """Captures `frames` (callables that enqueue one frame on the stream they are given) in order
    on a new stream into a single HIP graph (through the HIP runtime the library already runs
    on), replays it once to warm up and returns the device time of one more replay."""
    hip = C.CDLL("libamdhip64.so.7")
    vp = C.c_void_p

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {rc}")

    stream, graph, exe, e0, e1 = vp(), vp(), vp(), vp(), vp()
    ok(hip.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
    ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
    ok(hip.hipStreamBeginCapture(stream, 1), "hipStreamBeginCapture")  # thread-local mode
    for frame in frames:
        frame(stream.value)
    ok(hip.hipStreamEndCapture(stream, C.byref(graph)), "hipStreamEndCapture")
    ok(hip.hipGraphInstantiateWithFlags(C.byref(exe), graph, C.c_ulonglong(0)), "hipGraphInstantiateWithFlags")
    ok(hip.hipGraphLaunch(exe, stream), "hipGraphLaunch")
    ok(hip.hipEventRecord(e0, stream), "hipEventRecord")
    ok(hip.hipGraphLaunch(exe, stream), "hipGraphLaunch")
    ok(hip.hipEventRecord(e1, stream), "hipEventRecord")
    ok(hip.hipStreamSynchronize(stream), "hipStreamSynchronize")
    ms = C.c_float(0.0)
    ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
    hip.hipGraphExecDestroy(exe)
    hip.hipGraphDestroy(graph)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipStreamDestroy(stream)
    return float(ms.value)


def plain_frame(view, field):
    def enqueue(stream):
        view.set_field(field)
        view.render(stream=stream)
    return enqueue


def bench_stress(surrogate_steps: int = 0, train: bool = False) -> dict:
    """3840x2160 at h = 0.001 (5001 RK4 iterations per ray) against the Cornell box plus the
    1,310,720-face level-8 icosphere, lens and hot-air, overlay on, 8 outline rays x 64 records:
    one warm-up and one timed frame per field.  With surrogate_steps > 0, also the surrogate +
    fallback frame of each field at the same size and step (surrogate_leg), and then every frame
    -- the plain frames of both fields and their surrogate frames -- in one captured HIP graph."""
    wl = configs.STRESS_4K
    verts, faces = wl.mesh()
    view = L.MainView(wl.width, wl.height)
    t0 = time.perf_counter()
    view.update_model(verts, faces)
    mesh_s = time.perf_counter() - t0
    view.update_camera(wl.camera())
    view.update_settings(wl.settings())
    view.set_integrator(wl.integrator())
    out = {"workload": wl.name, "faces": int(len(faces)), "mesh_upload_seconds": mesh_s, "per_field": {}}
    for name in wl.fields:
        view.set_field(name)
        view.render()
        c = view.counters()
        # the timed frame's outline search reads this frame's exponents
        searched_hi = int((view.read_exponents() >= 0.8).sum())
        view.reset_timings()
        view.render()
        rms, pms = view.timings()
        fl = kernel_flops(name, c)
        r_ms = float(rms[0])
        out["per_field"][name] = {"raycast_ms": r_ms, "frame_ms": float(pms[0]), "rk4_steps": c["rk4_steps"],
                                  "triangle_tests": c["triangle_tests"],
                                  "gsteps_per_s": c["rk4_steps"] / (r_ms * 1e-3) / 1e9,
                                  "tflops": fl / (r_ms * 1e-3) / 1e12,
                                  # outline searches (of 8) that found a highlighted texel, and the
                                  # highlighted texels (exponent >= 0.8) in the buffer they searched
                                  "outline_rays_found": int(view.read_outline()[:, 2].sum()),
                                  "searched_exponents_ge_0.8": searched_hi}
        fr = view.read_frame()
        out["per_field"][name]["accuracy"] = sample_accuracy(wl.camera(), wl.settings(), wl.integrator(), name,
                                                             verts, faces, wl.width, wl.height, fr.mapping,
                                                             fr.hit_faces, 64, seed=2)
        log(f"stress {name}: {rms[0]:.1f} ms")
    frames = [plain_frame(view, name) for name in wl.fields]
    # the plain frames of both fields captured into one HIP graph on a caller's stream and
    # replayed (the graph keeps each capture's field; the frames share the output buffers)
    out["one_graph_all_fields_ms"] = one_graph_ms(frames)
    log(f"stress one graph ({'+'.join(wl.fields)}): {out['one_graph_all_fields_ms']:.1f} ms")
    if surrogate_steps > 0:
        sviews = []
        out["surrogate"] = {}
        for name in wl.fields:
            sv, res = surrogate_leg(name, surrogate_steps, wl, train)
            out["surrogate"][name] = res
            sviews.append(sv)
            frames.append(lambda stream, sv=sv: sv.render_surrogate(stream=stream))
        # every frame of the stress configuration in one graph: plain lens and hot air with the
        # mesh, overlay and outline rays, then the surrogate + fallback frames of both fields
        out["one_graph_plain_and_surrogate_ms"] = one_graph_ms(frames)
        out["one_graph_frames"] = [f"plain {n}" for n in wl.fields] + [f"surrogate {n}" for n in wl.fields]
        log(f"stress one graph (plain + surrogate frames): {out['one_graph_plain_and_surrogate_ms']:.1f} ms")
        for sv in sviews:
            sv.close()
    view.close()
    return out


def bench_batch(verts, faces, replays: int) -> dict:
    """65 views (64 drawn from seed 0 + one Enhance pass at h/4) in one hipGraph replay of the
    persistent kernels, escape mode, max 8000 RK4 steps, launch shape from the autotuner."""
    views = configs.batch_views(num_views=64, seed=0)
    view = L.MainView(64, 64)
    view.update_model(verts, faces)
    view.set_integrator(configs.batch_integrator(max_steps=8000))
    view.set_batch(views)
    t0 = time.perf_counter()
    tuning, tuned_ms = view.autotune_batch(downscale=4)
    tune_s = time.perf_counter() - t0
    view.render_batch()
    view.synchronize()
    c = view.counters()
    # accuracy of three views (one per field kind among the first ones) at 256 pixels each
    accuracy = {}
    seen = set()
    for i, bv in enumerate(views):
        if bv.field in seen or len(accuracy) == 3:
            continue
        seen.add(bv.field)
        fr = view.read_batch_view(i)
        ig = configs.batch_integrator(max_steps=8000)
        ig.h_initial = bv.h_initial
        name = {v: k for k, v in L.FIELDS.items()}[bv.field]
        accuracy[f"view{i}_{name}_{bv.width}x{bv.height}_h{bv.h_initial:g}"] = sample_accuracy(
            bv.camera, L.Settings.make(field_weight=bv.field_weight), ig, name, verts, faces, bv.width, bv.height,
            fr.mapping, fr.hit_faces, 256, seed=3)
    view.reset_timings()
    t0 = time.perf_counter()
    for _ in range(replays):
        view.render_batch()
    view.synchronize()
    wall = (time.perf_counter() - t0) / replays
    ms, _ = view.timings()
    dev_ms = float(np.mean(ms))
    rays = sum(v.width * v.height for v in views)
    # the same views one after another through the plain kernel (escape mode, same steps)
    plain_ms = 0.0
    per_field = {}
    names = {v: k for k, v in L.FIELDS.items()}
    ig = configs.batch_integrator(max_steps=8000)
    for bv in views:
        view.resize_texture(bv.width, bv.height)
        view.update_camera(bv.camera)
        view.update_settings(L.Settings.make(field_weight=bv.field_weight))
        ig.h_initial = bv.h_initial
        view.set_integrator(ig)
        view.set_field(bv.field)
        view.reset_timings()
        view.render()
        ms1 = float(view.timings()[0][0])
        plain_ms += ms1
        pf = per_field.setdefault(names[bv.field], {"views": 0, "rays": 0, "rk4_steps": 0, "plain_sequential_ms": 0.0})
        pf["views"] += 1
        pf["rays"] += bv.width * bv.height
        pf["rk4_steps"] += view.counters()["rk4_steps"]
        pf["plain_sequential_ms"] += ms1
    for pf in per_field.values():
        pf["plain_gsteps_per_s"] = pf["rk4_steps"] / (pf["plain_sequential_ms"] * 1e-3) / 1e9
    view.close()
    return {"workload": "65 views (seed 0): resolutions {256^2..2048^2}, eye radius U(2.5,4), fields swirl/lens/"
                        "hot-air/warp, h in {0.01,0.005,0.0025} + one Enhance view at h/4; escape mode, max 8000",
            "views": len(views), "rays": rays, "rk4_steps": c["rk4_steps"], "device_ms": dev_ms,
            "wall_ms": wall * 1e3, "gsteps_per_s": c["rk4_steps"] / (dev_ms * 1e-3) / 1e9,
            "mrays_per_s": rays / (dev_ms * 1e-3) / 1e6, "plain_kernel_sequential_ms": plain_ms,
            "speedup_vs_plain_sequential": plain_ms / dev_ms if dev_ms > 0 else None,
            "per_field": per_field,
            "tuning": tuning.as_dict(), "autotune_full_size_ms": tuned_ms, "autotune_seconds": tune_s,
            "accuracy": accuracy}


def surrogate_leg(field: str, steps: int, wl=configs.OVERLAY_1080P, train: bool = False):
    """The flow-map surrogate of the workload's step h for one field (configs.surrogate_config;
    the saved weights under linon_amd/surrogates/ when they match it and `train` is off, else
    trained here on 2^24 uniform entry states labelled by the FP32 RK4 kernel), then the
    workload's view without the mesh (mapping + overlay), rendered twice: plain RK4 (escape mode,
    exits on the sphere) and surrogate + fallback.  Returns the view (holding the surrogate) and
    the speed-up and, over the texels that took the surrogate's exit, the fraction within the
    1e-3 tolerance of the plain frame."""
    view = L.MainView(wl.width, wl.height)
    view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    view.update_camera(wl.camera())
    view.update_settings(wl.settings())
    cfg = configs.surrogate_config(field, wl, steps=steps)
    saved = configs.saved_surrogate(field, wl)
    if saved is not None and not train:
        view.load_surrogate(saved)
        tr = dict(view.surrogate_stats(), source=f"saved weights ({saved.name})")
    else:
        tr = dict(view.train_surrogate(cfg), source="trained in this run")
    ig = L.Integrator.make(h_initial=wl.h, t_max=1e30, adaptive=False, max_steps=wl.max_steps,
                           mode=L.MODE_ESCAPE, escape_radius=2.0, samples_per_ray=0, sample_steps=0,
                           need_mapping=True, clip_exit=True)
    view.set_integrator(ig)
    view.set_field(field)
    view.render()
    view.reset_timings()
    view.render()
    plain_ms = float(view.timings()[1][0])
    plain = view.read_mapping()
    view.render_surrogate()            # warm
    view.render_surrogate()
    st = view.surrogate_stats()
    sur = view.read_mapping()
    acc = view.read_surrogate_accepted()
    a = sur[..., :3].astype(np.float64)[acc]
    b = plain[..., :3].astype(np.float64)[acc]
    rel = (np.abs(a - b) / np.maximum(1.0, np.abs(b))).max(axis=-1) if acc.any() else np.zeros(0)
    res = {"train": {k: tr.get(k) for k in ("source", "train_loss", "label_seconds", "train_seconds",
                                            "holdout_err_q50", "holdout_err_q99", "holdout_frac_within_tol")},
           "h": wl.h, "max_steps": wl.max_steps, "coarse_h": cfg.coarse_h, "tolerance": cfg.tolerance,
           "plain_frame_ms": plain_ms, "surrogate_frame_ms": st["frame_ms"],
           "speedup": plain_ms / st["frame_ms"] if st["frame_ms"] > 0 else None,
           # exit-point error (max over x, y, z of |d| / max(1, |plain|)) of the texels that took
           # the surrogate's exit, against the plain frame
           "accepted_within_tol": float((rel <= cfg.tolerance).mean()) if acc.any() else None,
           "accepted_err_q50_q99_max": ([float(np.quantile(rel, q)) for q in (0.5, 0.99, 1.0)]
                                        if acc.any() else None),
           "nothing_accepted": not bool(acc.any()),
           "rays": int(st["frame_rays"]), "entered": int(st["entered"]), "accepted": int(st["accepted"]),
           "accepted_frac_of_entered": (int(st["accepted"]) / int(st["entered"]) if st["entered"] else None),
           "conf_fallback": int(st["conf_fallback"]), "lyapunov_fallback": int(st["lyapunov_fallback"]),
           "rejudged": int(st["rejudged"]), "rejudge_err_q50": st["rejudge_err_q50"],
           "rejudge_err_q99": st["rejudge_err_q99"], "rejudge_err_max": st["rejudge_err_max"],
           "phases_ms": {k: st[k] for k in ("approach_ms", "coarse_ms", "mlp_ms", "fallback_ms", "overlay_ms")},
           # the fused MLP kernel's rate (mlp_ms times that kernel alone, with HIP events on the
           # frame's stream); algorithmic FLOP per ray: 2 x (198*256 + 3*256*256 + 256*7) = 498,176
           "mlp_tflops": (int(st["entered"]) * 498176 / (st["mlp_ms"] * 1e-3) / 1e12
                          if st["mlp_ms"] > 0 else None)}
    log(f"surrogate {field} {wl.width}x{wl.height}: {res['speedup']}x, accepted "
        f"{res['accepted']}/{res['entered']}, within tol {res['accepted_within_tol']}")
    return view, res


def bench_surrogate(steps: int, fields, wl=configs.OVERLAY_1080P, train: bool = False) -> dict:
    """surrogate_leg for each field."""
    out = {}
    for field in fields:
        view, out[field] = surrogate_leg(field, steps, wl, train)
        view.close()
    return out


if __name__ == "__main__":
    sys.exit(main())
