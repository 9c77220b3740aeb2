"""GPU: the surrogate's BF16-MFMA GEMM against numpy, and the learned flow-map surrogate
(training, frame with confidence / Lyapunov fallback to RK4, re-judging)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import linon_amd as L
from linon_amd import configs, scenes
from oracle import oracle as O

from gpu_common import EXIT_RTOL, PROBE_VARIANTS, HipGraph, compare, entry_states, rel_err, run_oracles

pytestmark = pytest.mark.gpu


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even float32 -> bf16 bit patterns."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def run_gemm(A16, B16, M, N, K, a_trans, b_trans, splits=1, bias=None, silu=False):
    lib = L.load_library()
    out = np.zeros((M, N), np.float32)
    out16 = np.zeros((M, N), np.uint16)
    bias_p = None if bias is None else C.c_void_p(np.ascontiguousarray(bias, np.float32).ctypes.data)
    st = lib.linon_selftest_gemm(0, C.c_void_p(A16.ctypes.data), C.c_void_p(B16.ctypes.data), M, N, K, a_trans,
                                 b_trans, splits, bias_p, 1 if silu else 0, C.c_void_p(out.ctypes.data),
                                 C.c_void_p(out16.ctypes.data))
    assert st == 0
    return out, out16


@pytest.mark.parametrize("M,N,K,a_trans,b_trans,splits", [
    (64, 64, 32, 0, 0, 1),       # one tile: pins the MFMA operand/result lane layouts
    (200, 96, 224, 0, 1, 1),     # ragged M/N, forward layout (A [M][K], B stored [N][K])
    (130, 256, 256, 0, 0, 1),    # backward dX layout (B stored [K][N])
    (224, 256, 4096, 1, 0, 4),   # weight-gradient layout (A stored [K][M]) with split-K
    (1000, 256, 224, 0, 1, 1),   # forward layout through the 128x128 fast path (ragged M, K = 224)
    (4096, 256, 256, 0, 1, 1),   # fast path, full tiles
])
def test_mfma_gemm_matches_numpy(M, N, K, a_trans, b_trans, splits):
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    A16 = bf16_bits(A.T.copy() if a_trans else A)
    B16 = bf16_bits(B.T.copy() if b_trans else B)
    ref = bf16_value(bf16_bits(A)).astype(np.float64) @ bf16_value(bf16_bits(B)).astype(np.float64)
    print(f"[case] gemm M={M} N={N} K={K} a_trans={a_trans} b_trans={b_trans} splits={splits}", flush=True)
    got, _ = run_gemm(A16, B16, M, N, K, a_trans, b_trans, splits)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4 * np.sqrt(K))


@pytest.mark.parametrize("M,N,K", [(96, 64, 64), (300, 256, 224)])   # generic path, fast path
def test_gemm_bias_silu_epilogue(M, N, K):
    rng = np.random.default_rng(5)
    A, B = rng.normal(size=(M, K)).astype(np.float32), rng.normal(size=(K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    z = bf16_value(bf16_bits(A)).astype(np.float64) @ bf16_value(bf16_bits(B)).astype(np.float64) + bias
    b_trans = 1 if N % 128 == 0 else 0   # the fast path needs B stored [N][K]
    B16 = bf16_bits(B.T.copy()) if b_trans else bf16_bits(B)
    got_z, got16 = run_gemm(bf16_bits(A), B16, M, N, K, 0, b_trans, 1, bias, silu=True)
    np.testing.assert_allclose(got_z, z, rtol=1e-4, atol=1e-3)
    silu = z / (1 + np.exp(-z))
    np.testing.assert_allclose(bf16_value(got16), silu, rtol=1e-2, atol=1e-2)


@pytest.fixture(scope="module")
def trained():
    view = L.MainView(96, 64)
    view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    cfg = L.SurrogateConfig.make("swirl", train_rays=1 << 18, batch=8192, steps=400, h=0.01, max_steps=2000,
                                 conf_threshold=0.9, tolerance=1e-2, rejudge_fraction=0.25, seed=1)
    print("[case] surrogate training (swirl, 2^18 rays, 400 steps)", flush=True)
    stats = view.train_surrogate(cfg)
    yield view, cfg, stats
    view.close()


def test_training_learns_the_flow_map(trained):
    view, cfg, stats = trained
    print(stats, flush=True)
    assert np.isfinite(stats["train_loss"])
    # untrained: exit-point error ~ the sphere's size (relative error ~1); trained: well below
    assert 0.0 < stats["holdout_err_q50"] < 0.2   # exactly 0 would mean the evaluation never ran
    assert stats["train_rays"] == cfg.train_rays


def check_flowmap(field, ig, states, label):
    """GPU flow map (linon_flowmap) against the oracle's escape-mode flow map: the exit point
    within EXIT_RTOL of the FP32 oracle where FP32 rounding cannot move it by 1e-5 (all FP32
    variants of the oracle within 1e-5 of FP64), elsewhere as accurate against FP64 as 4x the
    worst FP32 variant, on >= 99.9% of the states; iteration counts equal on >= 99%."""
    fid = L.FIELDS[field]
    got, steps = L.flowmap(field, ig, states)
    igo = O.as_oracle_struct(ig, O.IntegratorIn)
    r32, s32 = O.flowmap(fid, igo, states, threads=8)
    r64, _ = O.flowmap(fid, igo, states.astype(np.float64), threads=8, precision="f64")
    e_probe = rel_err(r32, r64)
    for var in PROBE_VARIANTS:
        e_probe = np.maximum(e_probe, rel_err(O.flowmap(fid, igo, states, threads=8, precision=var)[0], r64))
    eg, d = rel_err(got, r64), rel_err(got, r32)
    well = e_probe <= 1e-5
    ok = np.where(well, d <= EXIT_RTOL, eg <= np.maximum(4.0 * e_probe, EXIT_RTOL))
    info = dict(label=label, n=len(states), well=float(well.mean()), ok=float(ok.mean()),
                agree_1e4=float((d <= EXIT_RTOL).mean()), steps_equal=float((steps == s32).mean()),
                gpu_vs_f64_q50_q99=[float(np.quantile(eg, 0.5)), float(np.quantile(eg, 0.99))],
                f32_vs_f64_q50_q99=[float(np.quantile(e_probe, 0.5)), float(np.quantile(e_probe, 0.99))])
    print(info, flush=True)
    assert ok.mean() >= 0.999, info
    assert (steps == s32).mean() >= 0.99, info
    return got, r32


@pytest.mark.parametrize("clip_exit", [0, 1])
@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_flowmap_matches_the_oracle(field, clip_exit):
    """The RK4 flow map the surrogate learns and falls back to (escape mode, no mesh, h 0.005):
    entry states on the sphere, exits against the oracle.  With clip_exit every ray that
    leaves ends on the sphere."""
    ig = L.Integrator.make(h_initial=0.005, t_max=1e30, adaptive=False, max_steps=4000, mode=L.MODE_ESCAPE,
                           escape_radius=2.0, samples_per_ray=0, sample_steps=0, need_mapping=True,
                           clip_exit=bool(clip_exit))
    print(f"[case] flow map {field} clip_exit={clip_exit}", flush=True)
    got, r32 = check_flowmap(field, ig, entry_states(2048, seed=L.FIELDS[field]), f"flowmap-{field}-clip{clip_exit}")
    if clip_exit:
        # rays that leave (lens captures most of them near its singularity: they end at max_steps)
        left = np.abs(np.linalg.norm(r32[:, :3].astype(np.float64), axis=1) - 2.0) <= 1e-5
        assert left.sum() >= 200
        # (a captured lens ray may be set free on one side of the rounding and not the other)
        on = np.abs(np.linalg.norm(got[:, :3].astype(np.float64), axis=1) - 2.0) <= 1e-5
        assert on[left].mean() >= 0.99


def test_labels_are_the_rk4_flow_map(trained):
    """The surrogate's labels and RK4 legs are the flow map at the surrogate's own step
    parameters (escape mode at h, need_mapping, exits clipped to the sphere, max_steps): that
    flow map, on states drawn as the training set draws them, against the oracle's."""
    view, cfg, _ = trained
    ig = L.Integrator.make(h_initial=cfg.h, t_max=1e30, adaptive=False, max_steps=cfg.max_steps,
                           mode=L.MODE_ESCAPE, escape_radius=cfg.radius, samples_per_ray=0, sample_steps=0,
                           need_mapping=True, clip_exit=True)
    print("[case] surrogate labels against the oracle", flush=True)
    check_flowmap("swirl", ig, entry_states(1024, seed=99, radius=cfg.radius, inset=cfg.h), "labels-swirl")


def test_surrogate_frame_against_plain_rk4(trained):
    """Rays the surrogate sends to RK4 reproduce the plain escape-mode frame bit for bit;
    accepted rays are within what the re-judged sample says."""
    view, cfg, _ = trained
    cam = L.Camera.look_at((0.2, 0.3, 3.0), (0, 0, 0), fov_deg=60.0)
    view.update_camera(cam)
    view.update_settings(L.Settings.make(overlay_mode=1))
    # plain RK4 frame: escape mode, need_mapping, exits on the sphere, no mesh
    ig = L.Integrator.make(h_initial=cfg.h, t_max=1e30, adaptive=False, max_steps=cfg.max_steps, mode=L.MODE_ESCAPE,
                           escape_radius=cfg.radius, samples_per_ray=0, sample_steps=0, need_mapping=True,
                           clip_exit=True)
    view.set_integrator(ig)
    view.set_field("swirl")
    view.render()
    plain_frame = view.read_frame()
    plain = plain_frame.mapping
    # ... which is the oracle's frame (so every fallback ray below is too)
    empty_v, empty_f = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    ref, ref64 = run_oracles(cam, L.Settings.make(overlay_mode=1), ig, L.FIELDS["swirl"], empty_v, empty_f, 96, 64)
    info = compare(plain_frame, ref, ref64, "surrogate-plain-frame")
    assert info["exit_pass"] and info["q99_ok"], info
    view.render_surrogate()
    st = view.surrogate_stats()
    sur = view.read_mapping()
    print(st, flush=True)
    n = st["frame_rays"]
    assert st["accepted"] + st["conf_fallback"] + st["lyapunov_fallback"] == st["entered"]
    diff = np.linalg.norm(sur[..., :3].astype(np.float64) - plain[..., :3], axis=-1)
    rel = diff / np.maximum(1.0, np.linalg.norm(plain[..., :3].astype(np.float64), axis=-1))
    # every ray that did not take the surrogate's answer is the plain RK4 result: the integrator
    # rounds alike in every kernel; the field body itself may be contracted differently by the
    # compiler in the surrogate's kernels, which moves a few rays by rounding amounts
    not_taken = n - st["accepted"]
    assert (diff == 0).sum() >= 0.995 * not_taken
    assert (rel <= 1e-4).sum() >= not_taken
    if st["rejudged"]:
        assert np.isfinite(st["rejudge_err_max"])
        assert st["rejudge_err_q50"] <= st["rejudge_err_max"]


def _device_free_bytes() -> int:
    """Free memory of device 0 as the HIP runtime reports it (device-wide)."""
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_closing_a_view_frees_its_surrogate():
    """linon_destroy releases every device buffer of the context, the trained surrogate's
    included (2^22 training states and labels, weights, optimizer moments, activations: several
    hundred MB).  Three create / train / close cycles leave the device's free memory where it
    was after the first."""
    def cycle():
        with L.MainView(64, 48) as view:
            view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
            view.train_surrogate(L.SurrogateConfig.make("warp", train_rays=1 << 22, batch=8192, steps=5, h=0.01,
                                                        max_steps=500))
    print("[case] surrogate create/train/close cycles", flush=True)
    cycle()   # the first use of each kernel and the runtime's own allocations
    before = _device_free_bytes()
    for _ in range(3):
        cycle()
    after = _device_free_bytes()
    assert before - after < 128 << 20, (before, after)


def test_fused_inference_equals_the_training_path(trained):
    """The fused inference kernel of a frame (features, four BF16-MFMA layers, FP32 output layer
    in one launch, activations in LDS) returns exactly the training path's outputs (feature
    kernel, four GEMM launches, output-layer kernel), ragged tiles included."""
    view, cfg, _ = trained
    for n, seed in ((1000, 7), (4096, 8)):
        s = entry_states(n, seed, radius=cfg.radius, inset=cfg.h)
        print(f"[case] fused inference n={n}", flush=True)
        a = view.surrogate_infer(s, fused=True)
        b = view.surrogate_infer(s, fused=False)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_save_and_load_give_the_same_surrogate(trained):
    """A saved surrogate (configuration, frequencies, FP32 weights, output scales) loaded into
    another context predicts the same bits; a damaged blob is refused."""
    view, cfg, stats = trained
    blob = view.save_surrogate()
    s = entry_states(512, 9, radius=cfg.radius, inset=cfg.h)
    ref = view.surrogate_infer(s)
    with L.MainView(32, 24) as other:
        other.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
        other.load_surrogate(blob)
        assert np.array_equal(other.surrogate_infer(s), ref)
        assert other.surrogate_stats()["holdout_err_q50"] == pytest.approx(stats["holdout_err_q50"])
        other.render_surrogate()
        bad = bytearray(blob)
        bad[0] ^= 0xFF
        with pytest.raises(L.LinonError):
            other.load_surrogate(bytes(bad))
        with pytest.raises(L.LinonError):
            other.load_surrogate(blob[:-4])


@pytest.fixture(scope="module")
def trained_coarse():
    """The multi-fidelity form: the network learns the residual of the flow map at h = 0.01 over
    the flow map at coarse_h = 0.1 (which is itself within 1e-3 of it for swirl)."""
    view = L.MainView(96, 64)
    view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    # the Lyapunov gate's threshold is the overlay's 0.0955 at the 1080-row pitch; lambda grows as
    # the pixel pitch, so at 64 rows it is 1080 / 64 times that
    cfg = L.SurrogateConfig.make("swirl", train_rays=1 << 18, batch=8192, steps=600, h=0.01, max_steps=2000,
                                 conf_threshold=0.9, tolerance=1e-3, rejudge_fraction=0.25, seed=2, coarse_h=0.1,
                                 lyapunov_threshold=0.0955 * 1080 / 64)
    print("[case] multi-fidelity surrogate training (swirl, coarse h 0.1)", flush=True)
    stats = view.train_surrogate(cfg)
    yield view, cfg, stats
    view.close()


def test_multifidelity_surrogate_accepts_rays_within_tolerance(trained_coarse):
    """Accepted rays: a majority of the entered ones, and on >= 99% of them the exit point is
    within the 1e-3 tolerance of the plain RK4 frame (which the oracle pins); the re-judged
    sample agrees.  Every other texel is the plain frame's."""
    view, cfg, stats = trained_coarse
    print(stats, flush=True)
    assert stats["holdout_frac_within_tol"] > 0.9
    cam = L.Camera.look_at((0.2, 0.3, 3.0), (0, 0, 0), fov_deg=60.0)
    view.update_camera(cam)
    view.update_settings(L.Settings.make(overlay_mode=1))
    ig = L.Integrator.make(h_initial=cfg.h, t_max=1e30, adaptive=False, max_steps=cfg.max_steps, mode=L.MODE_ESCAPE,
                           escape_radius=cfg.radius, samples_per_ray=0, sample_steps=0, need_mapping=True,
                           clip_exit=True)
    view.set_integrator(ig)
    view.set_field("swirl")
    view.render()
    plain = view.read_mapping()
    view.render_surrogate()
    st = view.surrogate_stats()
    sur = view.read_mapping()
    acc = view.read_surrogate_accepted()
    print(st, flush=True)
    assert st["accepted"] == int(acc.sum()) and st["accepted"] > 0.5 * st["entered"] > 0
    rel = rel_err(sur, plain)
    assert (rel[acc] <= cfg.tolerance).mean() >= 0.99, np.quantile(rel[acc], [0.5, 0.99, 1.0])
    assert st["rejudged"] > 0 and st["rejudge_err_q99"] <= cfg.tolerance
    assert (rel[~acc] <= 1e-4).all()


def test_plain_and_surrogate_frames_in_one_graph(trained_coarse):
    """A caller captures a plain frame of one view (mesh, overlay) and the surrogate + fallback
    frame of another view into one HIP graph on its own stream.  With eager frames of both views
    between the replays, every replay leaves both views' outputs as their eager frames left
    them, bit for bit, and the surrogate's counts as the eager frame's (they stay on the
    device, so the replay rewrites them)."""
    sview, cfg, _ = trained_coarse
    cam = L.Camera.look_at((0.2, 0.3, 3.0), (0, 0, 0), fov_deg=60.0)
    sview.update_camera(cam)
    sview.update_settings(L.Settings.make(overlay_mode=1))
    sview.set_integrator(L.Integrator.make(h_initial=cfg.h, t_max=1e30, adaptive=False, max_steps=cfg.max_steps,
                                           mode=L.MODE_ESCAPE, escape_radius=cfg.radius, samples_per_ray=0,
                                           sample_steps=0, need_mapping=True, clip_exit=True))
    sview.set_field("swirl")
    sview.render_surrogate()
    want_s = sview.read_mapping()
    want_acc = sview.read_surrogate_accepted()
    want_st = sview.surrogate_stats()
    assert want_st["accepted"] > 0
    v, f = scenes.cornell_box_centered()
    with L.MainView(64, 48) as pv:
        pv.update_model(v, f)
        pv.update_camera(L.Camera.look_at((0.0, 0.0, 3.0), (0, 0, 0), fov_deg=60.0))
        pv.update_settings(L.Settings.make(overlay_mode=1))
        pv.set_integrator(L.Integrator.make())
        pv.set_field("lens")
        pv.render()
        want_p = pv.read_frame()
        print("[case] plain + surrogate frames in one graph", flush=True)
        g = HipGraph([lambda st: pv.render(stream=st), lambda st: sview.render_surrogate(stream=st)])
        try:
            for rep in range(3):
                pv.set_field("swirl")            # eager work on both views between the replays
                pv.render()
                sview.render()                   # the plain RK4 frame: other exits where accepted
                assert not np.array_equal(sview.read_mapping().view(np.uint32), want_s.view(np.uint32))
                pv.set_field("lens")
                g.replay()
                got = pv.read_frame()
                assert np.array_equal(got.mapping.view(np.uint32), want_p.mapping.view(np.uint32)), rep
                assert np.array_equal(got.final, want_p.final), rep
                assert np.array_equal(got.hit_faces, want_p.hit_faces), rep
                assert np.array_equal(sview.read_mapping().view(np.uint32), want_s.view(np.uint32)), rep
                assert np.array_equal(sview.read_surrogate_accepted(), want_acc), rep
                st = sview.surrogate_stats()
                for k in ("entered", "accepted", "conf_fallback", "lyapunov_fallback"):
                    assert st[k] == want_st[k], (rep, k)
                # the re-judged sample is drawn afresh per frame (the capture is a later frame)
                assert st["rejudged"] > 0 and st["rejudge_err_q99"] <= cfg.tolerance, st
        finally:
            g.close()


@pytest.mark.parametrize("field", ["swirl", "lens", "hot_air", "warp"])
def test_saved_surrogates_at_1080p(field):
    """The saved surrogate weights the bench loads, on the 1080p workload's view without the mesh:
    swirl, warp and hot air accept most of the rays that enter the sphere (lens's captured rays
    leave the network unconfident, and it may accept none), >= 99% of the accepted exits are
    within the 1e-3 tolerance of the plain RK4 escape-mode frame, the re-judged sample's q99 is
    within it, and every texel that did not take the surrogate's exit is the plain frame's to
    1e-4 (the surrogate's RK4 legs round as the plain kernel does)."""
    wl = configs.OVERLAY_1080P
    path = configs.saved_surrogate(field, wl)
    assert path is not None, f"no saved weights for {field}"
    with L.MainView(wl.width, wl.height) as view:
        view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
        view.update_camera(wl.camera())
        view.update_settings(wl.settings())
        view.load_surrogate(path)
        view.set_integrator(L.Integrator.make(h_initial=wl.h, t_max=1e30, adaptive=False, max_steps=wl.max_steps,
                                              mode=L.MODE_ESCAPE, escape_radius=2.0, samples_per_ray=0,
                                              sample_steps=0, need_mapping=True, clip_exit=True))
        view.set_field(field)
        print(f"[case] saved surrogate {field} at 1080p", flush=True)
        view.render()
        plain = view.read_mapping()
        view.render_surrogate()
        st = view.surrogate_stats()
        sur = view.read_mapping()
        acc = view.read_surrogate_accepted()
    print({k: st[k] for k in ("entered", "accepted", "conf_fallback", "lyapunov_fallback", "rejudged",
                              "rejudge_err_q99")}, flush=True)
    assert st["accepted"] == int(acc.sum())
    rel = rel_err(sur, plain)
    assert (rel[~acc] <= 1e-4).mean() >= 0.999
    if field == "lens":
        assert st["accepted"] == 0 or (rel[acc] <= 1e-3).mean() >= 0.99
        return
    assert st["accepted"] > 0.6 * st["entered"] > 0
    assert (rel[acc] <= 1e-3).mean() >= 0.99, np.quantile(rel[acc], [0.5, 0.99, 1.0])
    assert st["rejudged"] > 0 and st["rejudge_err_q99"] <= 1e-3


def test_surrogate_frame_with_no_entered_rays(trained_coarse):
    """A camera that looks away from the bounding sphere: no ray enters, every launch of the
    frame reads a zero count on the device and does nothing, and the mapping is the plain
    escape-mode frame's."""
    view, cfg, _ = trained_coarse
    view.update_camera(L.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 6.0), fov_deg=30.0))
    view.update_settings(L.Settings.make(overlay_mode=1))
    view.set_integrator(L.Integrator.make(h_initial=cfg.h, t_max=1e30, adaptive=False, max_steps=cfg.max_steps,
                                          mode=L.MODE_ESCAPE, escape_radius=cfg.radius, samples_per_ray=0,
                                          sample_steps=0, need_mapping=True, clip_exit=True))
    view.set_field("swirl")
    print("[case] surrogate frame, no ray enters", flush=True)
    view.render()
    plain = view.read_mapping()
    view.render_surrogate()
    st = view.surrogate_stats()
    assert st["entered"] == 0 and st["accepted"] == 0 and st["rejudged"] == 0, st
    assert not view.read_surrogate_accepted().any()
    assert np.array_equal(view.read_mapping().view(np.uint32), plain.view(np.uint32))
