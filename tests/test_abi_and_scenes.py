"""CPU checks of the C ABI library (it loads and exports every symbol include/linon.h
declares; no compute calls) and of the host-side scene inputs."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

import linon_amd as L
from linon_amd import scenes

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "linon.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(linon_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_mainview_surface():
    names = declared_functions()
    for n in ("linon_create", "linon_render", "linon_set_camera", "linon_set_settings", "linon_set_mesh",
              "linon_set_field_wgsl", "linon_render_high_accuracy", "linon_resize"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds the same set
    assert set(L.SIGNATURES) == set(declared_functions())
    assert lib.linon_abi_version() == 1


def test_struct_layouts_match_the_reference_uniforms():
    assert C.sizeof(L.Camera) == 112     # CameraUniform, src/main_view.rs:12-19
    assert C.sizeof(L.Settings) == 24    # Settings, src/main_view.rs:43-51
    assert C.sizeof(L.Integrator) == 48
    assert L.Settings.field_weight.offset == 0 and L.Settings.overlay_mode.offset == 12
    assert L.Settings.lyapunov_scaling.offset == 20
    assert L.Camera.view_matrix.offset == 48


def test_null_context_is_an_argument_error():
    """Every entry point that takes a context rejects a NULL one with LINON_ERR_ARGUMENT before
    it touches the device (so this runs without a GPU); destroy(NULL) is a no-op and
    last_error(NULL) still returns a message."""
    lib = L.load_library()
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "linon.h").read_text(), flags=re.S)
    decls = re.findall(r"linon_status\s+(linon_[a-z_0-9]+)\s*\(((?:const )?linon_context\*[^;]*)\);", text)
    assert len(decls) > 30
    for name, params in decls:
        f = getattr(lib, name)
        f.restype, f.argtypes = C.c_int, None
        assert f(*([None] * (params.count(",") + 1))) == 1, name   # LINON_ERR_ARGUMENT
    lib.linon_destroy(None)
    err = lib.linon_last_error
    err.restype, err.argtypes = C.c_char_p, [C.c_void_p]
    assert err(None)
    # the translator entry points take no context and reject a NULL body the same way
    assert lib.linon_wgsl_translate(None, None, 0, None) == 1
    assert lib.linon_wgsl_compile_check(None, None, 0) == 1


def test_create_without_gpu_fails_cleanly():
    import os
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible")
    with pytest.raises(L.LinonError):
        L.MainView(16, 16)


def test_cornell_box_matches_the_reference_table():
    d = json.loads((ROOT / "tests" / "golden" / "cornell_digest.json").read_text())
    v, f = scenes.cornell_box_raw()
    assert list(v.shape) == d["vertices_shape"] and list(f.shape) == d["faces_shape"]
    assert hashlib.sha256(v.astype(np.float32).tobytes()).hexdigest() == d["vertices_sha256"]
    assert hashlib.sha256(f.astype(np.uint32).tobytes()).hexdigest() == d["faces_sha256"]


def test_startup_scene_normalisation():
    v, f, c = scenes.reference_startup_scene()
    assert v.min() == pytest.approx(-1.0) and v.max() == pytest.approx(1.0)
    np.testing.assert_allclose(c, [-0.49241877, 0.50469315, -0.49530685], rtol=1e-6)
    vc, fc = scenes.cornell_box_centered()
    np.testing.assert_allclose((vc.min(0) + vc.max(0)) / 2, 0, atol=1e-6)


def test_icosphere_counts_and_seeded_displacement():
    for level in (0, 1, 3, 5):
        v, f = scenes.icosphere_unit(level)
        assert len(f) == 20 * 4 ** level and len(v) == 10 * 4 ** level + 2
        # closed 2-manifold: every edge shared by exactly two faces
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e[:, 0] * len(v) + e[:, 1], return_counts=True)
        assert np.all(cnt == 2)
    v, f = scenes.displaced_icosphere(5, 0.4, 0.05, seed=0)
    r = np.linalg.norm(v.astype(np.float64), axis=1)
    assert r.min() >= 0.35 - 1e-6 and r.max() <= 0.45 + 1e-6
    v2, _ = scenes.displaced_icosphere(5, 0.4, 0.05, seed=0)
    assert np.array_equal(v, v2)
    v3, _ = scenes.displaced_icosphere(5, 0.4, 0.05, seed=1)
    assert not np.array_equal(v, v3)
    vb, fb = scenes.benchmark_scene(5)
    assert len(fb) == 36 + 20480


def test_splitmix64_reference_values():
    # SplitMix64 from seed 0: published first outputs
    z = scenes.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_parse_obj_first_model_fans_and_renumbers(tmp_path):
    """OBJ ingest: quads become fans, v/vt/vn corners and negative indices resolve, only the
    first model is taken and its vertices are renumbered in order of first use."""
    text = """# a unit square and a triangle in two objects
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vn 0 0 1
o square
f 1/1/1 2//1 3 4/1
o tri
v 5 5 5
f -1 1 2
"""
    v, f = scenes.parse_obj(text)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert v.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    p = tmp_path / "m.obj"
    p.write_text("v 2 0 0\nv 0 4 0\nv 0 0 0\nv 9 9 9\nf 3 1 2\n")
    v2, f2, c2 = scenes.load_obj_scene(p)
    # vertices in first-use order (3, 1, 2); the unused vertex 4 is dropped before normalising
    assert f2.tolist() == [[0, 1, 2]]
    np.testing.assert_array_equal(v2, scenes.normalize_vertices(np.array([[0, 0, 0], [2, 0, 0], [0, 4, 0]], np.float32)))
    np.testing.assert_array_equal(c2, scenes.get_center(v2))
    with pytest.raises(ValueError):
        scenes.parse_obj("v 0 0 0\nf 1 2 3\n")


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says so, before any GPU
    work."""
    import subprocess
    import sys
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_flowmap_noalloc_rejects_more_than_its_staging():
    """linon_flowmap_noalloc checks n against its fixed device staging before it touches a
    device (so this runs without a GPU) and leaves a message for linon_last_error(NULL)."""
    lib = L.load_library()
    ig = L.Integrator.make(mode=L.MODE_ESCAPE)
    x = np.zeros((1, 6), np.float32)
    st = lib.linon_flowmap_noalloc(0, L.FIELDS["swirl"], 0, C.byref(ig), 1.0, C.c_void_p(x.ctypes.data),
                                   L.FLOWMAP_NOALLOC_MAX + 1, C.c_void_p(x.ctypes.data), None)
    assert st == 1
    msg = lib.linon_last_error(None).decode()
    assert str(L.FLOWMAP_NOALLOC_MAX) in msg and "staging" in msg
    assert lib.linon_flowmap_noalloc(0, L.FIELDS["swirl"], 0, None, 1.0, None, 1, None, None) == 1
