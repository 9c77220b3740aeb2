"""GPU: the committed golden cases (tests/golden/oracle_fixtures.npz: inputs and the oracle's
FP32 / FMA-built FP32 / FP64 outputs) replayed through the C ABI, held to the parity bar of
tests/gpu_common.py."""
from __future__ import annotations

import numpy as np
import pytest

import linon_amd as L

from gpu_common import HIT_MIN_AGREE, compare
from test_golden_fixtures import NAMES, Z, case_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def view():
    v = L.MainView(32, 32)
    yield v
    v.close()


@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden_case(view, name):
    cam, st, ig, fid, v, f, W, H = case_inputs(Z, name)
    print(f"[case] golden {name} {W}x{H} field={fid}", flush=True)
    view.resize_texture(W, H)
    view.update_model(v, f)
    view.update_camera(cam)
    view.update_settings(st)
    view.set_integrator(ig)
    view.set_field(fid)
    view.render()
    g = view.read_frame()
    p = name + "/"
    ref = {"mapping": Z[p + "mapping_f32"], "hit": Z[p + "hit"], "noise_flag": Z[p + "noise_flag"],
           "color8": Z[p + "color8"]}
    ref64 = {"mapping": Z[p + "mapping_f64"]}
    for var in ("f32fma", "f32up", "f32dn"):
        ref64["mapping_" + var], ref64["hit_" + var] = Z[p + "mapping_" + var], Z[p + "hit_" + var]
    info = compare(g, ref, ref64, f"golden-{name}")
    assert info["gated"] > 0, info
    assert info["hit_ok"], info
    assert info["exit_pass"], info
    assert info["q99_ok"], info
    assert info["color_maxdiff"] <= 1, info
