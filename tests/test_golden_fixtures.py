"""The CPU oracle reproduces the committed golden cases (tests/golden/oracle_fixtures.npz,
written by tests/golden/make_oracle_fixtures.py) bit for bit: any change to the restatement's
arithmetic, loop or tie-breaking shows up here before it reaches a GPU comparison."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import linon_amd as L
from oracle import oracle as O

FIX = Path(__file__).with_name("golden") / "oracle_fixtures.npz"


def load_cases():
    z = np.load(FIX, allow_pickle=False)
    return z, [str(n) for n in z["cases"]]


def case_inputs(z, name):
    p = name + "/"
    cam = L.Camera.from_buffer_copy(z[p + "camera"].tobytes())
    st = L.Settings.from_buffer_copy(z[p + "settings"].tobytes())
    ig = L.Integrator.from_buffer_copy(z[p + "integrator"].tobytes())
    W, H = (int(x) for x in z[p + "size"])
    return cam, st, ig, int(z[p + "field"][0]), z[p + "verts"], z[p + "faces"], W, H


Z, NAMES = load_cases()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    cam, st, ig, fid, v, f, W, H = case_inputs(Z, name)
    c, s, i = (O.as_oracle_struct(cam, O.CameraIn), O.as_oracle_struct(st, O.SettingsIn),
               O.as_oracle_struct(ig, O.IntegratorIn))
    p = name + "/"
    r32 = O.render(c, s, i, fid, v, f, W, H, threads=4)
    for key, arr in (("hit", r32["hit"]), ("steps", r32["steps"]), ("color8", r32["color8"]),
                     ("mapping_f32", r32["mapping"]), ("noise_flag", r32["noise_flag"])):
        assert np.array_equal(arr, Z[p + key], equal_nan=key.startswith("mapping")), key
    r64 = O.render(c, s, i, fid, v, f, W, H, threads=4, precision="f64")
    assert np.array_equal(r64["mapping"], Z[p + "mapping_f64"], equal_nan=True)
    assert (Z[p + "hit"] >= -1).all() and Z[p + "mapping_f32"].shape == (H, W, 4)


def test_fixtures_cover_hits_and_misses():
    hits = np.concatenate([Z[n + "/hit"].ravel() for n in NAMES])
    assert (hits >= 0).mean() > 0.2 and (hits == -1).mean() > 0.05
