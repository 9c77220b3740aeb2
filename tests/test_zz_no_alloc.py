"""GPU: linon_flowmap_noalloc, the entry point that allocates no device memory (it stages in
the library's fixed per-device array), returns exactly what linon_flowmap returns, on two calls
in a row, the second with fewer states and another field."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import linon_amd as L

from gpu_common import entry_states

pytestmark = pytest.mark.gpu


def _free_bytes() -> int:
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_flowmap_noalloc_equals_flowmap():
    ig = L.Integrator.make(h_initial=0.005, t_max=1e30, adaptive=False, max_steps=4000, mode=L.MODE_ESCAPE,
                           escape_radius=2.0, samples_per_ray=0, sample_steps=0, need_mapping=True, clip_exit=True)
    calls = [("hot_air", L.FLOWMAP_NOALLOC_MAX, 1), ("swirl", 1000, 2)]
    for field, n, seed in calls:
        print(f"[case] flowmap_noalloc {field} n={n}", flush=True)
        s = entry_states(n, seed)
        ref, ref_steps = L.flowmap(field, ig, s)
        got, got_steps = L.flowmap(field, ig, s, noalloc=True)
        assert np.array_equal(got, ref), field
        assert np.array_equal(got_steps, ref_steps), field
    # repeated calls take no device memory
    s = entry_states(4096, 4)
    L.flowmap("warp", ig, s, noalloc=True)
    before = _free_bytes()
    for _ in range(3):
        L.flowmap("warp", ig, s, noalloc=True)
    assert _free_bytes() == before
    # beyond the fixed staging: rejected with a message naming the limit
    with pytest.raises(L.LinonError) as e:
        L.flowmap("swirl", ig, entry_states(L.FLOWMAP_NOALLOC_MAX + 1, 3), noalloc=True)
    assert e.value.status == 1 and str(L.FLOWMAP_NOALLOC_MAX) in e.value.message
