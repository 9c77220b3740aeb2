"""Trains the flow-map surrogate of each benchmark field on the GPU and saves its weights.

    python -m linon_amd.train_surrogates [--workload 1080p|stress] [--fields swirl,warp,...]
                                         [--steps N] [--out DIR]

Each field: 2^24 uniform entry states on the radius-2 sphere labelled by the FP32 RK4 kernel at
the workload's step and step cap (1080p: h 0.005, 4000 steps; stress: h 0.001, 20000 steps;
exits clipped to the sphere), Adam with a cosine learning rate over N steps of 65,536 rays, then
the hold-out error.  Without --fields, every field of the workload.  The weights (configuration,
frequencies, FP32 master weights, output scales, training statistics) go to
DIR/<field>_h<h>_max<steps>_coarse<coarse_h>.lnsg, which bench.py loads instead of training.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import MainView
from . import configs

WORKLOADS = {"1080p": configs.OVERLAY_1080P, "stress": configs.STRESS_4K}

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="1080p")
    ap.add_argument("--fields", default=None)
    ap.add_argument("--steps", type=int, default=configs.SURROGATE_STEPS)
    ap.add_argument("--out", default=str(configs.SURROGATE_DIR))
    a = ap.parse_args()
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    wl = WORKLOADS[a.workload]
    for field in (a.fields.split(",") if a.fields else wl.fields):
        with MainView(64, 64) as view:
            view.update_model(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
            cfg = configs.surrogate_config(field, wl, steps=a.steps)
            print(f"[train] {field}: coarse_h {cfg.coarse_h}, {a.steps} steps ...", flush=True)
            st = view.train_surrogate(cfg)
            path = out / configs.surrogate_blob_name(field, wl)
            view.save_surrogate(path)
            print(json.dumps({"field": field, "file": path.name, **{k: st[k] for k in (
                "train_loss", "label_seconds", "train_seconds", "holdout_err_q50", "holdout_err_q99",
                "holdout_frac_within_tol")}}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
