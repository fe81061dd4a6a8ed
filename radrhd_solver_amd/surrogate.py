"""Trained c2p surrogate models (RHLLC_GUESS_MLP).

A model is trained on the device by the library's own kernels
(``rhllc_mlp_train``: BF16 MFMA forward/backward, Adam) and stored as raw
float32 (RHLLC_MLP_MODEL floats) with a JSON sidecar holding the training
configuration and the loss curve summary.  ``load_or_train`` returns the
stored model for a gamma, training and storing it first when there is none.

    python -m radrhd_solver_amd.surrogate --gamma 1.6666666666666667 --steps 20000 --batch 65536
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import capi
from ._lib import PKG_DIR

MODEL_DIR = os.path.join(PKG_DIR, "models")
DEFAULTS = {"steps": 20000, "batch": 65536, "lr": 3e-3, "seed": 20261202, "norm_samples": 1 << 20}


def model_path(gamma: float) -> str:
    return os.path.join(MODEL_DIR, f"c2p_mlp_gamma{gamma:.6f}.f32")


def load(gamma: float):
    """The stored model for gamma (float32 array), or None."""
    p = model_path(gamma)
    if not os.path.exists(p):
        return None
    m = np.fromfile(p, dtype=np.float32)
    if m.size != capi.MLP_MODEL or abs(float(m[-1]) - gamma) > 1e-6 * gamma:
        raise ValueError(f"{p}: not a {capi.MLP_MODEL}-float model for gamma {gamma}")
    return m


def train(gamma: float, steps=None, batch=None, lr=None, seed=None, norm_samples=None, device=0):
    cfg = dict(DEFAULTS)
    for k, v in (("steps", steps), ("batch", batch), ("lr", lr), ("seed", seed), ("norm_samples", norm_samples)):
        if v is not None:
            cfg[k] = v
    t0 = time.perf_counter()
    model, hist = capi.mlp_train(gamma=gamma, device=device, **cfg)
    cfg["seconds"] = round(time.perf_counter() - t0, 2)
    cfg["gamma"] = gamma
    cfg["loss_first"] = float(np.mean(hist[:10]))
    cfg["loss_last"] = float(np.mean(hist[-100:]))
    return model, cfg


def save(model, cfg) -> str:
    os.makedirs(MODEL_DIR, exist_ok=True)
    p = model_path(cfg["gamma"])
    np.asarray(model, dtype=np.float32).tofile(p)
    with open(p + ".json", "w") as f:
        json.dump(cfg, f, indent=1)
    return p


def load_or_train(gamma: float, **kw):
    """(model, info): the stored model, else one trained now (and stored)."""
    m = load(gamma)
    if m is not None:
        info_path = model_path(gamma) + ".json"
        info = json.load(open(info_path)) if os.path.exists(info_path) else {}
        info["source"] = "stored"
        return m, info
    m, cfg = train(gamma, **kw)
    save(m, cfg)
    cfg["source"] = "trained now"
    return m, cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gamma", type=float, default=5.0 / 3.0)
    for k, v in DEFAULTS.items():
        ap.add_argument("--" + k.replace("_", "-"), type=type(v), default=v)
    a = ap.parse_args()
    model, cfg = train(a.gamma, a.steps, a.batch, a.lr, a.seed, a.norm_samples)
    print(json.dumps(cfg), flush=True)
    print(save(model, cfg), flush=True)


if __name__ == "__main__":
    main()
