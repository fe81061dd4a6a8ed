#!/usr/bin/env python3
"""Benchmark of the RHLLC step on one MI355X.

Workload (the full-form configuration the headline metric is quoted on):
relativistic blast wave in the unit cube on 159^3 hexes x 6 Kuhn tets =
24,118,074 tets (jittered 5%, randomly permuted), rho = 1, p = 1000 inside
r < 0.1 else 1, gamma = 5/3, CFL 0.4, with a seeded FP32 radiation field
(Ur log-uniform [1e-3, 1], |Fr| <= Ur, kappa_a, kappa_s in [0.1, 10]) applied
as an explicit 4-momentum source after every hydro update.

A "step" is one hipGraph replay of face kernel -> cell kernel -> dt kernel
over the whole mesh with the state resident in HBM.  `value` is the
cell-updates per second of the mixed step (FP32 storage of W, geometry and
face flux; FP64 U and arithmetic) with the c2p surrogate start (the stored
model for gamma 5/3, trained on the device first when there is none).  Beside
it, on the same mesh and state: the mixed step started from the previous
pressure (`mixed_prev`) and the plain FP64 step (`fp64`), with the per-field
L1 difference of the conserved variables against FP64 after the same steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--hexes 159]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s RHLLC step @24M tets: mixed+MLP-c2p vs plain FP64, at L1<=1e-5"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--hexes", type=int, default=159, help="hexes per box edge (6 tets each)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU baseline steps (0: auto, ~10-30 s)")
    ap.add_argument("--kernel-steps", type=int, default=10, help="eager steps timed per kernel with HIP events")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the calibration copy kernel (1 GiB read + 1 GiB write per launch)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def per_kernel_bytes(N, F, mixed, rad):
    """Algorithmic bytes per launch: each array element counted once per kernel
    that touches it, gathers once, at the storage precision of the variant."""
    if mixed:
        face = 32 * N + 44 * F  # cell record read (v, p | s - D, D, cs2, ext); ids 8 + S*n 12 + packed flux write 24
        cell = 148 * N + 24 * F  # ids+signs 16 + (V, closure defect) 16 + U r/w 80 + p32 read 4 + record write 32; packed flux read
    else:
        face = 80 * N + 80 * F  # U+W read; ids 8 + (n,S) 32 + flux write 40
        cell = 153 * N + 40 * F  # ids 16 + sign 1 + V 8 + U r/w 80 + p read 8 + W write 40; flux read
    if rad:
        cell += 24 * N
    return face, cell


def run_gpu(mesh, W0, rad, precision, guess, steps, warmup, kernel_steps, model=None):
    from radrhd_solver_amd import capi

    h = mesh.h_min()
    t0 = time.time()
    s = capi.Solver(mesh, gamma=5.0 / 3.0, cfl=0.4, h_min=h, precision=precision, guess=guess, reorder=True)
    s.upload_W(W0)
    if rad is not None:
        s.set_radiation(*rad)
    if model is not None:
        s.set_mlp(model)
    log(f"solver ready ({'mixed' if precision else 'fp64'}, guess {guess}) in {time.time() - t0:.1f} s")
    s.run(warmup, graph=True)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps, graph=True)
    s.sync()
    dt = time.perf_counter() - t0
    stats = s.stats()
    _, U = s.download()
    # Per-kernel device time with HIP events on the solver's stream (eager launches).
    kt = s.time_kernels(kernel_steps)
    s.close()
    return {"seconds": dt, "stats": stats, "U": U, "kernels": kt}


def cpu_baseline(mesh, W0, rad, steps_req):
    import oracle as O

    st = O.State(mesh, W0)
    h = mesh.h_min()
    Ur, Fr, ka, ks = rad
    threads = O.lib().orc_num_threads()
    st.run_rad(0.4, h, 0, Ur, Fr, ka, ks)  # p2c only
    # Warm-up step, then a bounded sample of timed steps (~10-30 s of CPU work).
    tau = 0.4 * h
    t0 = time.perf_counter()
    speed = _cpu_step(st, tau, rad)
    one = time.perf_counter() - t0
    steps = steps_req or max(2, min(20, int(20.0 / max(one, 1e-3))))
    tau = 0.4 * h / speed
    t0 = time.perf_counter()
    for _ in range(steps):
        speed = _cpu_step(st, tau, rad)
        tau = 0.4 * h / speed
    dt = time.perf_counter() - t0
    return {
        "value": mesh.n_cells * steps / dt / 1e6,
        "unit": "Mcell-updates/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"{steps} timed steps (after 1 warm-up) of the same {mesh.n_cells}-tet full-form workload, "
        f"FP64 AoS OpenMP restatement of Hllc3dStab + radiation source ({dt:.1f} s)",
    }


def _cpu_step(st, tau, rad):
    Ur, Fr, ka, ks = rad
    return st.step_rad(tau, Ur, Fr, ka, ks)


def load_traffic(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(key)


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py: --gpus {args.gpus} is not supported: the step runs on one MI355X; "
            "sharding across GPUs is designed but not built.",
            file=sys.stderr,
        )
        sys.exit(2)
    from radrhd_solver_amd import capi
    from radrhd_solver_amd import mesh as M

    if args.copy_probe:
        gbps = capi.copy_probe(1 << 30, 20)
        print(json.dumps({"copy_probe_GBps": round(gbps, 1), "bytes_per_launch": 2 << 30}), flush=True)
        return
    n = args.hexes
    t0 = time.time()
    mesh = M.kuhn_box(n, n, n, seed=SEED)
    W0 = M.init_state(mesh, M.INIT_BLAST)
    rad = M.init_radiation(mesh.n_cells, seed=SEED)
    N, F = mesh.n_cells, mesh.n_faces
    log(f"mesh {n}^3 x 6 = {N} tets, {F} faces in {time.time() - t0:.1f} s")

    from radrhd_solver_amd import surrogate

    t0 = time.time()
    model, minfo = surrogate.load_or_train(5.0 / 3.0)
    log(f"c2p surrogate: {minfo.get('source')} ({time.time() - t0:.1f} s)")
    mixed = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_MLP, args.steps, args.warmup, args.kernel_steps,
                    model=model)
    prev = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_PREVIOUS, args.steps, args.warmup, args.kernel_steps)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu(mesh, W0, rad, capi.PREC_FP64, capi.GUESS_REFERENCE, args.steps, args.warmup,
                       args.kernel_steps)
    V = mesh.cells["V"]

    def rel_l1(a, b):
        return [float((np.abs(a[:, k] - b[:, k]) * V).sum() / max((np.abs(b[:, k]) * V).sum(), 1e-300))
                for k in range(5)]

    ms = mixed["seconds"] / args.steps * 1e3
    value = N * args.steps / mixed["seconds"] / 1e6
    fb, cb = per_kernel_bytes(N, F, True, True)
    kt = mixed["kernels"]
    dom, dom_ms, dom_bytes = ("cell", kt["cell_ms"], cb) if kt["cell_ms"] >= kt["face_ms"] else ("face", kt["face_ms"], fb)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = load_traffic(f"mixed_{dom}")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: seeded jittered/permuted Kuhn-tet mesh, blast-wave state, seeded radiation field",
        "config": {
            "workload": f"full-form blast wave, {n}^3 hexes x 6 = {N} tets, {F} faces, radiation source, "
            "mixed precision (FP32 W/geometry/flux storage, FP64 U and arithmetic), surrogate c2p start, "
            "hipGraph-replayed step",
            "n_cells": N,
            "n_faces": F,
            "c2p_start": "MLP surrogate where tau changed > 2% in the step, previous pressure elsewhere",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_cell (mixed, radiation)" if dom == "cell" else "k_face (mixed)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_ms": round(dom_ms, 5),
            "algorithmic_bytes_per_launch": dom_bytes,
            "step_bytes": fb + cb,
            "step_GBps": round((fb + cb) / (ms * 1e-3) / 1e9, 1),
        },
        "kernels_ms": {k: round(v, 5) for k, v in kt.items()},
        "c2p": {k: mixed["stats"][k] for k in ("c2p_fail", "fix_calls", "fix_fail", "fallback_cells", "mlp_starts")},
        "surrogate": {k: minfo.get(k) for k in ("source", "steps", "batch", "lr", "loss_first", "loss_last", "seconds")},
        "mixed_prev": {
            "value": round(N * args.steps / prev["seconds"] / 1e6, 2),
            "ms_per_step": round(prev["seconds"] / args.steps * 1e3, 4),
            "kernels_ms": {k: round(v, 5) for k, v in prev["kernels"].items()},
            "c2p": {k: prev["stats"][k] for k in ("c2p_fail", "fix_calls", "fix_fail", "fallback_cells")},
        },
    }
    if fp64 is not None:
        fv = N * args.steps / fp64["seconds"] / 1e6
        l1 = rel_l1(mixed["U"], fp64["U"])
        out["mixed_prev"]["rel_l1_vs_fp64"] = rel_l1(prev["U"], fp64["U"])
        out["mixed_prev"]["speedup_vs_fp64"] = round(out["mixed_prev"]["value"] / fv, 3)
        fb64, cb64 = per_kernel_bytes(N, F, False, True)
        out["fp64"] = {
            "value": round(fv, 2),
            "ms_per_step": round(fp64["seconds"] / args.steps * 1e3, 4),
            "kernels_ms": {k: round(v, 5) for k, v in fp64["kernels"].items()},
            "step_GBps": round((fb64 + cb64) / (fp64["seconds"] / args.steps) / 1e9, 1),
        }
        out["speedup_vs_fp64"] = round(value / fv, 3)
        out["l1_vs_fp64"] = {"fields": ["D", "mx", "my", "mz", "E"], "rel_l1": l1, "after_steps": args.steps + args.warmup}
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(mesh, W0, rad, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
