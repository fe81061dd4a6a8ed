#!/usr/bin/env python3
"""Benchmark of the RHLLC step on one MI355X.

Workload (the full-form configuration the headline metric is quoted on):
relativistic blast wave in the unit cube on 159^3 hexes x 6 Kuhn tets =
24,118,074 tets (jittered 5%, randomly permuted), rho = 1, p = 1000 inside
r < 0.1 else 1, gamma = 5/3, CFL 0.4, with a seeded FP32 radiation field
(Ur log-uniform [1e-3, 1], |Fr| <= Ur, kappa_a, kappa_s in [0.1, 10]) applied
as an explicit 4-momentum source after every hydro update.

A "step" is one hipGraph replay of the step over the whole mesh with the
state resident in HBM: the tiled kernel (each workgroup computes the faces of
256 cells into LDS and updates those cells), the surrogate-list kernel and the
dt kernel.  `value` is the cell-updates per second of the mixed step (FP32
cell records, face vectors in FP32 plus residual bits; FP64 U, 1/V, fluxes and
arithmetic) with the c2p surrogate start (the stored model for gamma 5/3, trained on the
device first when there is none).  Beside it, on the same mesh and state: the
mixed step started from the previous pressure (`mixed_prev`), the plain FP64
step under the tiled schedule (`fp64`) and under the two-kernel schedule
(`fp64_two_kernel`: face kernel -> flux array -> cell kernel), and the
per-field L1 of the conserved variables against FP64 after the same steps.

The roofline uses the step's algorithmic bytes of the two-kernel structure
at the variant's storage precision (FP64 233 N + 120 F, mixed 161 N + 64 F,
plus 24 N for the radiation fields): the tiled kernel does all of a step's
work, so its launch is charged the whole step's bytes.  `design_bytes` is
what this build's kernel itself must move at least; `traffic` is the HBM
bytes rocprofv3 counted for the same kernel build (profiles/traffic.json,
used only when its source hash matches the current kernels).

With --eos taub the same workload runs with the Taub-Mathews EOS (the
reference's shipped one); the mixed step then starts c2p from the previous
pressure (the surrogate is trained for the ideal gas) and the line carries its
own metric name.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--hexes 159] [--eos ideal|taub]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s RHLLC step @24M tets: mixed+MLP-c2p vs plain FP64, at L1<=1e-5"
METRIC_TAUB = "Mcell-updates/s RHLLC step @24M tets, EOS TAUB: mixed (previous-pressure c2p start) vs plain FP64, at L1<=1e-5"
HBM_PEAK_GBS = 8000.0
SEED = 20261202
KERNEL_SOURCES = ("rhllc_kernels.hip", "rhllc_physics.h", "flux_ext.h", "rhllc_internal.h", "mlp.h")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed graph steps first: after an idle period the device clock takes ~15 steps to settle")
    ap.add_argument("--hexes", type=int, default=159, help="hexes per box edge (6 tets each)")
    ap.add_argument("--eos", choices=("ideal", "taub"), default="ideal",
                    help="equation of state (taub: the mixed step starts c2p from the previous pressure)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 legs")
    ap.add_argument("--two-kernel-mixed", action="store_true", help="also time the mixed two-kernel schedule")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU baseline timed steps (0: auto, >= 5, ~10-30 s)")
    ap.add_argument("--kernel-steps", type=int, default=20, help="eager steps timed per kernel with HIP events")
    ap.add_argument("--stats-steps", type=int, default=5, help="steps with statistics after the timed region")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the calibration copy kernel (1 GiB read + 1 GiB write per launch)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def model_bytes(N, F, mixed, rad):
    """Algorithmic bytes of one step by kernel of the two-kernel structure (face, cell):
    each array element counted once per kernel that touches it, gathers once, at the
    storage precision of the variant."""
    if mixed:
        face = 20 * N + 44 * F  # W32 read; ids 8 + (n, S) FP32 16 + flux32 write 20
        cell = 141 * N + 20 * F  # ids 16 + sign 1 + V 8 + U64 r/w 80 + p64 r/w 16 + W32 write 20; flux32 read
    else:
        face = 80 * N + 80 * F  # U + W read; ids 8 + (n, S) 32 + flux write 40
        cell = 153 * N + 40 * F  # ids 16 + sign 1 + V 8 + U r/w 80 + p read 8 + W write 40; flux read
    if rad:
        cell += 24 * N  # FP32 Ur, Fr, kappa_a, kappa_s
    return face, cell


def design_bytes(N, F, mixed, rad, tiled):
    """Compulsory bytes of this build's kernels (per step) at their actual layout."""
    r = 24 * N if rad else 0
    if tiled:
        if mixed:  # records read 32 + U r/w 80 + 1/V 8 + slots 8 + records written 32; far cell 4 + S n 16
            return 160 * N + 20 * F + r
        return 176 * N + 40 * F + r  # U + W read 80, U + W written 80, V 8, slots 8; ids 8 + (n, S) 32
    if mixed:  # face: record 32, ids 8 + S n 16 + packed flux 24; cell: ids 16 + 1/V 8 + U r/w 80
        return (32 * N + 48 * F) + (140 * N + 24 * F) + r  # + p32 4 + record 32; packed flux read
    return (80 * N + 80 * F) + (153 * N + 40 * F) + r


def kernel_source_sha():
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "radrhd_solver_amd", "csrc", "hip", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def load_traffic(kernel, n_cells):
    """Counted HBM bytes of `kernel` from profiles/traffic.json, if measured on the current kernel
    sources and on a mesh of this size."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, "no profiles/traffic.json"
    with open(path) as f:
        t = json.load(f)
    if t.get("kernel_source_sha") != kernel_source_sha():
        return None, "profiles/traffic.json was measured on other kernel sources"
    if t.get("n_cells") != n_cells:
        return None, f"profiles/traffic.json was measured at {t.get('n_cells')} cells"
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None, f"no counters for {kernel}"
    return k["total_bytes"], f"rocprofv3 FETCH_SIZE/WRITE_SIZE, commit {t.get('commit')}"


def run_gpu(mesh, W0, rad, precision, guess, steps, warmup, kernel_steps, stats_steps, model=None,
            schedule=None, eos=0):
    from radrhd_solver_amd import capi

    sched = capi.SCHED_TILED if schedule is None else schedule
    h = mesh.h_min()
    t0 = time.time()
    s = capi.Solver(mesh, gamma=5.0 / 3.0, cfl=0.4, h_min=h, precision=precision, guess=guess, reorder=True,
                    schedule=sched, eos=eos)
    s.upload_W(W0)
    if rad is not None:
        s.set_radiation(*rad)
    if model is not None:
        s.set_mlp(model)
    ready = time.time() - t0
    log(f"solver ready ({'mixed' if precision else 'fp64'}, guess {guess}, schedule {sched}) in {ready:.1f} s")
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
    # Statistics of a few more steps (other kernels: outside the timed region).
    c2p = None
    if stats_steps > 0:
        s.set_collect_stats(True)
        s.reset_stats()
        s.run(stats_steps, graph=False)
        s.sync()
        st = s.stats()
        s.set_collect_stats(False)
        n = mesh.n_cells * stats_steps
        c2p = {
            "steps": stats_steps,
            "max_residual": st["max_residual"],
            "newton_per_cell": st["newton_iters"] / n,
            "newton_hist": {str(k): v for k, v in enumerate(st["newton_hist"]) if v},
            "fallback_cells": st["fallback_cells"],
            "mlp_starts": st["mlp_starts"],
            "c2p_fail": st["c2p_fail"],
            "fix_fail": st["fix_fail"],
        }
    s.close()
    return {"seconds": dt, "stats": stats, "U": U, "kernels": kt, "ready_s": ready, "stats_run": c2p}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, W0, rad, steps_req, eos=0):
    """The oracle's FP64 AoS OpenMP step (the reference's loop structure and records), -O3
    timing build, on this host's cores: 2 warm-up steps, then the median of >= 5 steps."""
    import oracle as O

    st = O.State(mesh, W0, eos=eos, fast=True)
    h = mesh.h_min()
    Ur, Fr, ka, ks = rad
    threads = st.L.orc_num_threads()
    st.run_rad(0.4, h, 0, Ur, Fr, ka, ks)  # p2c only
    tau = 0.4 * h
    t0 = time.perf_counter()
    for _ in range(2):
        speed = st.step_rad(tau, Ur, Fr, ka, ks)
        tau = 0.4 * h / speed
    one = (time.perf_counter() - t0) / 2
    steps = steps_req or max(5, min(25, int(20.0 / max(one, 1e-3))))
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        speed = st.step_rad(tau, Ur, Fr, ka, ks)
        times.append(time.perf_counter() - t0)
        tau = 0.4 * h / speed
    med = statistics.median(times)
    return {
        "value": mesh.n_cells / med / 1e6,
        "unit": "Mcell-updates/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"median of {steps} timed steps after 2 warm-up steps of the same {mesh.n_cells}-tet full-form "
        f"workload ({sum(times):.1f} s): the oracle's FP64 AoS OpenMP restatement of Hllc3dStab + radiation "
        f"source",
        "build": O.FAST_FLAGS,
        "omp_threads": int(threads),
        "cpu": cpu_model(),
    }


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

    taub = args.eos == "taub"
    eos = capi.EOS_TAUB if taub else capi.EOS_IDEAL
    leg = dict(steps=args.steps, warmup=args.warmup, kernel_steps=args.kernel_steps, stats_steps=args.stats_steps,
               eos=eos)
    if taub:
        model, minfo = None, {}
        mixed = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_PREVIOUS, **leg)
        prev = mixed
    else:
        t0 = time.time()
        model, minfo = surrogate.load_or_train(5.0 / 3.0)
        log(f"c2p surrogate: {minfo.get('source')} ({time.time() - t0:.1f} s)")
        mixed = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_MLP, model=model, **leg)
        prev = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_PREVIOUS, **leg)
    mixed2k = None
    if args.two_kernel_mixed:
        mixed2k = run_gpu(mesh, W0, rad, capi.PREC_MIXED, capi.GUESS_PREVIOUS if taub else capi.GUESS_MLP,
                          model=model, schedule=capi.SCHED_TWO_KERNEL, **leg)
    fp64 = fp64_2k = None
    if not args.no_fp64:
        fp64 = run_gpu(mesh, W0, rad, capi.PREC_FP64, capi.GUESS_REFERENCE, **leg)
        fp64_2k = run_gpu(mesh, W0, rad, capi.PREC_FP64, capi.GUESS_REFERENCE, schedule=capi.SCHED_TWO_KERNEL,
                          **leg)
    V = mesh.cells["V"]

    def rel_l1(a, b):
        return [float((np.abs(a[:, k] - b[:, k]) * V).sum() / max((np.abs(b[:, k]) * V).sum(), 1e-300))
                for k in range(5)]

    def rate(r):
        return N * args.steps / r["seconds"] / 1e6

    def ms(r):
        return r["seconds"] / args.steps * 1e3

    def kms(r):
        return {k: round(v, 5) for k, v in r["kernels"].items()}

    value = rate(mixed)
    fb, cb = model_bytes(N, F, True, True)
    step_bytes = fb + cb
    kt = mixed["kernels"]
    kernel = "k_tile<2, true, false, 1>" if taub else "k_tile<3, true, false, 0>"
    dom_ms = kt["face_ms"]  # the tiled kernel (faces and cells of the step)
    achieved = step_bytes / (dom_ms * 1e-3) / 1e9
    traffic, traffic_src = load_traffic(kernel, N)
    out = {
        "metric": METRIC_TAUB if taub else METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms(mixed), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: seeded jittered/permuted Kuhn-tet mesh, blast-wave state, seeded radiation field",
        "config": {
            "workload": f"full-form blast wave, {n}^3 hexes x 6 = {N} tets, {F} faces, radiation source, "
            f"{'EOS Taub-Mathews' if taub else 'ideal gas gamma 5/3'}, mixed precision (FP32 cell records, face "
            "vectors FP32 plus residuals, FP64 U, fluxes and arithmetic), tiled step (faces into LDS, cells from "
            "LDS), hipGraph-replayed",
            "n_cells": N,
            "n_faces": F,
            "c2p_start": "previous pressure (the surrogate is trained for the ideal gas)" if taub
            else "MLP surrogate where tau changed > 5% in the step, previous pressure elsewhere",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"{kernel} (tiled step, mixed, {'previous-pressure' if taub else 'surrogate'} start, radiation)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel_ms": round(dom_ms, 5),
            "algorithmic_bytes_per_launch": step_bytes,
            "bytes_model": "two-kernel structure at storage precision: mixed 161 N + 64 F + 24 N (radiation)",
            "design_bytes": design_bytes(N, F, True, True, True),
            "step_GBps": round(step_bytes / (ms(mixed) * 1e-3) / 1e9, 1),
        },
        "kernels_ms": kms(mixed),
        "c2p": {k: mixed["stats"][k] for k in ("c2p_fail", "fix_calls", "fix_fail", "fallback_cells", "mlp_starts")},
        "c2p_stats_run": mixed["stats_run"],
        "surrogate": {k: minfo.get(k) for k in ("source", "steps", "batch", "lr", "loss_first", "loss_last", "seconds")},
        "ready_s": round(mixed["ready_s"], 2),
        "mixed_prev": {
            "value": round(rate(prev), 2),
            "ms_per_step": round(ms(prev), 4),
            "kernels_ms": kms(prev),
            "c2p": {k: prev["stats"][k] for k in ("c2p_fail", "fix_calls", "fix_fail", "fallback_cells")},
            "c2p_stats_run": prev["stats_run"],
        },
    }
    if mixed2k is not None:
        out["mixed_two_kernel"] = {"value": round(rate(mixed2k), 2), "ms_per_step": round(ms(mixed2k), 4),
                                   "kernels_ms": kms(mixed2k)}
    if fp64 is not None:
        l1 = rel_l1(mixed["U"], fp64["U"])
        out["mixed_prev"]["rel_l1_vs_fp64"] = rel_l1(prev["U"], fp64["U"])
        fb64, cb64 = model_bytes(N, F, False, True)
        out["fp64"] = {
            "value": round(rate(fp64), 2),
            "ms_per_step": round(ms(fp64), 4),
            "kernels_ms": kms(fp64),
            "step_GBps": round((fb64 + cb64) / (ms(fp64) * 1e-3) / 1e9, 1),
            "c2p_stats_run": fp64["stats_run"],
        }
        out["fp64_two_kernel"] = {
            "value": round(rate(fp64_2k), 2),
            "ms_per_step": round(ms(fp64_2k), 4),
            "kernels_ms": kms(fp64_2k),
            "step_GBps": round((fb64 + cb64) / (ms(fp64_2k) * 1e-3) / 1e9, 1),
            "bitwise_equal_to_tiled": bool(np.array_equal(fp64["U"], fp64_2k["U"])),
        }
        best = max(rate(fp64), rate(fp64_2k))
        out["speedup_vs_fp64"] = round(value / best, 3)
        out["speedup_vs_fp64_note"] = "against the faster of the two FP64 schedules"
        out["mixed_prev"]["speedup_vs_fp64"] = round(rate(prev) / best, 3)
        out["l1_vs_fp64"] = {"fields": ["D", "mx", "my", "mz", "E"], "rel_l1": l1,
                             "after_steps": args.steps + args.warmup}
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(mesh, W0, rad, args.cpu_steps, eos)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
